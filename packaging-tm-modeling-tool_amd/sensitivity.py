"""Tornado sensitivity on the device: the outputs of one Draft analysis and the
tornado table built from the base, low and high cases
(sensitivityOutputs / runSensitivityCase / tornadoFromCases,
src/core/14-sensitivity.js:124-172).

A case is a thermal path solved on the device (path.run_path) from its plan
and step data; its outputs are the JEITA warpage of the 25 degC field, the
warpage at the reflow peak, the largest interior peak of the top-surface
stress of the silicon parts at 25 degC, and the worst-bump dW proxy from the
bump screening of the case (screening.screen_joints over each die's bump
field, driven by the case's own path increments).  The adjoint pair of
hip_solver.Model gives the derivatives of other objectives from one extra
solve; the tornado here is the reference's finite-difference form, case by
case.
"""
from __future__ import annotations

import math

import numpy as np

from . import path as pth
from . import post
from . import screening as scr
from .meshspec import PART

OUTPUT_KEYS = ("warpRT", "warpPeak", "dieStress", "dW")


def sensitivity_outputs(model, mesh, steps, plan, step_of, fields, bga, geom, peak, dW=None):
    """Outputs of one solved case (sensitivityOutputs, src/core/14-sensitivity.js:124-138).

    fields: path.run_path's fields (25 degC and the evaluation temperatures);
    step_of(step) the hip_solver.Step of a plan step; the stress histories use
    the step increments as the analysis keeps them after finishPath (single
    precision, src/core/11-analysis.js:255), read back with model.path_du.
    dW: the worst-bump dW proxy of the case's screening, or None (-inf when
    the case has no cycling, as the reference's)."""
    w25 = post.warpage_jeita(mesh, fields[25], bga)["signed"]
    w_peak = post.warpage_jeita(mesh, fields[peak], bga)["signed"] if peak in fields else math.nan
    chain = [(step_of(steps[k]), model.path_du(k).astype(np.float64)) for k in plan["chain"]]
    view = post.StressView(model, chain, 25)
    s_peak = -math.inf
    for d in range(len(geom["siParts"])):
        s_peak = max(s_peak, post.die_surface_stress(view, geom, d, "top")["peakInterior"])
    return {"warpRT": w25, "warpPeak": w_peak, "dieStress": s_peak, "dW": -math.inf if dW is None else dW}


def worst_bump_dw(model, steps, plan, step_of, temps, zmid, dies, solder_params, solder, cycling):
    """The worst-bump dW proxy of a solved case (sensitivityOutputs' dW: the largest
    top dW of screenBumps over the dies, src/core/14-sensitivity.js:134-135,
    src/core/13-fatigue.js:359-364): each die's flip-chip bump field screened at
    the bump layer's mid-plane zmid over the cycling temperatures `temps`, the
    strain tables from the path increments as the analysis keeps them (single
    precision); -inf without cycling or screened sites."""
    dW = -math.inf
    if len(temps) <= 1:
        return dW

    def pairs(ids):
        return [(step_of(steps[k]), model.path_du(k).astype(np.float64)) for k in ids]
    chain, up, down = pairs(plan["chain"]), pairs(plan["up"]), pairs(plan["down"])
    for d in dies:
        r = scr.screen_joints(model, chain, up, down, temps, d, zmid, PART["SOLDER"], solder_params, solder, cycling,
                              variant=getattr(model, "variant", "hip"))
        if len(r["top"]):
            dW = max(dW, float(r["dW"][r["top"][0]]))
    return dW


def run_case(model, mesh, steps, plan, step_of, temps_all, bga, geom, peak, tol=1e-8, precision=None, screen=None):
    """One sensitivity case (runSensitivityCase, src/core/14-sensitivity.js:150-157) on the
    device: the path solved step by step with a warm start along it, then its outputs.
    screen: None (dW left at -inf), or the keyword arguments of worst_bump_dw past
    the path (temps, zmid, dies, solder_params, solder, cycling)."""
    fields, _, _, recs = pth.run_path(model, steps, plan, temps_all, step_of, tol=tol, precision=precision)
    dW = worst_bump_dw(model, steps, plan, step_of, **screen) if screen is not None else None
    out = sensitivity_outputs(model, mesh, steps, plan, step_of, fields, bga, geom, peak, dW)
    return out, recs


def path_field_gradient(model, steps, ids, step_of, dJdu, tol=1e-10, precision=None):
    """Log-derivatives with respect to every group's E and CTE of an objective of
    a path field u_T = sum over the steps ids of their increments (the fields of
    finishPath, src/core/11-analysis.js:236-247), given dJ/du_T: each step k
    solves K_k du_k = f_k, so dJ/dtheta = sum_k lambda_k . (df_k/dtheta -
    dK_k/dtheta du_k) with K_k lambda_k = dJ/du_T, one forward re-solve and one
    adjoint solve per step (hip_solver.Model.adjoint_forward /
    adjoint_backward_u).  Returns (dlnE, dlnA) per group: scaling a group's
    stiffness (and with it its thermal stress) or its thermal strain by f
    changes J by dlnE[g] ln f or dlnA[g] ln f to first order."""
    from . import hip_solver as hs
    precision = hs.PREC_FP64 if precision is None else precision
    dE = dA = 0.0
    for k in ids:
        model.adjoint_forward(step_of(steps[k]), 0, np.zeros(0, np.int32), 8.0, tol=tol, precision=precision)
        e, a, _, rec = model.adjoint_backward_u(dJdu)
        if not rec["converged"]:
            raise hs.SolverError(f"adjoint solve of step {k} did not converge")
        dE, dA = dE + e, dA + a
    return dE, dA


def tornado_from_cases(inputs, base, cases, keys=OUTPUT_KEYS):
    """Tornado rows per output (tornadoFromCases, src/core/14-sensitivity.js:159-172):
    for every input with both cases, the low and high changes from the base, the
    values, and the swing |hi - lo|; rows by decreasing swing (a stable order for
    equal swings), the first ten and all of them.

    inputs: [{id, label}]; base: {output: value}; cases: {id: {lo, hi, loValue,
    hiValue, baseValue}} with lo / hi the outputs of the two cases."""
    out = {}
    for k in keys:
        rows = []
        for inp in inputs:
            c = cases.get(inp["id"])
            if not c:
                continue
            lo, hi = c["lo"][k] - base[k], c["hi"][k] - base[k]
            rows.append({"id": inp["id"], "label": inp.get("label"), "lo": lo, "hi": hi, "loValue": c.get("loValue"),
                         "hiValue": c.get("hiValue"), "baseValue": c.get("baseValue"), "swing": abs(hi - lo)})
        # JS sorts NaN comparisons as equal; keep such rows where they are
        rows = sorted(rows, key=lambda r: -r["swing"] if r["swing"] == r["swing"] else 0.0)
        out[k] = {"base": base[k], "rows": rows[:10], "all": rows}
    return out
