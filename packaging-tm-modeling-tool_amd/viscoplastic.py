"""Host side of the bump-array cycling workload: the global viscoplastic
Newton (fcbga_vp_*) stepped through the temperature cycles.

The device does every time step (Anand points, B-bar solder tangents, inexact
Newton with the multigrid PCG inner solve); the host keeps what is per step
or per cycle:

* the history: stress-free at Tref, then ramps and dwells of equal duration,
  each a fixed number of steps (`cycle_history`);
* the step data of the elastic groups for the total load from Tref
  (`elastic_step`, prepareStep's output for temperature-independent tables,
  src/core/09-assembly.js:346-365);
* cutbacks: a step that does not converge (or whose plastic increment exceeds
  the limit) is restored and split in two, recursively (the reference's
  submodel doubles its sub-steps on failure, src/core/13-fatigue.js:1059-1075);
* the per-bump work density per cycle: W averaged (volume-weighted) over the
  solder points of the die-side and substrate-side layers of each bump (the
  25 um bands of the reference's submodel, src/core/13-fatigue.js:1010-1023),
  dW = the larger of the two layers' increments over the cycle.
"""
from __future__ import annotations

import time

import numpy as np

from . import hip_solver as hs
from . import meshspec

SUB_T_AVG = 0.025              # mm, interface averaging layer (src/core/00-constants.js:114)
INELASTIC_STEP_MAX = 2e-4      # plastic increment limit of a step (src/core/00-constants.js:113)


def solder_of(spec: dict) -> dict:
    g = spec["solder_group"]
    m = spec["materials"][g]
    return {"group": g, "anand": m["anand"], "nu": m["nu"], "E_table": m["E_table"], "cte": m["cte"],
            "Tref": spec["Tref"]}


def elastic_step(spec: dict, T: float) -> hs.Step:
    """Group data for the total load from the stress-free Tref to T."""
    D, s0, born = meshspec.group_step_data(spec, spec["Tref"], T, 0)
    return hs.Step(spec["Tref"], T, 0, D, s0, born)


def cycle_history(spec: dict) -> dict:
    """Times, temperatures and cycle index of every history point: from Tref a
    ramp to Tmax, then per cycle a dwell at Tmax, a ramp to Tmin, a dwell at
    Tmin and (but for the last cycle) a ramp back to Tmax; each ramp and dwell
    lasts segment_s and takes steps_per_segment equal steps."""
    c = spec["cycle"]
    n, seg = c["steps_per_segment"], c["segment_s"]
    t, T, cyc = [0.0], [spec["Tref"]], [0]

    def segment(Ta, Tb, k):
        for q in range(1, n + 1):
            t.append(t[-1] + seg / n)
            T.append(Ta + (Tb - Ta) * q / n)
            cyc.append(k)

    for k in range(1, c["cycles"] + 1):
        segment(T[-1], c["Tmax"], k)
        segment(c["Tmax"], c["Tmax"], k)
        segment(c["Tmax"], c["Tmin"], k)
        segment(c["Tmin"], c["Tmin"], k)
    return {"t": np.array(t), "T": np.array(T), "cycle": np.array(cyc), "cycles": c["cycles"]}


def bump_layers(mesh, spec: dict, elems: np.ndarray):
    """For every solder element (in the engine's order): its bump, whether it
    lies in the die-side or substrate-side averaging layer, and its point
    weight (volume / 8)."""
    b = spec.get("bumps")  # None: one joint (a submodel)
    k0, k1 = spec["bump_layer"]
    zt, zaf = mesh.zs[k0], mesh.zs[k1]
    bump = np.empty(len(elems), np.int64)
    die = np.zeros(len(elems), bool)
    sub = np.zeros(len(elems), bool)
    w = np.empty(len(elems))
    for q, e in enumerate(elems):
        i, j, k = mesh.e_ijk[e]
        xc, yc, zc = [(a[v] + a[v + 1]) / 2 for a, v in ((mesh.xs, i), (mesh.ys, j), (mesh.zs, k))]
        bump[q] = meshspec.bump_index(b, xc, yc) if b else 0
        die[q] = zc > zaf - SUB_T_AVG
        sub[q] = zc < zt + SUB_T_AVG
        w[q] = (mesh.xs[i + 1] - mesh.xs[i]) * (mesh.ys[j + 1] - mesh.ys[j]) * (mesh.zs[k + 1] - mesh.zs[k]) / 8
    return bump, die, sub, w


def layer_means(points: np.ndarray, layers, nb: int, field: int = 8):
    """Per bump: volume-weighted mean of a point field (8: W, 7: ebar) over the
    die-side and substrate-side layers."""
    bump, die, sub, w = layers
    v = points[:, :, field].sum(axis=1)  # the 8 points of an element share its weight
    out = []
    for mask in (die, sub):
        num = np.bincount(bump[mask], weights=(w * v)[mask], minlength=nb)
        den = np.bincount(bump[mask], weights=(8 * w)[mask], minlength=nb)
        out.append(num / np.where(den > 0, den, np.nan))
    return out[0], out[1]


def run(vp: hs.ViscoplasticModel, spec: dict, nsteps: int = None, precision: int = hs.PREC_MIXED,
        max_depth: int = 4, dmax_limit: float = None, log=None, prescribe=None, field=None, **newton):
    """Step the model through the cycle history (its first nsteps steps);
    returns the record: per-step Newton counts and times, cutbacks, and the
    per-bump work densities per cycle.  A step is cut back when Newton fails
    and, with dmax_limit (the reference submodel's is INELASTIC_STEP_MAX),
    when its largest plastic increment exceeds the limit.  prescribe(T), when
    given, returns the constrained DOFs' displacement at T (a submodel's
    boundary), set before every (sub-)step; field(T), when given, the
    boundary field at every node, whose increment over the step is added to
    the whole displacement as the Newton's starting point (the boundary jump
    then does not land on the boundary elements alone)."""
    hist = cycle_history(spec)
    n = len(hist["T"]) - 1 if nsteps is None else min(nsteps, len(hist["T"]) - 1)
    nb = spec["bumps"]["n"] ** 2 if spec.get("bumps") else 1
    _, elems = vp.points()
    layers = bump_layers(vp.mesh, spec, elems)
    Wd = [np.zeros(nb)]
    Ws = [np.zeros(nb)]
    steps = []
    cutbacks = [0]
    t0 = time.perf_counter()

    def advance(T0, T1, dt, depth):
        vp.save()
        if prescribe is not None:
            vp.set_prescribed(prescribe(T1))
        if field is not None:
            vp.set_displacement(vp.displacement() + field(T1) - field(T0))
        rec = vp.step(elastic_step(spec, T1), T1, dt, precision=precision, **newton)
        ok = rec["converged"] and (dmax_limit is None or rec["dmax"] <= dmax_limit)
        if ok:
            steps.append(dict(rec, T=T1, dt=dt, depth=depth))
            return
        if depth >= max_depth:
            raise hs.SolverError(f"viscoplastic step to {T1:.2f} C did not converge after {depth} cutbacks")
        vp.restore()
        cutbacks[0] += 1
        Tm = 0.5 * (T0 + T1)
        advance(T0, Tm, dt / 2, depth + 1)
        advance(Tm, T1, dt / 2, depth + 1)

    for k in range(1, n + 1):
        advance(hist["T"][k - 1], hist["T"][k], hist["t"][k] - hist["t"][k - 1], 0)
        if log:
            s = steps[-1]
            log(f"step {k}/{n} T={hist['T'][k]:.1f} newton {s['newton']} pcg {s['pcg_iters']} "
                f"dmax {s['dmax']:.2e} {s['ms']:.1f} ms")
        last_of_cycle = k == n or hist["cycle"][k + 1] != hist["cycle"][k]
        if last_of_cycle:
            pts, _ = vp.points()
            d, s_ = layer_means(pts, layers, nb)
            Wd.append(d)
            Ws.append(s_)
    wall = time.perf_counter() - t0
    per_cycle = []
    for c in range(1, len(Wd)):
        dd, ds = Wd[c] - Wd[c - 1], Ws[c] - Ws[c - 1]
        per_cycle.append({"dWdie": dd, "dWsub": ds, "dW": np.fmax(dd, ds)})
    return {
        "steps": n, "solves": len(steps), "cutbacks": cutbacks[0],
        "newton": int(sum(s["newton"] for s in steps)), "pcg_iters": int(sum(s["pcg_iters"] for s in steps)),
        "s": wall, "ms_per_step": 1e3 * wall / max(n, 1),
        "ms_setup": float(sum(s["ms_setup"] for s in steps)), "ms_pcg": float(sum(s["ms_pcg"] for s in steps)),
        "per_cycle": per_cycle, "step_records": steps,
    }


def submodel_boundary_values(mesh, spec: dict):
    """prescribe(T) for a batched submodel spec (cases.cfg5_spec): the seeded
    boundary field at the coordinates of the constrained DOFs' nodes."""
    from . import cases
    nodes = mesh.constraint_dofs[0::3] // 3
    shape = cases.cfg5_boundary(spec, mesh.coords[nodes], spec["Tref"] + 1.0).reshape(-1)  # linear in T - Tref

    def prescribe(T):
        return (T - spec["Tref"]) * shape
    return prescribe


def submodel_field(mesh, spec: dict):
    """field(T): the boundary field at every node (the Newton's predictor)."""
    from . import cases
    shape = cases.cfg5_boundary(spec, mesh.coords, spec["Tref"] + 1.0).reshape(-1)  # linear in T - Tref

    def field(T):
        return (T - spec["Tref"]) * shape
    return field


def run_submodel(spec: dict, variant: str = "hip", precision: int = hs.PREC_MIXED, nsteps: int = None, **newton):
    """One batched bump submodel through its cycle: mesh, engine, boundary,
    steps; returns its record with the wall time, DOF and the joint's dW."""
    t0 = time.perf_counter()
    mesh = meshspec.build_mesh(spec)
    vp = hs.ViscoplasticModel(mesh, solder_of(spec), variant=variant)
    try:
        rec = run(vp, spec, nsteps=nsteps, precision=precision, prescribe=submodel_boundary_values(mesh, spec),
                  field=submodel_field(mesh, spec), **newton)
    finally:
        vp.close()
    pc = rec.pop("per_cycle")
    rec.pop("step_records")
    rec.update({"id": spec["draw"]["id"], "n": spec["draw"]["n"], "ndof": mesh.ndof, "solder_elements": vp.nsolder,
                "wall_s": time.perf_counter() - t0,
                "dW": float(pc[-1]["dW"][0]) if pc else None, "dWdie": float(pc[-1]["dWdie"][0]) if pc else None,
                "dWsub": float(pc[-1]["dWsub"][0]) if pc else None})
    return rec


def run_submodel_batch(specs, workers: int = 4, variant: str = "hip", **kw):
    """Many submodels at once (the reference gives each BumpSubmodel to a helper
    worker, src/worker.js:222-246): one host thread and one engine (its own
    device stream) per submodel in flight, `workers` in flight; records in the
    order of specs."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max(1, workers)) as ex:
        return list(ex.map(lambda sp: run_submodel(sp, variant=variant, **kw), specs))
