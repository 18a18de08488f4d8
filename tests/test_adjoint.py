"""Adjoint sensitivities (the forward/backward pair that replaces the
finite-difference tornado of runSensitivityCase, src/core/14-sensitivity.js:
8-172): gradients of the warpage proxy and of the die von Mises p-norm with
respect to every group's E and CTE and every element's CTE, checked against
central finite differences with 1e-3 relative steps of the same solver (the
reference pins only the sign and ranking of its own, different, objectives)."""
import copy

import numpy as np
import pytest

from conftest import golden_spec
from fcbga_amd import cases
from fcbga_amd import hip_solver as hs
from fcbga_amd import meshspec

H = 1e-3


def _case(variant, name="small_box"):
    spec = golden_spec(name) if name != "cfg4" else cases.cfg4_spec()
    mesh = meshspec.build_mesh(spec)
    model = hs.Model(mesh, variant=variant)
    D, s0, born = meshspec.group_step_data(spec, spec["Ta"], spec["Tb"], 0)
    if name == "cfg4":
        obj = spec["objective"]
        k_top, vm, p = obj["k_top"], np.nonzero(np.isin(mesh.e_group, obj["vm_groups"]))[0], obj["p"]
    else:
        k_top, vm, p = len(spec["zs"]) - 1, np.nonzero(mesh.e_group == 0)[0], 8.0
    return spec, mesh, model, (D, s0, born), (k_top, vm.astype(np.int32), p)


def _J(model, spec, data, obj, prec, tol):
    D, s0, born = data
    k_top, vm, p = obj
    J, _, rec = model.adjoint_forward(hs.Step(spec["Ta"], spec["Tb"], 0, D, s0, born), k_top, vm, p, tol=tol,
                                      maxit=5000, precision=prec)
    assert rec["converged"]
    return J


def _fd(model, spec, data, obj, prec, tol, groups=None, h=H, richardson=False):
    """Central differences in ln E_g (D and s0 scaled) and ln CTE_g (s0 scaled)
    with relative step h; richardson: (4 d(h) - d(2h)) / 3, which cancels the
    O(h^2) term so that larger steps can be used."""
    D, s0, born = data
    groups = range(len(D)) if groups is None else groups

    def central(g, scale_D, step):
        vals = []
        for sgn in (1, -1):
            Dp, sp = D.copy(), s0.copy()
            if scale_D:
                Dp[g] *= 1 + sgn * step
            sp[g] *= 1 + sgn * step
            vals.append(_J(model, spec, (Dp, sp, born), obj, prec, tol))
        return (vals[0] - vals[1]) / (2 * step)

    fE, fA = {}, {}
    for g in groups:
        for out, scale_D in ((fE, True), (fA, False)):
            d1 = central(g, scale_D, h)
            out[g] = (4 * d1 - central(g, scale_D, 2 * h)) / 3 if richardson else d1
    return fE, fA


def _adjoint(model, spec, data, obj, prec, tol, gJ):
    _J(model, spec, data, obj, prec, tol)
    dE, dA, de, rec = model.adjoint_backward(gJ)
    assert rec["converged"]
    return dE, dA, de


def _check(adj, fd, rtol, floor=1e-3):
    """Per parameter: |adjoint - fd| <= rtol |fd|, for parameters whose
    sensitivity is above floor x the largest one (smaller ones: absolute
    rtol x largest)."""
    big = max(abs(v) for v in fd.values())
    for g, v in fd.items():
        tolg = rtol * (abs(v) if abs(v) >= floor * big else big)
        assert abs(adj[g] - v) <= tolg, (g, adj[g], v)


@pytest.mark.parametrize("comp", [0, 1])
def test_adjoint_matches_central_differences(comp):
    spec, mesh, model, data, obj = _case("hostemu")
    gJ = np.eye(2)[comp]
    dE, dA, _ = _adjoint(model, spec, data, obj, hs.PREC_FP64, 1e-13, gJ)
    fE, fA = _fd(model, spec, data, obj, hs.PREC_FP64, 1e-13)
    _check(dE, {g: v[comp] for g, v in fE.items()}, 1e-5)
    _check(dA, {g: v[comp] for g, v in fA.items()}, 1e-5)


def test_element_cte_gradient():
    """Per-element CTE sensitivities add up to the group's, and one element's
    equals the CTE sensitivity of a group holding only that element (the
    element split off into a copy of its group)."""
    spec, mesh, model, data, obj = _case("hostemu")
    gJ = np.array([1.0, 0.02])
    dE, dA, de = _adjoint(model, spec, data, obj, hs.PREC_FP64, 1e-13, gJ)
    for g in range(len(dA)):
        assert abs(de[mesh.e_group == g].sum() - dA[g]) <= 1e-12 * np.abs(dA).max()
    D, s0, born = data
    for e in (int(obj[1][3]), int(np.nonzero(mesh.e_group == 1)[0][5])):
        m2 = copy.deepcopy(mesh)
        g = int(m2.e_group[e])
        m2.e_group = m2.e_group.copy()
        m2.e_group[e] = len(m2.groups)
        m2.groups = list(m2.groups) + [dict(m2.groups[g], name=m2.groups[g]["name"] + "_split")]
        data2 = (np.vstack([D, D[g]]), np.vstack([s0, s0[g]]), np.append(born, born[g]))
        model2 = hs.Model(m2, variant="hostemu")
        _, fA = _fd(model2, spec, data2, obj, hs.PREC_FP64, 1e-13, groups=[len(D)])
        fd = float(gJ @ fA[len(D)])
        assert abs(de[e] - fd) <= 1e-5 * abs(fd) + 1e-9 * np.abs(de).max(), (e, de[e], fd)


def test_backward_is_linear_and_survives_other_solves():
    spec, mesh, model, data, obj = _case("hostemu")
    D, s0, born = data
    g1, g2 = np.array([1.0, 0.0]), np.array([0.0, 1.0])
    a = _adjoint(model, spec, data, obj, hs.PREC_FP64, 1e-13, g1)
    b = _adjoint(model, spec, data, obj, hs.PREC_FP64, 1e-13, g2)
    _J(model, spec, data, obj, hs.PREC_FP64, 1e-13)
    # another solve on the model in between: backward rebuilds the forward's operator
    model.solve_step(hs.Step(spec["Ta"], spec["Tb"], 0, 2 * D, s0, born), tol=1e-8)
    c, cA, ce, _ = model.adjoint_backward(3 * g1 - 2 * g2)
    assert np.abs(c - (3 * a[0] - 2 * b[0])).max() <= 1e-8 * np.abs(c).max()
    assert np.abs(cA - (3 * a[1] - 2 * b[1])).max() <= 1e-8 * np.abs(cA).max()
    assert np.abs(ce - (3 * a[2] - 2 * b[2])).max() <= 1e-8 * np.abs(ce).max()


def test_backward_without_forward_and_bad_objective():
    spec, mesh, model, data, obj = _case("hostemu")
    with pytest.raises(hs.SolverError):
        model.adjoint_backward([1.0, 0.0])
    D, s0, born = data
    with pytest.raises(hs.SolverError):
        model.adjoint_forward(hs.Step(spec["Ta"], spec["Tb"], 0, D, s0, born), len(spec["zs"]), [])
    with pytest.raises(hs.SolverError):
        model.adjoint_forward(hs.Step(spec["Ta"], spec["Tb"], 0, D, s0, born), 0, [mesh.ne])


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [hs.PREC_FP64, hs.PREC_MIXED])
def test_gpu_adjoint_small_box(prec):
    """The device pair (adjoint solve in either precision) against central
    differences of the device's own FP64 forward solves, which use no part of
    the adjoint: every group's E and CTE sensitivity of both objectives, and
    the per-element CTE sensitivities of two elements, each split off into a
    group of its own and differenced there."""
    spec, mesh, model, data, obj = _case("hip")
    fE, fA = _fd(model, spec, data, obj, hs.PREC_FP64, 1e-13)
    for comp in (0, 1):
        gJ = np.eye(2)[comp]
        dE, dA, _ = _adjoint(model, spec, data, obj, prec, 1e-12, gJ)
        _check(dE, {g: v[comp] for g, v in fE.items()}, 1e-5)
        _check(dA, {g: v[comp] for g, v in fA.items()}, 1e-5)
    gJ = np.array([1.0, 0.02])
    dE, dA, de = _adjoint(model, spec, data, obj, prec, 1e-12, gJ)
    for g in range(len(dA)):
        assert abs(de[mesh.e_group == g].sum() - dA[g]) <= 1e-10 * np.abs(dA).max()
    D, s0, born = data
    for e in (int(obj[1][3]), int(np.nonzero(mesh.e_group == 1)[0][5])):
        m2 = copy.deepcopy(mesh)
        g = int(m2.e_group[e])
        m2.e_group = m2.e_group.copy()
        m2.e_group[e] = len(m2.groups)
        m2.groups = list(m2.groups) + [dict(m2.groups[g], name=m2.groups[g]["name"] + "_split")]
        data2 = (np.vstack([D, D[g]]), np.vstack([s0, s0[g]]), np.append(born, born[g]))
        _, fe = _fd(hs.Model(m2, variant="hip"), spec, data2, obj, hs.PREC_FP64, 1e-13, groups=[len(D)])
        fd = float(gJ @ fe[len(D)])
        assert abs(de[e] - fd) <= 1e-5 * abs(fd) + 1e-9 * np.abs(de).max(), (e, de[e], fd)


@pytest.mark.gpu
def test_gpu_cfg4_adjoint_vs_central_differences():
    """The lidded mesh with 18 groups: gradients of both objectives with
    respect to the 36 group parameters from one forward and one adjoint solve
    (mixed precision, 1e-10) against central differences to 1e-4 relative.
    On this mesh J carries FP64 rounding amplified by the stiffness contrast
    (about 1e-9 of J, the same for solves converged to 1e-17 and 1e-21), which
    1e-3 steps turn into up to 3e-4 errors of the smaller sensitivities; the
    differences therefore use 4e-3 and 8e-3 steps with Richardson
    extrapolation.  Sensitivities below 1 % of the largest are held to 1e-4 of
    the largest instead."""
    spec, mesh, model, data, obj = _case("hip", "cfg4")
    gJ = np.array([1.0, 1e-3])
    dE, dA, de = _adjoint(model, spec, data, obj, hs.PREC_MIXED, 1e-10, gJ)
    assert np.isfinite(de).all() and len(de) == mesh.ne
    # the 961,600 per-element CTE sensitivities add up, group by group, to the
    # group CTE sensitivities checked against differences below (a single
    # element moves J by less than the differences resolve at this size; the
    # small-box tests difference single elements)
    for g in range(len(dA)):
        assert abs(de[mesh.e_group == g].sum() - dA[g]) <= 1e-9 * np.abs(dA).max(), g
    fE, fA = _fd(model, spec, data, obj, hs.PREC_FP64, 1e-12, h=4e-3, richardson=True)
    _check(dE, {g: float(gJ @ v) for g, v in fE.items()}, 1e-4, floor=1e-2)
    _check(dA, {g: float(gJ @ v) for g, v in fA.items()}, 1e-4, floor=1e-2)


@pytest.mark.gpu
def test_node_addon_sensitivity_pair():
    """The JS host path of the pair: node -> HipSensitivity -> Node-API addon
    -> C ABI -> HIP, against central differences of its own forward."""
    import json
    import os
    import shutil
    import subprocess
    node = shutil.which("node")
    if node is None:
        pytest.skip("node not installed")
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "js", "napi_adjoint.js")
    out = subprocess.run([node, script, "plate_16"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    rec = json.loads(out.stdout.strip().splitlines()[-1])
    assert rec["rel_err"] <= 1e-5 and rec["elem_sum_rel_err"] <= 1e-10, rec
    # the backward from the caller's dJ/du (the proxy's own gradient) equals the built-in one
    assert rec["dJdu_rel_err"] <= 1e-10, rec


def test_path_field_gradient_matches_central_differences():
    """An objective of a path field (the sum of two step increments with
    different step data, as finishPath sums them) through the per-step adjoint
    of a caller-given dJ/du (fcbga_adjoint_backward_u): every group's ln E
    derivative against central differences of the two re-solved steps."""
    from fcbga_amd import sensitivity as sens
    spec = golden_spec("small_box")
    mesh = meshspec.build_mesh(spec)
    model = hs.Model(mesh, variant="hostemu")
    d1 = meshspec.group_step_data(spec, 150, 25, 0)
    d2 = meshspec.group_step_data(spec, 25, -40, 0)
    steps = [dict(Ta=150, Tb=25, data=d1), dict(Ta=25, Tb=-40, data=(2 * d2[0], 1.5 * d2[1], d2[2]))]

    def step_of(s, grp=None, f=1.0):
        D, s0, born = (a.copy() for a in s["data"])
        if grp is not None:
            D[grp] *= f
            s0[grp] *= f
        return hs.Step(s["Ta"], s["Tb"], 0, D, s0, born)
    gvec = np.random.default_rng(0).normal(size=mesh.ndof)

    def J(grp=None, f=1.0):
        u = 0
        for s in steps:
            du, _ = model.solve_step(step_of(s, grp, f), tol=1e-13, precision=hs.PREC_FP64)
            u = u + du
        return gvec @ u
    dE, _ = sens.path_field_gradient(model, steps, [0, 1], step_of, gvec, tol=1e-13)
    h = 1e-4
    fd = np.array([(J(g, np.exp(h)) - J(g, np.exp(-h))) / (2 * h) for g in range(len(d1[0]))])
    big = np.abs(fd).max()
    assert np.abs(dE - fd).max() <= 1e-5 * big, (dE, fd)
    with pytest.raises(hs.SolverError):
        model.adjoint_backward_u(np.zeros(3))
