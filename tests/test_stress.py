"""Stress recovery: element stress histories (accumulateStep/stateStress,
src/core/09-assembly.js:494-571), the die surface stress maps and joint loads
(dieSurfaceStress/jointLoads, src/core/12-post.js:180-309).

Golden data: <name>_stress_golden.json, written by tests/golden/gen_golden.js
from the reference's own functions on its own step solution (<name>_u.f64.gz).
The oracle restates the element recovery through the strain at each point
(sum_m phi_m D E_m = D eps), an independent route to the same numbers."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_f64, golden_spec
from fcbga_amd import hip_solver as hs
from fcbga_amd import meshspec, post

NAMES = ["small_box", "cfg1"]


def stress_golden(name):
    with open(os.path.join(GOLDEN, name + "_stress_golden.json")) as f:
        return json.load(f)


def arr(v):
    return np.array([np.nan if x is None else x for x in v], np.float64)


def close(a, b, rtol, scale=None):
    """Equal NaN pattern and |a - b| <= rtol * scale (default max |b|)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(b)
    if not ok.any():
        return
    s = np.abs(b[ok]).max() if scale is None else scale
    err = np.abs(a[ok] - b[ok]).max()
    assert err <= rtol * s, (err, s)


def _setup(name):
    spec = golden_spec(name)
    mesh = meshspec.build_mesh(spec)
    step = hs.step_for_spec(spec)
    u = golden_f64(name + "_u.f64.gz")
    return spec, mesh, step, u


@pytest.mark.parametrize("name", NAMES)
def test_oracle_element_recovery_matches_reference(oracle, name):
    spec, mesh, step, u = _setup(name)
    D, s0, born = meshspec.group_step_data(spec, spec["Ta"], spec["Tb"], 0)
    g = stress_golden(name)
    for el in g["elements"]:
        e = el["e"]
        i, j, k = (int(v) for v in mesh.e_ijk[e])
        hx, hy, hz = mesh.xs[i + 1] - mesh.xs[i], mesh.ys[j + 1] - mesh.ys[j], mesh.zs[k + 1] - mesh.zs[k]
        grp = int(mesh.e_group[e])
        du24 = u.reshape(-1, 3)[mesh.conn[e]].reshape(-1)
        q = np.zeros(33)
        xez = np.array([p["xez"] for p in el["points"]])
        sig, eps = oracle.element_step_points(hx, hy, hz, oracle.shape_matrices(D=D[grp]), D[grp], s0[grp],
                                              bool(born[grp]), du24, q, xez)
        st = np.array(el["state"])
        close(q, st[:33], 1e-11)
        close(sig, np.array([p["sig"] for p in el["points"]]), 1e-10)
        close(eps, np.array([p["eps"] for p in el["points"]]), 1e-11)


def check_post(name, view, g, rtol=1e-9):
    """Die surface maps and joint loads of the view against the golden (rtol
    relative to the largest |value| of each map)."""
    spec = golden_spec(name)
    sd = spec["stress_die"]
    x0, x1, y0, y1 = sd["rect"]
    geom = {"siParts": [{"name": "die", "rect": {"x0": x0, "x1": x1, "y0": y0, "y1": y1}, "ztop": sd["ztop"],
                         "zb": sd["zb"], "gkey": sd["gkey"], "part": meshspec.PART["DIE"]}]}
    for which in ("top", "bottom"):
        r = post.die_surface_stress(view, geom, 0, which)
        ref = g["die_" + which]
        assert (r["nx"], r["ny"]) == (ref["nx"], ref["ny"])
        scale = np.nanmax(np.abs(arr(ref["smax"])))
        for key in ("smax", "sxx", "syy"):
            close(r[key], arr(ref[key]), rtol, scale)
        for key in ("peakInterior", "gpPeakInterior", "gpPeakAll"):
            close([r[key]], [ref[key]], rtol, scale)
        assert (r["peakX"], r["peakY"]) == (ref["peakX"], ref["peakY"])
        close(r["corners"], arr(ref["corners"]), rtol, scale)
        close(r["midEdges"], arr(ref["midEdges"]), rtol, scale)
    sj = spec["stress_joints"]
    field = dict(nbx=sj["nbx"], nby=sj["nby"], bx0=sj["bx0"], by0=sj["by0"], px=sj["px"], py=sj["py"],
                 peripheral=False, rows=0)
    r = post.joint_loads(view, field, sj["zmid"], sj["part"], sj["A"])
    ref = g["joints"]
    for key in ("N", "V", "szz"):
        close(r[key], arr(ref[key]), rtol)
    assert (r["worstNidx"], r["worstVidx"]) == (ref["worstNidx"], ref["worstVidx"])


@pytest.mark.parametrize("name", NAMES)
def test_hostemu_stress_post_matches_reference(name):
    spec, mesh, step, u = _setup(name)
    model = hs.Model(mesh, variant="hostemu")
    g = stress_golden(name)
    elems = [el["e"] for el in g["elements"]]
    st = model.state_accumulate(step, u, elems)
    for q, el in enumerate(g["elements"]):
        close(st[q], np.array(el["state"]), 1e-11)
        xez = np.array([p["xez"] for p in el["points"]])
        sig, eps = model.state_stress(elems, st, np.full(len(xez), q), xez)
        close(sig, np.array([p["sig"] for p in el["points"]]), 1e-11)
        close(eps, np.array([p["eps"] for p in el["points"]]), 1e-11)
    check_post(name, post.StressView(model, [(step, u)], spec["Tb"]), g)


def test_state_accumulate_is_additive_and_validates(oracle):
    """Two half steps add up to one step (linearity of the recovery in du and
    s0); q_only leaves the stress terms untouched; bad element ids and point
    rows are rejected."""
    spec, mesh, step, u = _setup("small_box")
    model = hs.Model(mesh, variant="hostemu")
    D, s0, born = meshspec.group_step_data(spec, spec["Ta"], spec["Tb"], 0)
    half = hs.Step(spec["Ta"], spec["Tb"], 0, D, 0.5 * s0, born)
    elems = np.arange(mesh.ne, dtype=np.int32)
    full = model.state_accumulate(step, u, elems)
    two = model.state_accumulate(half, 0.5 * u, elems)
    two = model.state_accumulate(half, 0.5 * u, elems, two)
    close(two, full, 1e-12)
    qo = model.state_accumulate(step, u, elems, q_only=True)
    close(qo[:, :33], full[:, :33], 1e-14)
    assert not qo[:, 33:].any()
    with pytest.raises(hs.SolverError):
        model.state_accumulate(step, u, [mesh.ne])
    with pytest.raises(hs.SolverError):
        model.state_stress(elems[:2], full[:2], [2], [[0, 0, 0]])


def test_max_principal_closed_form():
    """Smith's closed form against numpy's symmetric eigensolver, including
    diagonal and repeated-eigenvalue tensors."""
    rng = np.random.default_rng(3)
    s = rng.standard_normal((200, 6)) * 50
    s[:20, 3:] = 0
    s[20:30] = [7, 7, 7, 0, 0, 0]
    full = np.zeros((len(s), 3, 3))
    idx = [(0, 0, 0), (1, 1, 1), (2, 2, 2), (3, 0, 1), (4, 1, 2), (5, 2, 0)]
    for c, a, b in idx:
        full[:, a, b] = s[:, c]
        full[:, b, a] = s[:, c]
    ref = np.linalg.eigvalsh(full)[:, -1]
    assert np.abs(post.max_principal(s) - ref).max() <= 1e-10 * np.abs(ref).max()


def test_gp_to_node_extrapolates_trilinear_fields():
    """A trilinear field sampled at the Gauss points extrapolates exactly to
    its nodal values."""
    rng = np.random.default_rng(5)
    c = rng.standard_normal(8)
    f = lambda x, y, z: c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * y * z + c[6] * x * z + \
        c[7] * x * y * z
    gp = f(post.HEX_XI * post.GP2, post.HEX_ETA * post.GP2, post.HEX_ZETA * post.GP2)
    assert np.allclose(post.GP_TO_NODE @ gp, f(post.HEX_XI, post.HEX_ETA, post.HEX_ZETA), rtol=0, atol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_stress_post_matches_reference(name):
    spec, mesh, step, u = _setup(name)
    model = hs.Model(mesh, variant="hip")
    g = stress_golden(name)
    elems = [el["e"] for el in g["elements"]]
    st = model.state_accumulate(step, u, elems)
    for q, el in enumerate(g["elements"]):
        close(st[q], np.array(el["state"]), 1e-11)
    # the device contracts multiply-adds (FMA): sigma = sum phi S - S_th cancels
    # to a few 1e-9 of the map's range
    check_post(name, post.StressView(model, [(step, u)], spec["Tb"]), g, rtol=1e-7)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [hs.PREC_FP64, hs.PREC_MIXED])
def test_gpu_solve_to_die_peak_within_tenth_percent(prec):
    """End to end on the device: solve the bare-die step, recover the die
    surface stresses from the solution and compare the peak die and joint
    stresses with the reference's (0.1 % bar)."""
    spec, mesh, step, _ = _setup("cfg1")
    model = hs.Model(mesh, variant="hip")
    u, rec = model.solve_step(step, tol=spec["tol"], precision=prec)
    assert rec["converged"]
    g = stress_golden("cfg1")
    view = post.StressView(model, [(step, u)], spec["Tb"])
    sd = spec["stress_die"]
    x0, x1, y0, y1 = sd["rect"]
    geom = {"siParts": [{"name": "die", "rect": {"x0": x0, "x1": x1, "y0": y0, "y1": y1}, "ztop": sd["ztop"],
                         "zb": sd["zb"], "gkey": sd["gkey"], "part": meshspec.PART["DIE"]}]}
    for which in ("top", "bottom"):
        r = post.die_surface_stress(view, geom, 0, which)
        ref = g["die_" + which]
        assert abs(r["peakInterior"] - ref["peakInterior"]) <= 1e-3 * abs(ref["peakInterior"])
        assert abs(r["gpPeakAll"] - ref["gpPeakAll"]) <= 1e-3 * abs(ref["gpPeakAll"])
    sj = spec["stress_joints"]
    field = dict(nbx=sj["nbx"], nby=sj["nby"], bx0=sj["bx0"], by0=sj["by0"], px=sj["px"], py=sj["py"])
    r = post.joint_loads(view, field, sj["zmid"], sj["part"], sj["A"])
    ref = g["joints"]
    assert abs(r["worstN"] - ref["worstN"]) <= 1e-3 * abs(ref["worstN"])
    assert abs(r["worstV"] - ref["worstV"]) <= 1e-3 * abs(ref["worstV"])


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [hs.PREC_FP64, hs.PREC_MIXED])
def test_gpu_cfg2_solve_to_die_and_joint_peaks(prec):
    """The lidded 3,093,159-DOF mesh end to end on the device: the step solve
    (tolerance 1e-10), the device stress states of the die and bump-layer
    elements and the reference's post functions restated in post.py, against
    the reference's own solve and post on the same mesh
    (cfg2_stress_golden.json): peak die and joint stresses within 0.1 %, the
    die surface maps within 0.1 % of their range."""
    spec = golden_spec("cfg2")
    mesh = meshspec.build_mesh(spec)
    step = hs.step_for_spec(spec)
    model = hs.Model(mesh, variant="hip")
    u, rec = model.solve_step(step, tol=spec["tol"], precision=prec)
    assert rec["converged"]
    g = stress_golden("cfg2")
    view = post.StressView(model, [(step, u)], spec["Tb"])
    sd = spec["stress_die"]
    x0, x1, y0, y1 = sd["rect"]
    geom = {"siParts": [{"name": "die", "rect": {"x0": x0, "x1": x1, "y0": y0, "y1": y1}, "ztop": sd["ztop"],
                         "zb": sd["zb"], "gkey": sd["gkey"], "part": meshspec.PART["DIE"]}]}
    for which in ("top", "bottom"):
        r = post.die_surface_stress(view, geom, 0, which)
        ref = g["die_" + which]
        for key in ("peakInterior", "gpPeakInterior", "gpPeakAll"):
            assert abs(r[key] - ref[key]) <= 1e-3 * abs(ref[key]), (which, key, r[key], ref[key])
        for key in ("smax", "sxx", "syy"):
            close(r[key], arr(ref[key]), 1e-3)
    sj = spec["stress_joints"]
    field = dict(nbx=sj["nbx"], nby=sj["nby"], bx0=sj["bx0"], by0=sj["by0"], px=sj["px"], py=sj["py"])
    r = post.joint_loads(view, field, sj["zmid"], sj["part"], sj["A"])
    ref = g["joints"]
    assert abs(r["worstN"] - ref["worstN"]) <= 1e-3 * abs(ref["worstN"])
    assert abs(r["worstV"] - ref["worstV"]) <= 1e-3 * abs(ref["worstV"])
    close(r["N"], arr(ref["N"]), 1e-3)


@pytest.mark.gpu
def test_node_addon_state_view():
    """The JS host path of the stress histories: node -> HipStateView ->
    Node-API addon -> C ABI -> HIP, against the reference's element states and
    point stresses."""
    import shutil
    import subprocess
    node = shutil.which("node")
    if node is None:
        pytest.skip("node not installed")
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "js", "napi_stress.js")
    out = subprocess.run([node, script, "cfg1"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    rec = json.loads(out.stdout.strip().splitlines()[-1])
    assert rec["points"] == 12
    assert rec["q_rel_err"] <= 1e-11 and rec["sig_rel_err"] <= 1e-9 and rec["eps_rel_err"] <= 1e-9, rec
