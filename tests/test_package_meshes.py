"""The step solve on the reference's own package meshes: the bare-die and the
lidded default presets at the Draft mesh, built by the reference's
buildPackageMesh (graded lines, fill fractions, layered substrate bands),
with the FEModel's prepared group and section data of two thermal-path steps
(tests/golden/package_<name>_golden.json.gz, gen_golden.js 'package'), and a
stacked-module package with steps of its module stages
(package_modules_golden.json.gz, gen_golden.js 'package_modules') and its whole
thermal path with the reference's 25 C field (package_modules_path.json.gz).

Bar: the assembled load against the reference's to 1e-13, nodal
displacements within 1e-6 of the largest displacement of the reference's
solution (tol 1e-12) and JEITA warpage within 0.1 %, in FP64 and mixed
precision; the adjoint pair runs on the layered mesh and its group
gradients match central differences."""
import gzip
import os
import shutil
import subprocess
import json

import numpy as np
import pytest

import package_case as pc
from fcbga_amd import hip_solver as hs
from fcbga_amd import post

CASES = [("bare", 0), ("bare", 1), ("lidded", 0), ("lidded", 1)]


def _check(g, m, q, u, rec):
    s = g["steps"][q]
    uref = np.array(s["u"])
    assert rec["converged"]
    assert np.abs(u - uref).max() <= 1e-6 * np.abs(uref).max()
    w = post.warpage_jeita(m, u, pc.bga(g))["signed"]
    assert abs(w - s["warpage"]) <= 1e-3 * abs(s["warpage"])


@pytest.mark.parametrize("name,q", [("bare", 0), ("bare", 1)])
def test_host_emulation_solves_reference_package(name, q):
    g, m = pc.load(name)
    assert m.nsections > 0 and (m.e_section >= 0).any()
    M = hs.Model(m, variant="hostemu")
    st = pc.step(g, q)
    f = M.load(st)
    rref = np.array(g["steps"][q]["rhs"])
    assert np.abs(f - rref).max() <= 1e-13 * np.abs(rref).max()
    u, rec = M.solve_step(st, tol=1e-10, precision=hs.PREC_FP64)
    _check(g, m, q, u, rec)


def test_section_step_data_required():
    g, m = pc.load("bare")
    M = hs.Model(m, variant="hostemu")
    s = g["steps"][0]
    bad = hs.Step(s["Ta"], s["Tb"], s["stage"], np.array(s["D"]), np.array(s["s0"]), np.array(s["born"], np.uint8))
    with pytest.raises(hs.SolverError, match="layered sections"):
        M.solve_step(bad, tol=1e-8)


@pytest.mark.gpu
@pytest.mark.parametrize("name,q", CASES)
@pytest.mark.parametrize("prec", [hs.PREC_FP64, hs.PREC_MIXED])
def test_gpu_solves_reference_package(name, q, prec):
    g, m = pc.load(name)
    M = hs.Model(m, variant="hip")
    u, rec = M.solve_step(pc.step(g, q), tol=1e-10, precision=prec)
    _check(g, m, q, u, rec)


# the module package's steps on one model, module and package stages alternating, so that the
# stage's constraint set and multigrid aggregates switch both ways
MODULE_ORDER = (0, 3, 1, 4, 2)


def _check_module_step(g, q, f, u, rec, rel):
    s = g["steps"][q]
    rref, uref = np.array(s["rhs"]), np.array(s["u"])
    assert rec["converged"], (q, rec)
    # the stage's load: only born parts, zero on the held DOFs
    assert np.abs(f - rref).max() <= 1e-13 * np.abs(rref).max(), q
    assert np.abs(u - uref).max() <= rel * np.abs(uref).max(), (q, np.abs(u - uref).max() / np.abs(uref).max())
    # the module stage's held set: the module supports, the package 3-2-1 and every node no born element touches
    assert np.abs(u[np.array(s["held"])]).max() == 0, q
    # the aggregates follow the born parts: the reference's level count and coarsest size
    assert rec["levels"] == len(s["levels"]) and rec["coarsest_dof"] == s["levels"][-1], (q, rec["levels"], s["levels"])


def test_host_emulation_module_stages():
    """The stacked-module package (the reference's V18 model: one Si-base module with two
    chiplets on a 14 mm substrate): the first step of each module stage (chiplet join, wafer
    underfill, mold cure), in which only the born module parts are assembled and everything
    else is held, and two package steps, on one model.  Loads to 1e-13, displacements within
    1e-9 of the reference's (both at tol 1e-12), held DOFs exactly zero, the reference's
    multigrid levels, and its PCG iteration counts within one."""
    g, m = pc.load("modules")
    assert len(m.module_constraint_dofs) > len(m.constraint_dofs)
    M = hs.Model(m, variant="hostemu")
    stages = set()
    for q in MODULE_ORDER:
        st = pc.step(g, q)
        f = M.load(st)
        u, rec = M.solve_step(st, tol=1e-12, precision=hs.PREC_FP64)
        _check_module_step(g, q, f, u, rec, 1e-9)
        assert abs(rec["iters"] - g["steps"][q]["iters"]) <= 1, (q, rec["iters"], g["steps"][q]["iters"])
        stages.add(g["steps"][q]["stage"])
    assert stages == {-3, -2, -1, 1, 2}


def test_module_constraints_validated():
    g, m = pc.load("modules")
    m.module_constraint_dofs = np.array([3 * m.nn], np.int32)
    with pytest.raises(hs.SolverError, match="module constraint"):
        hs.Model(m, variant="hostemu")


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [hs.PREC_FP64, hs.PREC_MIXED])
def test_gpu_module_stages(prec):
    """The module package's steps on the GPU in both precisions, module and package stages
    alternating on one model (the iteration graphs captured for one stage's aggregates are
    dropped when the next stage builds its own)."""
    g, m = pc.load("modules")
    M = hs.Model(m, variant="hip")
    for q in MODULE_ORDER:
        print("module package step", q, "stage", g["steps"][q]["stage"], flush=True)
        st = pc.step(g, q)
        f = M.load(st)
        u, rec = M.solve_step(st, tol=1e-10, precision=prec)
        _check_module_step(g, q, f, u, rec, 1e-6)


def _modules_path():
    with gzip.open(os.path.join(pc.GOLDEN, "package_modules_path.json.gz")) as f:
        return json.load(f)


@pytest.mark.parametrize("variant", [pytest.param("hip", marks=pytest.mark.gpu), "hostemu"])
def test_module_package_thermal_path(variant):
    """The module package's whole thermal path as the reference's analysis runs it (preview
    evaluations: chiplet join, wafer underfill and mold cure as module stages, then the
    package stages), solved step by step with warm starts along the path (the stored
    solutions of the module stages carry into the package stages) and summed on the device:
    the 25 C field within 1e-6 of the largest displacement of the reference's and its JEITA
    warpage within 0.1 %."""
    from fcbga_amd import path as pth
    g, m = pc.load("modules")
    P = _modules_path()
    assert P["same_mesh"]
    plan = P["plan"]
    assert any(P["steps"][k]["stage"] < 0 for k in plan["chain"]) and any(P["steps"][k]["stage"] >= 0 for k in plan["chain"])
    M = hs.Model(m, variant=variant)

    def step_of(s):
        return hs.Step(s["Ta"], s["Tb"], s["stage"], np.array(s["D"]), np.array(s["s0"]), np.array(s["born"], np.uint8),
                       np.array(s["secW"]), np.array(s["secV"]), np.array(s["secBorn"], np.uint8))
    prec = hs.PREC_MIXED if variant == "hip" else hs.PREC_FP64
    fields, _, _, recs = pth.run_path(M, P["steps"], plan, P["temps"], step_of, tol=1e-10, precision=prec)
    u, uref = fields[25], np.array(P["u25"])
    assert np.abs(u - uref).max() <= 1e-6 * np.abs(uref).max(), np.abs(u - uref).max() / np.abs(uref).max()
    w = post.warpage_jeita(m, u, pc.bga(g))["signed"]
    assert abs(w - P["warpage25"]) <= 1e-3 * abs(P["warpage25"]), (w, P["warpage25"])


@pytest.mark.gpu
def test_gpu_adjoint_on_layered_mesh():
    """Group gradients of the warpage proxy on the layered bare-die mesh (the
    layered substrate group included: its bands' stiffness and thermal load
    scale together) against central differences of the forward objective."""
    g, m = pc.load("bare")
    M = hs.Model(m, variant="hip")
    s = g["steps"][1]
    kt = m.nz - 1
    base = pc.step(g, 1)
    J0, _, _ = M.adjoint_forward(base, kt, np.zeros(0, np.int32), 8.0, tol=1e-12, precision=hs.PREC_FP64)
    dlnE, dlnA, _, _ = M.adjoint_backward(np.array([1.0, 0.0]))
    D, s0 = np.array(s["D"]).reshape(-1, 36), np.array(s["s0"]).reshape(-1, 6)
    W, V = np.array(s["secW"]).reshape(-1, 108), np.array(s["secV"]).reshape(-1, 18)
    layered = [i for i, k in enumerate(g["groupKinds"]) if k == "layered"]
    h = 1e-4
    worst = 0.0
    for grp in range(len(D)):
        Jp = []
        for sg in (1, -1):
            f = np.exp(sg * h)
            D2, s02, W2, V2 = D.copy(), s0.copy(), W.copy(), V.copy()
            if grp in layered:
                W2 *= f
                V2 *= f
            else:
                D2[grp] *= f
                s02[grp] *= f
            st = hs.Step(s["Ta"], s["Tb"], s["stage"], D2, s02, np.array(s["born"], np.uint8), W2, V2,
                         np.array(s["secBorn"], np.uint8))
            J, _, _ = M.adjoint_forward(st, kt, np.zeros(0, np.int32), 8.0, tol=1e-12, precision=hs.PREC_FP64)
            Jp.append(J[0])
        fd = (Jp[0] - Jp[1]) / (2 * h)
        worst = max(worst, abs(fd - dlnE[grp]) / max(np.abs(dlnE).max(), 1e-300))
    assert worst <= 1e-4, worst


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["bare", "lidded", "modules"])
def test_node_addon_install_on_reference_package(name):
    """The JS drop-in on a reference-built mesh: install(Analysis) replaces
    solveStep, the step data of the layered sections come from the model's
    own sections (sectionWmean) and prepared fields, and both steps' du match
    the reference's solution."""
    node = shutil.which("node")
    if node is None:
        pytest.skip("node not installed")
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "js", "napi_package.js")
    out = subprocess.run([node, script, name], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    rec = json.loads(out.stdout.strip().splitlines()[-1])
    for r in rec["steps"]:
        assert r["converged"] and r["rel_err"] <= 1e-6, rec
