"""Warm start along a thermal path (LinearSolver._warmStart/_storeSolution,
src/core/10-solver.js:885-926): each step's load is projected onto the
stored, orthonormalised previous solutions before the PCG, and a converged
solution joins the basis.  Golden: <name>_warm_golden.json and
<name>_warm_u.f64.gz from the reference's own LinearSolver on a 4-step path
with temperature-dependent moduli (a different matrix per step).  The
initial relative residual (the record's iters0) depends only on the matrix,
the load and the basis, so it is compared per step; iteration counts differ
by the smoother's colour order."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_f64, golden_spec
from fcbga_amd import hip_solver as hs
from fcbga_amd import meshspec


def _golden(name):
    with open(os.path.join(GOLDEN, name + "_warm_golden.json")) as f:
        g = json.load(f)
    U = golden_f64(name + "_warm_u.f64.gz").reshape(len(g["steps"]), -1)
    return g, U


def _step(st):
    D = np.array([x["D"] for x in st["groups"]])
    s0 = np.array([x["s0"] for x in st["groups"]])
    born = np.array([x["born"] for x in st["groups"]], np.uint8)
    return hs.Step(st["Ta"], st["Tb"], 0, D, s0, born)


def run_path(variant, name, prec):
    g, U = _golden(name)
    model = hs.Model(meshspec.build_mesh(golden_spec(name)), variant=variant)
    for k, st in enumerate(g["steps"]):
        u, rec = model.solve_step(_step(st), tol=g["tol"], maxit=3000, precision=prec, warm=20)
        assert rec["converged"] and rec["warm"] == st["warm"]
        # the basis carries each step's solver error (tol 1e-12 of the solution),
        # so later iters0 agree to that error relative to their size
        assert abs(rec["r0"] - st["iters0"]) <= 1e-5 * st["iters0"] + 1e-11, (k, rec["r0"], st["iters0"])
        assert np.abs(u - U[k]).max() <= 1e-8 * np.abs(U[k]).max()
    return model, g


@pytest.mark.parametrize("name", ["small_box", "plate_16"])
def test_warm_start_follows_reference(name):
    run_path("hostemu", name, hs.PREC_FP64)


def test_warm_reset_and_cold_solves():
    model, g = run_path("hostemu", "small_box", hs.PREC_FP64)
    st = _step(g["steps"][1])
    # warm = 0: cold, the basis is neither used nor extended
    _, rec = model.solve_step(st, tol=1e-12, maxit=3000, warm=0)
    assert rec["r0"] == 1.0 and rec["warm"] == 4
    # a solution already in the span is not stored (as the reference's
    # _storeSolution); a new one is, and the limit then drops the oldest
    _, rec = model.solve_step(st, tol=1e-12, maxit=3000, warm=2)
    assert rec["warm"] == 4 and rec["r0"] < 1e-9
    D2 = st.D.reshape(-1, 36).copy()
    D2[1] *= 0.5  # one group softer: a solution outside the span
    other = hs.Step(st.c.Ta, st.c.Tb, 0, D2, st.s0.reshape(-1, 6), st.born)
    _, rec = model.solve_step(other, tol=1e-12, maxit=3000, warm=2)
    assert rec["warm"] == 2 and rec["r0"] < 1.0
    model.warm_reset()
    _, rec = model.solve_step(st, tol=1e-12, maxit=3000, warm=20)
    assert rec["r0"] == 1.0 and rec["warm"] == 1
    with pytest.raises(hs.SolverError):
        model.solve_step(st, tol=1e-12, warm=-1)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [hs.PREC_FP64, hs.PREC_MIXED])
@pytest.mark.parametrize("name", ["small_box", "plate_16"])
def test_gpu_warm_start_follows_reference(name, prec):
    run_path("hip", name, prec)
