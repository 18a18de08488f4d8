"""Thermal path: plan, step solves along it and summation (temperatureSamples,
thermalPathPoints, Analysis.evaluationTemperatures / buildPath / run /
finishPath: src/core/11-analysis.js:29-261).

Golden data (tests/golden/gen_golden.js 'path'): the reference's own functions
on sample inputs and its buildPath / finishPath methods on configuration
records of six kinds (bare, lidded, sensitivity and preview evaluations,
stacked modules with a wafer underfill or a mold), with synthetic step
increments for the sums."""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_spec
from fcbga_amd import hip_solver as hs
from fcbga_amd import meshspec
from fcbga_amd import path as P


def golden():
    with open(os.path.join(GOLDEN, "path_golden.json")) as f:
        return json.load(f)


def test_temperature_samples_and_path_points_match_reference():
    g = golden()
    for c in g["samples"]:
        assert P.temperature_samples(*c["args"]) == c["out"], c
    for c in g["points"]:
        assert P.thermal_path_points(*c["args"]) == c["out"], c


@pytest.mark.parametrize("k", range(6))
def test_build_path_and_sums_match_reference(k):
    c = golden()["plans"][k]
    cfg = dict(c["cfg"])
    if c["hasModules"]:
        cfg["modules"] = [{}]
    stage_tgs = {int(s): v for s, v in c["stageTgs"].items()}
    steps, path, temps = P.build_path(cfg, c["tgs"], stage_tgs, c["solidi"], c["evaluations"])
    assert temps == c["temps"]
    assert steps == c["steps"]
    for key in ("chain", "chainJoin", "cpi", "up", "down", "ddiJoin", "ddi"):
        assert path[key] == c["path"][key], key
    du = [np.array([k + 1, -0.5 * (k + 1) + 1e-3 * k, 0.25 * (k + 1) + 1e-3 * k * k]) for k in range(len(steps))]
    fields, u_cpi, u_ddi = P.finish_path(steps, path, du, temps["all"])
    assert sorted(fields) == sorted(float(T) for T in c["fields"])
    for T, v in c["fields"].items():
        assert np.array_equal(fields[float(T)], np.array(v)), T
    assert (u_cpi is None) == (c["uCPI"] is None) and (u_ddi is None) == (c["uDDI"] is None)
    if u_cpi is not None:
        assert np.array_equal(u_cpi, np.array(c["uCPI"]))
    if u_ddi is not None:
        assert np.array_equal(u_ddi, np.array(c["uDDI"]))


def _plate_path(variant):
    spec = golden_spec("plate_16")
    mesh = meshspec.build_mesh(spec)
    model = hs.Model(mesh, variant=variant)
    cfg = {"packageType": "bare", "process": {"Tsub": 217, "Tjoin": 217, "Tuf": 165, "Tattach": 150},
           "reflow": {"peak": 260, "reportTemps": [25, 150, 220, 260]}, "cycling": {"Tmin": -40, "Tmax": 125}}
    stage_tgs = {0: [], 1: [], 2: [], 3: []}
    steps, path, temps = P.build_path(cfg, [], stage_tgs, [217], "sensitivity")

    def step_of(s):
        return hs.step_for_spec(spec, s["Ta"], s["Tb"], 0)

    return spec, model, steps, path, temps, step_of


@pytest.mark.parametrize("variant", ["hostemu", pytest.param("hip", marks=pytest.mark.gpu)])
def test_path_run_sums_to_single_solves(variant):
    """With temperature-independent materials the thermal stress is linear in
    T, so the field at T summed along the path equals one solve from the
    stress-free temperature to T (the reference's own consistency of the
    incremental path, cfg1's 'their sum equals one solve')."""
    spec, model, steps, path, temps, step_of = _plate_path(variant)
    fields, _, _, recs = P.run_path(model, steps, path, temps["all"], step_of, tol=1e-12, precision=hs.PREC_FP64)
    assert all(r["converged"] for r in recs)
    assert recs[-1]["warm"] > 0 and min(r["r0"] for r in recs[1:]) < 1
    for T in (25, -40, 125, 260):
        ref, _ = model.solve_step(hs.step_for_spec(spec, 217, T, 0), tol=1e-13, precision=hs.PREC_FP64)
        assert np.abs(fields[T] - ref).max() <= 1e-9 * np.abs(ref).max(), T
    # the FP32 increments the stress histories use, and the same sums on the host
    du = [model.path_du(k).astype(np.float64) for k in range(len(steps))]
    f32, _, _ = P.finish_path(steps, path, du, temps["all"])
    assert np.abs(f32[125] - fields[125]).max() <= 1e-6 * np.abs(fields[125]).max()


def test_path_blocks_follow_the_worker_farm_rule():
    # runSweep (src/worker.js:114-117): max(2, ceil(n / (3 workers))) consecutive steps per block
    assert P.path_blocks(44, 3) == [list(range(k, min(44, k + 5))) for k in range(0, 44, 5)]
    assert P.path_blocks(5, 4) == [[0, 1], [2, 3], [4]]
    assert P.path_blocks(7, 1) == [[0, 1, 2], [3, 4, 5], [6]]


@pytest.mark.parametrize("variant", ["hostemu", pytest.param("hip", marks=pytest.mark.gpu)])
def test_path_over_model_instances(variant):
    """The plan's blocks over two model instances (their own warm-start bases
    and streams, one host thread each) sum to the fields of the one-model run,
    and with one instance the run is the one-model run bit for bit; an FP64
    increment handed to another model (the partial message) sums the same."""
    spec, model, steps, path, temps, step_of = _plate_path(variant)
    ref, _, _, _ = P.run_path(model, steps, path, temps["all"], step_of, tol=1e-12, precision=hs.PREC_FP64)
    one, _, _, _ = P.run_path_blocks([model], steps, path, temps["all"], step_of, tol=1e-12,
                                     precision=hs.PREC_FP64)
    for T in ref:
        assert np.array_equal(one[T], ref[T]), T
    other = hs.Model(meshspec.build_mesh(spec), variant=variant)
    two, _, _, recs = P.run_path_blocks([model, other], steps, path, temps["all"], step_of, tol=1e-12,
                                        precision=hs.PREC_FP64)
    assert all(r["converged"] for r in recs)
    scale = max(np.abs(v).max() for v in ref.values())  # (the field at the stress-free 217 C is ~0)
    for T in ref:
        assert np.abs(two[T] - ref[T]).max() <= 1e-9 * scale, T
    # step 0's increment as a message: FP64 from the model that kept it into a fresh one
    fresh = hs.Model(meshspec.build_mesh(spec), variant=variant)
    fresh.path_reset()
    d0 = model.path_sum([0])
    fresh.path_load(0, d0)
    assert np.array_equal(fresh.path_sum([0]), d0)
