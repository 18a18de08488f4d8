#!/usr/bin/env python3
"""Step-solve benchmark of the MI355X hex8 thermoelastic solver.

Workload: the lidded full-fidelity FCBGA (160 x 160 x 52 hexes, graded,
3,093,159 DOF) solved cold from the 150 C stress-free state to each of six
temperatures 25..260 C; one "step" is one complete step solve (device
assembly of the condensed element tables and load, multigrid set-up, PCG) in
mixed precision (FP32 multigrid cycle inside FP64 flexible PCG), to the
tolerance at which the JEITA warpage is within 1e-6 of the FP64 oracle (a
tight FP64 solve, tol 1e-12, of the same step): an untimed calibration pass
takes, for each precision, the loosest tolerance of 1e-4 .. 1e-10 that meets
1e-6, and every timed solve's warpage error is reported.  The mesh is
uploaded before the timed region; the displacements stay in HBM.  The six
steps are timed one at a time on one model instance (solve_s_mixed,
value_one_at_a_time) and with three solves in flight on three instances
(each its own stream, HIP graphs and host thread), as the reference spreads
the six temperatures over its workers; value = sum(DOF x PCG iterations) /
timed seconds of the latter.

Beside the timed mixed-precision solves the run also reports the plain FP64
PCG-MG solve of the same steps at its own 1e-6-warpage tolerance (the
comparator), the dominant kernel's roofline and the composite per-iteration
roofline from HIP-event timing, the bump-array viscoplastic cycling workload
(120 steps of the global Newton), a sample of the 256 batched bump
submodels (60 steps each, several at once), and the CPU oracle (a single-threaded C
restatement of the reference's FP64 PCG-MG) on a bounded sample of the same
workload, one process per temperature.

  python bench.py [--gpus 1] [--steps 6] [--warmup 1] [--config cfg2|cfg1] [--concurrent 3]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "hbm": 8.0e12}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def incidences_per_node(mesh):
    """Element-node incidences of every grid node (8 per element in total)."""
    nx, ny, nz = mesh.nx, mesh.ny, mesh.nz
    inc = np.zeros((nx, ny, nz), np.int64)
    slots = (mesh.cell_elems.reshape(nx - 1, ny - 1, nz - 1, 2) >= 0).sum(axis=3)
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                inc[di:di + nx - 1, dj:dj + ny - 1, dk:dk + nz - 1] += slots
    return inc


def traffic_from_profile(kernel):
    """HBM bytes per launch of a kernel from the committed counter summary
    (profiles/<kernel>_cfg2_hbm_traffic.txt: FETCH_SIZE and WRITE_SIZE per
    dispatch, scaled by the factors a plain copy of known size reads)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_cfg2_hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def profile_values(name):
    """key = value lines of a committed counter summary under profiles/ (None if absent)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            try:
                vals[k.strip()] = float(v.strip())
            except ValueError:
                pass
    return vals


SIMDS, VALU_CYCLES, CLOCK_HZ = 1024, 4, 2.4e9  # MI355X: 256 CUs x 4 SIMD16, wave64 FP32 op = 4 cycles


def valu_issue_roofline(avg_launch_us):
    """The fine sweep against its VALU issue bound: the committed counter figure
    of VALU instructions per launch (profiles/k_fine_col_sweep_cfg2_valu_issue.txt)
    x 4 cycles / 1,024 SIMDs at 2.4 GHz, over the measured launch time."""
    v = profile_values("k_fine_col_sweep_cfg2_valu_issue.txt")
    if not v or "valu_insts_per_launch" not in v:
        return None
    bound_us = v["valu_insts_per_launch"] * VALU_CYCLES / SIMDS / CLOCK_HZ * 1e6
    return {"bound": "valu_issue", "bound_us": bound_us, "avg_launch_us": avg_launch_us,
            "frac": bound_us / avg_launch_us, "valu_insts_per_wave": v["valu_insts_per_launch"] / v["waves_per_launch"],
            "note": "FP32 VALU instructions per launch from SQ_INSTS_VALU; packed FP32 is not used"}


def roofline_for(cls, ms_total, launches, mesh, dtype_cycle):
    """Roofline of a kernel class from its HIP-event time per launch.
    fine_sweep (the z-line smoother, one launch per colour): HBM-bound, 36 B per
    swept DOF (FP32 column factors 24, x read/write 8, b 4) plus the 4 B FP32 x
    of every DOF that the colour's residual reads.  apply / fine_resid (the
    operator): compute-bound, 144 flop per (node, element) incidence (3 rows x
    24 columns of the condensed matrix, 2 flop per MAC), 1152 per element."""
    inc = incidences_per_node(mesh)
    avg_s = ms_total / launches / 1e3
    if cls == "fine_sweep":
        exist = mesh.node_id.reshape(mesh.nx, mesh.ny, mesh.nz) >= 0
        per = [exist[a::2, b::2, :].sum() for a in (0, 1) for b in (0, 1)]
        nbytes = 3 * (36.0 * float(np.mean(per)) + 4.0 * mesh.nn)
        achieved = nbytes / avg_s
        tr = traffic_from_profile("k_fine_col_sweep")
        return {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK["hbm"] / 1e9, "unit": "GB/s",
                "frac": achieved / PEAK["hbm"], "traffic": tr, "kernel": cls, "dtype": dtype_cycle,
                "avg_launch_us": avg_s * 1e6, "launches": launches, "bytes_per_launch": nbytes,
                "issue": valu_issue_roofline(avg_s * 1e6)}
    if cls == "apply":
        flops = 144.0 * inc.sum()
        peak, unit_dt = PEAK["fp64"], "fp64"
    elif cls == "fine_resid":
        flops = 144.0 * inc.sum()
        peak, unit_dt = PEAK[dtype_cycle], dtype_cycle
    else:
        return None
    achieved = flops / avg_s
    return {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": None, "kernel": cls, "dtype": unit_dt,
            "avg_launch_us": avg_s * 1e6, "launches": launches, "flop_per_launch": flops}


def composite_roofline(ktimes, ksolves, rec_iters, ms_pcg, mesh, ncycle_sweeps=4):
    """Algorithmic time of one PCG iteration of the mixed solve over its
    measured time: sum over the fine-level kernels of max(flop / peak,
    bytes / HBM bandwidth) with the per-unit figures of the hot-path table
    (outer FP64 apply: 1152 flop per element, 16 B per DOF; z-line sweeps,
    2 pre + 2 post per cycle: 36 B per DOF each; FP32 residual apply: 1152 flop
    per element, 8 B per DOF; restriction + prolongation: 56 B per DOF; PCG
    vectors: 64 B per DOF, the fused minimum).  The shell levels and the
    coarsest solve are left out of the algorithmic sum (they are counted in
    the measured time), so the fraction is an upper bound of the fine
    level's share."""
    N, E = mesh.ndof, mesh.ne
    terms = {
        "outer_apply_fp64": max(1152.0 * E / PEAK["fp64"], 16.0 * N / PEAK["hbm"]),
        "fine_sweeps": ncycle_sweeps * 36.0 * N / PEAK["hbm"],
        "fine_residual_fp32": max(1152.0 * E / PEAK["fp32"], 8.0 * N / PEAK["hbm"]),
        "transfers": 56.0 * N / PEAK["hbm"],
        "pcg_vectors": 64.0 * N / PEAK["hbm"],
    }
    t_alg = sum(terms.values())
    t_meas = ms_pcg / 1e3 / max(rec_iters, 1)
    return {"alg_us_per_iter": 1e6 * t_alg, "measured_us_per_iter": 1e6 * t_meas, "frac": t_alg / t_meas,
            "terms_us": {k: 1e6 * v for k, v in terms.items()},
            "kernel_ms_per_iter": {k: v[0] / max(1, ksolves * rec_iters) for k, v in ktimes.items() if v[1]},
            "note": "fine-level kernels only in the algorithmic sum; shell levels, coarsest solve and launch gaps "
                    "are in the measured time"}


def warpage_oracle(model, step, du, mesh, bga, post):
    """FP64 oracle of a step: the same solve in FP64 to 1e-12."""
    model.solve_step_device(step, du.data_ptr(), tol=1e-12, maxit=4000, precision=0, precond=0)
    u = du.cpu().numpy().copy()
    return u, post.warpage_jeita(mesh, u, bga)["signed"]


def calibrate_tol(solve_fn, du, mesh, bga, post, w_ref, target=1e-6):
    """Loosest tolerance of 1e-4 .. 1e-10 whose warpage is within target of w_ref."""
    ladder = [1e-4, 1e-5, 1e-6, 1e-7, 1e-8, 1e-9, 1e-10]
    tried = []
    for tol in ladder:
        rec = solve_fn(tol)
        u = du.cpu().numpy()
        err = abs(post.warpage_jeita(mesh, u, bga)["signed"] - w_ref) / abs(w_ref)
        tried.append({"tol": tol, "iters": rec["iters"], "warpage_err": err})
        if err <= target:
            return tol, tried
    return ladder[-1], tried


def cfg3_timing(hs, cases, meshspec, nsteps=None):
    """The bump-array viscoplastic cycling workload: the global Newton through
    the 120-step history (two cycles -40 <-> 125 C), mixed-precision inner
    solves; per-bump inelastic work density of each cycle."""
    from fcbga_amd import viscoplastic as vpm
    spec = cases.cfg3_spec()
    mesh = meshspec.build_mesh(spec)
    V = hs.ViscoplasticModel(mesh, vpm.solder_of(spec), variant="hip")
    rec = vpm.run(V, spec, nsteps=nsteps)
    V.close()
    out = {k: rec[k] for k in ("steps", "solves", "cutbacks", "newton", "pcg_iters", "s", "ms_per_step",
                               "ms_setup", "ms_pcg")}
    out.update({"workload": spec["name"], "ndof": mesh.ndof, "solder_elements": int((mesh.e_group ==
                spec["solder_group"]).sum()), "points": 8 * int((mesh.e_group == spec["solder_group"]).sum()),
                "newton_per_step": rec["newton"] / max(rec["solves"], 1),
                "pcg_per_newton": rec["pcg_iters"] / max(rec["newton"], 1)})
    pc = rec["per_cycle"]
    if pc:
        dW = pc[-1]["dW"]
        b = spec["bumps"]["n"]
        corner = (b - 1) * b + (b - 1)  # the bump nearest the die corner
        out["dW_last_cycle"] = {"max": float(np.nanmax(dW)), "mean": float(np.nanmean(dW)),
                                "corner_bump": float(dW[corner]), "argmax_bump": int(np.nanargmax(dW)),
                                "per_cycle_max": [float(np.nanmax(c["dW"])) for c in pc], "unit": "MPa"}
    return out


def cfg5_timing(hs, cases, nsub=8, workers=4):
    """Batched bump submodels: the first nsub of the 256 seeded sphere-cap voxel
    submodels (24^3 .. 56^3 voxels), each through its 60-step cycle, `workers`
    at once (one host thread and device stream each).  Reports the sample's
    per-submodel times (p50, p99), its aggregate DOF x PCG iterations per
    second, and the 256-submodel time extrapolated from the sample in
    proportion to DOF (labelled as extrapolated)."""
    from fcbga_amd import viscoplastic as vpm
    draws = cases.cfg5_draws()
    specs = [cases.cfg5_spec(d) for d in draws[:nsub]]
    t0 = time.perf_counter()
    recs = vpm.run_submodel_batch(specs, workers=workers)
    wall = time.perf_counter() - t0
    times = np.array([r["wall_s"] for r in recs])
    dof = np.array([r["ndof"] for r in recs], np.float64)
    dof_all = float(sum(3 * (d["n"] + 1) ** 3 for d in draws))
    return {"submodels_total": len(draws), "sample": nsub, "workers": workers, "steps": recs[0]["steps"],
            "wall_s": wall, "p50_s": float(np.percentile(times, 50)), "p99_s": float(np.percentile(times, 99)),
            "dof_sample": int(dof.sum()), "dof_all": int(dof_all),
            "dof_iters_per_s": float(sum(r["ndof"] * r["pcg_iters"] for r in recs) / wall),
            "gp_updates_per_s": float(sum(8 * r["solder_elements"] * r["newton"] for r in recs) / wall),
            "newton_per_step": float(np.mean([r["newton"] / max(r["solves"], 1) for r in recs])),
            "cutbacks": int(sum(r["cutbacks"] for r in recs)),
            "dW_MPa": [r["dW"] for r in recs], "n": [r["n"] for r in recs],
            "all_256_s_extrapolated": wall * dof_all / dof.sum(),
            "note": "seeded std::mt19937_64 draws (seed 20261203); the 256-submodel time is extrapolated from the "
                    "sample in proportion to DOF"}


def _cpu_sample_child(args):
    """One process of the CPU baseline: the oracle's cold solve of one
    temperature, `iters` PCG iterations."""
    spec_name, T, iters = args
    sys.path.insert(0, ROOT)
    ge.package()
    from oracle import oracle as orc
    from fcbga_amd import cases, meshspec
    spec = cases.cfg2_spec() if spec_name == "cfg2" else cases.cfg1_spec()
    mesh = meshspec.build_mesh(spec)
    D, s0, born = meshspec.group_step_data(spec, spec["Ta"], T, 0)
    t0 = time.perf_counter()
    sysm = orc.System(mesh, D, s0, born)
    _, st = sysm.solve(tol=1e-30, maxit=iters)
    return mesh.ndof, st["iters"], time.perf_counter() - t0


def host_facts():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": ncpu}


def cpu_baseline(spec, mesh, iters=5):
    """Single-threaded C restatement of the reference FP64 PCG-MG on the same
    mesh: assembly + multigrid set-up + `iters` PCG iterations (bounded sample)."""
    from oracle import oracle as orc
    from fcbga_amd import meshspec
    if not os.path.exists(orc.LIB_PATH):
        orc.build()
    D, s0, born = meshspec.group_step_data(spec, spec["Ta"], spec["Tb"], 0)
    t0 = time.perf_counter()
    sysm = orc.System(mesh, D, s0, born)
    u, st = sysm.solve(tol=1e-30, maxit=iters)
    dt = time.perf_counter() - t0
    per_iter = st["ms_pcg"] / 1e3 / max(st["iters"], 1)
    return {"value": mesh.ndof * st["iters"] / dt, "unit": "DOF*iter/s", "cores": 1, "kind": "port",
            "sample": f"{spec['name']} full mesh ({mesh.ndof} DOF), one cold step 150->25 C: assembly + "
                      f"multigrid set-up + {st['iters']} PCG iterations of the oracle (C, FP64, 1 thread)",
            "seconds": dt, "assemble_s": st["ms_assemble"] / 1e3, "setup_s": st["ms_setup"] / 1e3,
            "pcg_s_per_iter": per_iter}


def anand_throughput(torch, hs, n=115200, reps=20):
    """Batched Anand update (FP64) on device-resident states: point updates per second."""
    P = hs.anand_params(5.87e6, 7460.0, 2.0, 0.0942, 58.3, 0.015, 9350.0, 1.5, 45.9, 0.4)
    g = torch.Generator(device="cpu").manual_seed(5)
    eps = (torch.randn(n, 6, generator=g, dtype=torch.float64) * 3e-3).cuda()
    state = torch.zeros(n, 11, dtype=torch.float64)
    state[:, 6] = 45.9
    state = state.cuda()
    trial = torch.empty_like(state)
    sig = torch.empty(n, 6, dtype=torch.float64, device="cuda")
    D = torch.empty(n, 36, dtype=torch.float64, device="cuda")
    dep = torch.empty(n, dtype=torch.float64, device="cuda")
    L = hs.load("hip")
    import ctypes as C

    def run():
        rc = L.fcbga_anand_update_device(C.byref(P), n, 25.0, 45.0, 46500.0, C.c_void_p(eps.data_ptr()),
                                         C.c_void_p(state.data_ptr()), C.c_void_p(trial.data_ptr()),
                                         C.c_void_p(sig.data_ptr()), C.c_void_p(D.data_ptr()),
                                         C.c_void_p(dep.data_ptr()), None)
        if rc != 0:
            raise RuntimeError(hs.last_error(L))
    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    return {"n_points": n, "ms_per_update": dt * 1e3, "gp_updates_per_s": n / dt, "dtype": "f64",
            "note": "AnandPoint.step with consistent tangent, one batch per global Newton iteration"}


def screening_case(scr):
    """Bump screening at the README's size: a 132 x 132 die bump field (17,424
    joints), the reference's default cycle (-40..125 C, 10 C/min, 15 min
    dwells, 3 cycles: 245 steps), SAC305 Anand constants, and a synthetic
    bump-layer strain table at five temperatures."""
    import numpy as np
    solder = scr.MaterialTables([-40, 25, 75, 125], [57.4, 46.5, 38.2, 30.0], 22.5, 0.4)
    d = {"nbx": 132, "nby": 132, "bx0": 0.1, "by0": 0.1, "px": 0.15, "py": 0.15, "peripheral": False, "rows": 0}
    sites, _ = scr.screening_sites(d)
    temps = [-40, 0, 25, 75, 125]
    x, y = sites["x"][:, None], sites["y"][:, None]
    dT = np.array(temps, np.float64)[None, :] - 25.0
    table = np.stack([1.6e-5 * dT * (1 + 0.2 * np.sin(x)), 1.3e-5 * dT * (1 + 0.2 * np.cos(y)),
                      (2.4e-5 + 2e-6 * x) * dT, 3e-6 * dT * np.sin(x + y), 6e-6 * dT * (y - 10),
                      6e-6 * dT * (x - 10)], axis=-1)
    hist = scr.cycle_history({"Tmin": -40, "Tmax": 125, "ramp": 10, "dwell": 15, "cycles": 3})
    return scr.step_data(hist, temps, solder), table, temps.index(25)


def screening_timing(hs, scr, reps=5):
    """fcbga_screen_joints on the screening case (host arrays in and out)."""
    P = hs.anand_params(5.87e6, 7460.0, 2.0, 0.0942, 58.3, 0.015, 9350.0, 1.5, 45.9, 0.4)
    steps, table, i25 = screening_case(scr)
    hs.screen_joints(P, steps, table, i25)
    t0 = time.perf_counter()
    for _ in range(reps):
        dW, _, _ = hs.screen_joints(P, steps, table, i25)
    dt = (time.perf_counter() - t0) / reps
    n = table.shape[0]
    return {"joints": n, "steps": steps["nsteps"] - 1, "cycles": steps["cycles"], "ms": dt * 1e3,
            "joint_steps_per_s": n * (steps["nsteps"] - 1) / dt, "dtype": "f64",
            "reference_s": 15.0, "reference_note": "README.md:378, 17,424 bumps, JS",
            "note": "one thread per joint through the cycle (stepColumn Newton around the Anand return map); "
                    "includes the host copies of the strain table and results"}


# the property curve each material of the lidded workload follows: the index of
# the group of the reference's lidded default preset (its prepared grids,
# tests/golden/prepare_golden.json) whose E(T) / E(25 C) and thermal strain
# shape it takes: the underfill's glass transition for the polymers, the
# silicon's rising CTE, the TIM's constant properties, the copper lid's CTE
# and the lid adhesive's curve
PATH_CURVE = {"buildup": 2, "core": 5, "bump_uf": 2, "si": 3, "tim": 4, "lid_cu": 5, "lid_adh": 6}


def path_groups(spec):
    """Temperature-dependent step data for the lidded workload's materials
    (the inputs of fcbga_prepare_step): each material's modulus and CTE at
    25 C scaled along the 1 C grid by the shape of the reference preset group
    PATH_CURVE names, E(T) = E25 Eref(T) / Eref(25) and F(T) = Fref(T)
    alpha25 / alpha_ref(25)."""
    import gzip
    from fcbga_amd import hip_solver as hs
    from fcbga_amd.meshspec import column_d, iso_d
    gold = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(gold, "prepare_golden.json")) as f:
        lay = json.load(f)
    with open(os.path.join(gold, "prepare_grids.f64.gz"), "rb") as f:
        flat = np.frombuffer(gzip.decompress(f.read()), dtype="<f8")
    n, i25 = hs.TGRID_N, 25 - hs.TGRID_MIN
    groups = []
    for mt in spec["materials"]:
        ref = lay["groups"][PATH_CURVE[mt["name"]]]
        F = flat[ref["F"][0]:ref["F"][0] + ref["F"][1]].reshape(n, 3)
        a25 = 0.5 * (F[i25 + 1, 0] - F[i25 - 1, 0])
        g = {"stage": mt.get("stage", 0), "F": F * (mt["cte"] * 1e-6 / a25)}
        Dfn = column_d if mt.get("column") else iso_d
        if ref["Dgrid"] is not None:
            Dr = flat[ref["Dgrid"][0]:ref["Dgrid"][0] + ref["Dgrid"][1]].reshape(n, 36)
            shape = Dr[:, 0] / Dr[i25, 0]
            g["Dgrid"] = np.stack([Dfn(mt["E"] * f, mt["nu"]).reshape(-1) for f in shape])
        else:
            g["Dconst"] = Dfn(mt["E"], mt["nu"]).reshape(-1)
        groups.append(g)
    return groups


def path_timing(torch, hs, model, spec, tol):
    """The lidded package's thermal path (Analysis.run): the reference's plan
    for a lidded package without Tg values (process 217 -> 165 -> 150 -> 25 C,
    sweeps to -40 and 260 C), every step solved on the device with one
    warm-start basis of 20 solutions along the path, fields summed in device
    memory.  Step data from temperature-dependent property grids
    (path_groups through fcbga_prepare_step), so that each step's stiffness
    and load shape differ; the same path with the 25 C properties throughout
    is timed beside it."""
    from fcbga_amd import path as P
    cfg = {"packageType": "lid", "process": {"Tsub": 217, "Tjoin": 217, "Tuf": 165, "Tattach": 150},
           "reflow": {"peak": 260, "reportTemps": [25, 150, 220, 260]}, "cycling": {"Tmin": -40, "Tmax": 125}}
    steps, path, temps = P.build_path(cfg, [], {0: [], 1: [], 2: [], 3: []}, [217], "full")
    groups = path_groups(spec)

    def step_of(s):
        return hs.prepare_step(groups, [], s["Ta"], s["Tb"], 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, _, rc = P.run_path(model, steps, path, temps["all"],
                              lambda s: hs.step_for_spec(spec, s["Ta"], s["Tb"], 0), tol=tol)
    torch.cuda.synchronize()
    dtc = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fields, _, _, recs = P.run_path(model, steps, path, temps["all"], step_of, tol=tol)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    iters = [r["iters"] for r in recs]
    # the worker-farm form: blocks of steps over three model instances at once
    extra = [hs.Model(model.mesh, variant="hip") for _ in range(2)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fb, _, _, rb = P.run_path_blocks([model] + extra, steps, path, temps["all"], step_of, tol=tol)
    torch.cuda.synchronize()
    dtb = time.perf_counter() - t0
    scale = max(np.abs(v).max() for v in fields.values())
    diff = max(np.abs(fb[T] - fields[T]).max() for T in fields) / scale
    del extra
    return {"steps": len(steps), "fields": len(fields), "tol": tol, "s": dt, "s_per_step": dt / len(steps),
            "iters_total": int(sum(iters)), "iters_mean": float(np.mean(iters)),
            "iters0_mean_warm": float(np.mean([r["r0"] for r in recs[1:]])),
            "materials": "temperature-dependent: each material's 25 C modulus and CTE along the property curve "
                         "of a group of the reference's lidded default preset (glass transitions of the polymers)",
            "note": "one warm-started mixed solve per step (basis of 20), fresh multigrid per step",
            "constant_properties": {"s": dtc, "iters_total": int(sum(r["iters"] for r in rc)),
                                    "note": "the same path with the 25 C properties throughout: every increment is "
                                            "then a multiple of one load shape and the projection onto the stored "
                                            "solutions already meets the tolerance on most steps"},
            "blocks": {"instances": 3, "s": dtb, "iters_total": int(sum(r["iters"] for r in rb)),
                       "fields_rel_diff_vs_one": float(diff),
                       "note": "blocks of consecutive steps from one queue over three model instances, each with "
                               "its own warm-start basis, stream and host thread (the reference's worker farm)"}}


def submodel_timing():
    """One bump submodel through two temperature cycles (BumpSubmodel.run):
    the corner bump of the bare-die preset set up by the reference's own code
    (8 x 8 x 12 B-bar elements, 3159 DOF, committed case data), 686 steps with
    4 cutbacks.  The reference's run of the same case took 61.5 s in Node on
    the build host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from submodel_case import build
    g, m, _ = build("hip")
    t0 = time.perf_counter()
    r = m.run()
    dt = time.perf_counter() - t0
    m.close()
    ref = g["result"]
    # the batch form in a process of its own (its streams then have the device's
    # hardware queues to themselves): eight copies of the submodel at once
    nb = 8
    env = dict(os.environ, GPU_MAX_HW_QUEUES="16")
    res = subprocess.run([sys.executable, "-u", os.path.abspath(__file__), "--submodel-batch", str(nb)],
                         capture_output=True, text=True, env=env, timeout=600)
    if res.returncode != 0:
        raise RuntimeError("submodel batch run failed: " + res.stderr[-2000:])
    b = json.loads(res.stdout.strip().splitlines()[-1])
    b["vs_one_at_a_time"] = nb * dt / b["s"]
    b["same_records"] = bool(b["same_records"] and b["dW"] == float(r["dW"]))
    return {"ndof": ref["ndof"], "steps": r["steps"], "cutbacks": r["cutbacks"], "s": dt,
            "ms_per_step": 1e3 * dt / r["steps"], "newton_per_step": r["newtonPerStep"],
            "dW_rel_err_vs_reference": abs(r["dW"] - ref["dW"]) / abs(ref["dW"]), "reference_s": 61.5,
            "note": "Newton + IC(0)-PCG per step, the CG in one 1024-thread workgroup (latency-bound at this size)",
            "batch": b}


def submodel_batch_child(nb):
    """The batch measurement's own process: nb copies of the corner-bump
    submodel run at once (submodel.run_batch, one host thread and device stream
    each; the parent sets GPU_MAX_HW_QUEUES=16 so that the streams do not share
    hardware queues).  Prints one JSON line."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    ge.package()
    from submodel_case import build
    from fcbga_amd.submodel import run_batch
    ms = [build("hip")[1] for _ in range(nb)]
    t0 = time.perf_counter()
    recs = run_batch(ms, workers=nb)
    dtb = time.perf_counter() - t0
    for x in ms:
        x.close()
    print(json.dumps({"submodels": nb, "s": dtb, "submodels_per_s": nb / dtb, "dW": float(recs[0]["dW"]),
                      "same_records": bool(all(x["dW"] == recs[0]["dW"] and x["steps"] == recs[0]["steps"]
                                               for x in recs)),
                      "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                      "note": "one host thread and device stream per submodel, in a process of its own"}))


def adjoint_timing(torch, hs, cases, meshspec, tol, reps=3):
    """Sensitivity pair on the lidded mesh with 18 groups: one forward (solve +
    objectives) and one backward (adjoint solve + per-element contractions)
    give the gradients for 36 group parameters and every element's CTE; central
    differences on the 36 group parameters take 72 forward solves."""
    spec = cases.cfg4_spec()
    mesh = meshspec.build_mesh(spec)
    model = hs.Model(mesh, variant="hip")
    st = hs.step_for_spec(spec)
    obj = spec["objective"]
    vm = np.nonzero(np.isin(mesh.e_group, obj["vm_groups"]))[0].astype(np.int32)
    gJ = np.array([1.0, 1e-3])
    tf, tb, itf, itb = [], [], [], []
    for r in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        J, _, rf = model.adjoint_forward(st, obj["k_top"], vm, obj["p"], tol=tol, precision=hs.PREC_MIXED)
        t1 = time.perf_counter()
        _, _, _, rb = model.adjoint_backward(gJ)
        t2 = time.perf_counter()
        if r:
            tf.append(t1 - t0); tb.append(t2 - t1); itf.append(rf["iters"]); itb.append(rb["iters"])
    fwd, bwd = float(np.mean(tf)), float(np.mean(tb))
    nparam = 2 * len(mesh.groups)
    return {"workload": f"{spec['name']}: warpage proxy + die von Mises p-norm (p = {obj['p']:g})",
            "params": nparam, "elem_params": mesh.ne, "forward_s": fwd, "backward_s": bwd,
            "forward_iters": itf[-1], "adjoint_iters": itb[-1], "fd_solves": 2 * nparam,
            "fd_equivalent_s": 2 * nparam * fwd, "speedup_vs_fd": 2 * nparam * fwd / (fwd + bwd),
            "J": [float(J[0]), float(J[1])]}


def operator_compare(hs, mesh, step, tol, solves=2):
    """The fine operator's element products on the matrix cores
    (FCBGA_ELEM_OP=mfma, type-batched v_mfma_16x16x4) against the node-row
    kernels on the vector units (the default): one model each, kernels
    launched eagerly with HIP-event timing, `solves` mixed solves each.
    Per-launch times of the outer FP64 operator and the cycle's FP32
    residual; the iteration counts must agree."""
    res = {}
    saved = os.environ.get("FCBGA_ELEM_OP")
    try:
        for op in ("mfma", "valu"):
            os.environ["FCBGA_ELEM_OP"] = op
            M = hs.Model(mesh, variant="hip")
            M.set_kernel_timing(True)
            for _ in range(solves):
                _, r = M.solve_step(step, tol=tol, precision=hs.PREC_MIXED)
            kt = M.kernel_times()
            res[op] = {"iters": r["iters"],
                       "apply_fp64_us": 1e3 * kt["apply"][0] / max(1, kt["apply"][1]),
                       "residual_fp32_us": 1e3 * kt["fine_resid"][0] / max(1, kt["fine_resid"][1])}
            M.close()
    finally:
        if saved is None:
            os.environ.pop("FCBGA_ELEM_OP", None)
        else:
            os.environ["FCBGA_ELEM_OP"] = saved
    res["default"] = "valu"
    res["note"] = ("matrix-core products do 2.6x the algorithmic work (24 rows padded to 32, type groups of 16, "
                   "halo cells of 4x4-column tiles and 8-plane chunks computed twice); the node-row kernels stay "
                   "the default")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg2", choices=["cfg1", "cfg2"])
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--sweeps", type=int, default=0, help="smoothing sweeps per level (0 = the reference's 2)")
    ap.add_argument("--concurrent", type=int, default=3,
                    help="solves in flight in the throughput pass (model instances, one stream and host thread each)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-path", action="store_true")
    ap.add_argument("--no-submodel", action="store_true")
    ap.add_argument("--no-cfg3", action="store_true")
    ap.add_argument("--no-cfg5", action="store_true")
    ap.add_argument("--cfg5-sample", type=int, default=8, help="submodels of the 256 timed in the batched leg")
    ap.add_argument("--fixed-tol", dest="tol_from_warpage", action="store_false",
                    help="time at --tol instead of the tolerance calibrated to 1e-6 warpage error")
    ap.add_argument("--submodel-batch", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--operator-compare", action="store_true",
                    help="only the matrix-core vs node-row operator comparison (eager launches), one JSON line")
    ap.add_argument("--no-operator-compare", action="store_true")
    ap.add_argument("--smoother", default="coarse-jacobi", choices=["coarse-jacobi", "gs"],
                    help="multigrid smoother: damped column Jacobi from the third shell level down (one launch "
                         "per sweep), or the reference's colour Gauss-Seidel on every level; the line also "
                         "times the other one, one solve at a time")
    ap.add_argument("--counter-pass", type=int, default=0,
                    help="only N mixed solves at the calibrated tolerance (1e-7) with every kernel launched "
                         "eagerly, no captured graphs: the workload for hardware-counter runs")
    args = ap.parse_args()
    if args.submodel_batch:
        submodel_batch_child(args.submodel_batch)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the solve runs on a single MI355X and this build has no "
              f"multi-GPU path (asked for --gpus {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    import torch
    ge.package()
    from fcbga_amd import cases, meshspec, post
    from fcbga_amd import hip_solver as hs

    spec = cases.cfg2_spec() if args.config == "cfg2" else cases.cfg1_spec()
    temps = spec.get("temperatures", [spec["Tb"]])
    mesh = meshspec.build_mesh(spec)
    log(f"{spec['name']}: {mesh.ndof} DOF, {mesh.ne} elements")
    du = torch.empty(mesh.ndof, dtype=torch.float64, device="cuda")  # torch's device context first
    steps = [hs.step_for_spec(spec, Ta=spec["Ta"], Tb=temps[k % len(temps)]) for k in range(max(args.steps, 1))]
    if args.operator_compare:
        print(json.dumps({"operator_compare": operator_compare(hs, mesh, steps[0], 1e-7)}), flush=True)
        return
    if args.counter_pass:
        m = hs.Model(mesh, variant="hip")
        m.set_kernel_timing(True)  # eager launches, each bracketed by events
        for k in range(args.counter_pass):
            rec = m.solve_step_device(steps[k % len(steps)], du.data_ptr(), tol=1e-7, maxit=2000,
                                      precision=hs.PREC_MIXED, precond=hs.PC_MG,
                                      smoother=hs.SMOOTH_COARSE_JACOBI if args.smoother == "coarse-jacobi"
                                      else hs.SMOOTH_GS)
            log(f"counter pass {k}: iters {rec['iters']}")
        # a plain 256 MiB device copy: the reference point for scaling hardware byte counters
        src = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda").fill_(1.0)
        torch.empty_like(src).copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"counter_pass": args.counter_pass, "kernel_classes": m.kernel_times()}), flush=True)
        return
    model = hs.Model(mesh, variant="hip")
    bga = meshspec.bga_sites(**spec["bga"])
    tols = {hs.PREC_MIXED: args.tol, hs.PREC_FP64: args.tol}

    SM = hs.SMOOTH_COARSE_JACOBI if args.smoother == "coarse-jacobi" else hs.SMOOTH_GS

    def solve(k, prec, tol=None, smoother=None):
        return model.solve_step_device(steps[k % len(steps)], du.data_ptr(), tol=tols[prec] if tol is None else tol,
                                       maxit=2000, precision=prec, precond=hs.PC_MG, sweeps=args.sweeps,
                                       smoother=SM if smoother is None else smoother)

    for w in range(args.warmup):
        rec = solve(w, hs.PREC_MIXED)
        log(f"warmup {w}: iters {rec['iters']} {rec['ms_total']:.1f} ms")
    # untimed calibration: the FP64 oracle of every step (tol 1e-12) and, per
    # precision, the loosest tolerance whose warpage is within 1e-6 of it
    w_ref, cal = [], {}
    for k in range(len(steps)):
        w_ref.append(warpage_oracle(model, steps[k], du, mesh, bga, post)[1])
    if args.tol_from_warpage:
        for prec in (hs.PREC_MIXED, hs.PREC_FP64):
            tols[prec], cal[prec] = calibrate_tol(lambda t: solve(0, prec, t), du, mesh, bga, post, w_ref[0])
            log(f"precision {prec}: tolerance {tols[prec]:g} for warpage error <= 1e-6 ({cal[prec]})")
    golden_w = None
    if args.config == "cfg2":
        with open(os.path.join(ROOT, "tests", "golden", "cfg2_golden.json")) as f:
            golden_w = json.load(f)["warpage"]["signed"]  # the reference's JS solve, 25 C, tol 1e-12
    recs = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        recs.append(solve(k, hs.PREC_MIXED))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # warpage error of the timed mixed solves (recomputed untimed, same tolerance, deterministic)
    werr_mixed = []
    for k in range(min(args.steps, len(steps))):
        solve(k, hs.PREC_MIXED)
        w = post.warpage_jeita(mesh, du.cpu().numpy(), bga)["signed"]
        werr_mixed.append(abs(w - w_ref[k]) / abs(w_ref[k]))
    # throughput pass: the same steps with K solves in flight on K instances of
    # the model (each its own stream, graphs and host thread), as the reference
    # spreads the six temperatures over its workers
    K = max(1, min(args.concurrent, args.steps))
    extra = [hs.Model(mesh, variant="hip") for _ in range(K - 1)]
    models_k = [model] + extra
    dus_k = [du] + [torch.empty(mesh.ndof, dtype=torch.float64, device="cuda") for _ in extra]

    def solve_on(mi, k):
        return models_k[mi].solve_step_device(steps[k % len(steps)], dus_k[mi].data_ptr(), tol=tols[hs.PREC_MIXED],
                                              maxit=2000, precision=hs.PREC_MIXED, precond=hs.PC_MG,
                                              sweeps=args.sweeps, smoother=SM)

    for mi in range(1, K):
        solve_on(mi, 0)  # warm-up: graph capture of each instance
    torch.cuda.synchronize()
    from concurrent.futures import ThreadPoolExecutor
    tc0 = time.perf_counter()
    with ThreadPoolExecutor(K) as ex:
        per = list(ex.map(lambda mi: [(k, solve_on(mi, k)) for k in range(mi, args.steps, K)], range(K)))
    torch.cuda.synchronize()
    tc1 = time.perf_counter()
    recs_conc = [r for _, r in sorted((kr for lst in per for kr in lst), key=lambda kr: kr[0])]
    del extra, models_k, dus_k
    # second timed pass with HIP events around the kernel classes (launched eagerly,
    # so the events bracket each kernel; the kernels themselves are the same)
    model.set_kernel_timing(True)
    torch.cuda.synchronize()
    for k in range(min(args.steps, 2)):
        solve(k, hs.PREC_MIXED)
    torch.cuda.synchronize()
    ktimes = model.kernel_times()
    model.set_kernel_timing(False)
    # the other smoother on the same steps, one solve at a time (its own captured graphs)
    OTHER = hs.SMOOTH_GS if SM == hs.SMOOTH_COARSE_JACOBI else hs.SMOOTH_COARSE_JACOBI
    solve(0, hs.PREC_MIXED, smoother=OTHER)
    torch.cuda.synchronize()
    to0 = time.perf_counter()
    recs_other = [solve(k, hs.PREC_MIXED, smoother=OTHER) for k in range(args.steps)]
    torch.cuda.synchronize()
    to1 = time.perf_counter()
    # a plain 256 MiB device copy: the reference point for scaling hardware byte counters
    src = torch.empty(32 * 1024 * 1024, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    elapsed = t1 - t0
    dof_iters = sum(mesh.ndof * r["iters"] for r in recs)
    elapsed_conc = tc1 - tc0
    value = sum(mesh.ndof * r["iters"] for r in recs_conc) / elapsed_conc
    log(f"{K} solves in flight: {1e3 * elapsed_conc / max(args.steps, 1):.1f} ms per solve; one at a time "
        f"{1e3 * elapsed / max(args.steps, 1):.1f} ms ({args.smoother}); other smoother one at a time "
        f"{1e3 * (to1 - to0) / max(args.steps, 1):.1f} ms, iters {recs_other[0]['iters']}")
    for k, r in enumerate(recs):
        log(f"step {k} T={temps[k % len(temps)]}: iters {r['iters']} relres {r['relres']:.2e} "
            f"asm {r['ms_assemble']:.1f} setup {r['ms_setup']:.1f} pcg {r['ms_pcg']:.1f} ms")
    log("kernel classes (ms total, launches):", json.dumps(ktimes))

    # dominant timed kernel class and its roofline
    timed = {k: v for k, v in ktimes.items() if v[1] > 0}
    dom = max(timed, key=lambda k: timed[k][0]) if timed else None
    roof = None
    for cls in ([dom] if dom else []) + ["fine_sweep", "apply"]:
        if cls in timed:
            roof = roofline_for(cls, timed[cls][0], timed[cls][1], mesh, "fp32")
            if roof:
                break
    if roof and dom and roof["kernel"] != dom:
        roof["note"] = f"largest timed class is {dom}, whose work is not a single operator launch"

    out = {
        "metric": "solve s & DOF*iter/s at ~4M DOF: mixed vs FP64 PCG-MG at 1e-6 warpage err",
        "value": value, "unit": "DOF*iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed_conc / max(args.steps, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": f"{spec['name']}: one cold thermoelastic step solve per step (150 C stress-free "
                               f"to T in {temps}), PCG to {tols[hs.PREC_MIXED]:g} scaled relative residual (the "
                               f"loosest tolerance with JEITA warpage within 1e-6 of the FP64 oracle), {K} solves "
                               f"in flight on {K} model instances",
                   "ndof": mesh.ndof, "elements": mesh.ne, "precision": "FP32 multigrid in FP64 flexible PCG"},
        "solves_in_flight": K,
        "tol_mixed": tols[hs.PREC_MIXED], "tol_fp64": tols[hs.PREC_FP64],
        "tol_calibration": {("mixed" if p == hs.PREC_MIXED else "fp64"): c for p, c in cal.items()},
        "warpage_rel_err_mixed_vs_fp64_oracle": werr_mixed,
        "warpage_fp64_oracle_vs_reference": (abs(w_ref[0] - golden_w) / abs(golden_w)) if golden_w else None,
        "value_one_at_a_time": dof_iters / elapsed,
        "solve_s_mixed": elapsed / max(args.steps, 1),
        "iters_mixed": [r["iters"] for r in recs],
        "iters_mixed_in_flight": [r["iters"] for r in recs_conc],
        "setup_ms_mixed": float(np.mean([r["ms_assemble"] + r["ms_setup"] for r in recs])),
        "smoother": args.smoother,
        "other_smoother": {"smoother": "gs" if OTHER == hs.SMOOTH_GS else "coarse-jacobi",
                           "solve_s_mixed": (to1 - to0) / max(args.steps, 1),
                           "iters_mixed": [r["iters"] for r in recs_other],
                           "value_one_at_a_time": sum(mesh.ndof * r["iters"] for r in recs_other) / (to1 - to0),
                           "note": "gs is the reference's colour Gauss-Seidel on every level; coarse-jacobi smooths "
                                   "the shell levels from the third down with four damped column-Jacobi sweeps, "
                                   "one launch each"},
        "roofline": roof,
        "composite_roofline": composite_roofline(ktimes, min(args.steps, 2), int(np.mean([r["iters"] for r in recs])),
                                                 float(np.mean([r["ms_pcg"] for r in recs])), mesh),
    }

    if not args.no_operator_compare:
        out["matrix_core_operator"] = operator_compare(hs, mesh, steps[0], tols[hs.PREC_MIXED])
        log("matrix-core operator:", json.dumps(out["matrix_core_operator"]))

    if not args.no_fp64:
        # FP64 comparator on the same steps at its own 1e-6-warpage tolerance
        t_fp64, it64, werr = [], [], []
        solve(0, hs.PREC_FP64)  # graph capture of the FP64 iteration
        for k in range(args.steps):
            torch.cuda.synchronize()
            ta = time.perf_counter()
            rec64 = solve(k, hs.PREC_FP64)
            torch.cuda.synchronize()
            t_fp64.append(time.perf_counter() - ta)
            it64.append(rec64["iters"])
            w_d = post.warpage_jeita(mesh, du.cpu().numpy(), bga)["signed"]
            werr.append(abs(w_d - w_ref[k % len(steps)]) / abs(w_ref[k % len(steps)]))
        out["solve_s_fp64"] = float(np.mean(t_fp64))
        out["iters_fp64"] = it64
        out["speedup_mixed_vs_fp64"] = out["solve_s_fp64"] / out["solve_s_mixed"]
        out["warpage_rel_err_fp64_vs_fp64_oracle"] = werr
        out["warpage_rel_err_mixed_vs_fp64_max"] = float(max(werr_mixed))
        log(f"FP64 comparator: {out['solve_s_fp64']:.3f} s/step at tol {tols[hs.PREC_FP64]:g}, mixed "
            f"{out['solve_s_mixed']:.3f} s/step at tol {tols[hs.PREC_MIXED]:g}")

    if not args.no_adjoint and args.config == "cfg2":
        out["adjoint"] = adjoint_timing(torch, hs, cases, meshspec, args.tol)
        a = out["adjoint"]
        log(f"adjoint ({a['params']} parameters): forward {a['forward_s']:.3f} s + backward {a['backward_s']:.3f} s "
            f"vs {a['fd_solves']} central-difference solves {a['fd_equivalent_s']:.2f} s")

    # Anand material-point update at the bump-array size (20 x 20 bumps x 36 solder hexes x 8 points)
    out["anand"] = anand_throughput(torch, hs)
    log(f"Anand: {out['anand']['gp_updates_per_s']:.3e} point updates/s ({out['anand']['n_points']} points, "
        f"{out['anand']['ms_per_update']:.3f} ms per batch)")

    if not args.no_path and args.config == "cfg2":
        # at the tolerance the reference's Analysis gives its solver (PCG_TOL,
        # src/core/00-constants.js:60, src/core/11-analysis.js:110)
        out["path"] = path_timing(torch, hs, model, spec, 1e-8)
        log(f"thermal path: {out['path']['steps']} steps in {out['path']['s']:.2f} s "
            f"({out['path']['iters_mean']:.1f} iterations per step, warm start); over 3 instances "
            f"{out['path']['blocks']['s']:.2f} s (fields differ by {out['path']['blocks']['fields_rel_diff_vs_one']:.1e})")

    if not args.no_submodel:
        out["submodel"] = submodel_timing()
        b = out["submodel"]["batch"]
        log(f"submodel: {out['submodel']['steps']} steps in {out['submodel']['s']:.1f} s "
            f"(dW vs reference {out['submodel']['dW_rel_err_vs_reference']:.1e}); "
            f"{b['submodels']} at once in {b['s']:.1f} s ({b['vs_one_at_a_time']:.1f}x one at a time)")

    if not args.no_cfg3:
        out["cfg3"] = cfg3_timing(hs, cases, meshspec)
        c3 = out["cfg3"]
        log(f"cfg3 bump array ({c3['ndof']} DOF, {c3['points']} Anand points): {c3['steps']} steps in "
            f"{c3['s']:.1f} s ({c3['ms_per_step']:.0f} ms per step, {c3['newton_per_step']:.2f} Newton per step, "
            f"{c3['pcg_per_newton']:.1f} PCG per Newton)")

    if not args.no_cfg5:
        out["cfg5"] = cfg5_timing(hs, cases, nsub=args.cfg5_sample)
        c5 = out["cfg5"]
        log(f"cfg5 batched submodels: {c5['sample']} of {c5['submodels_total']} ({c5['dof_sample']} DOF) in "
            f"{c5['wall_s']:.1f} s with {c5['workers']} in flight; p50 {c5['p50_s']:.1f} s p99 {c5['p99_s']:.1f} s; "
            f"{c5['dof_iters_per_s']:.3e} DOF*iter/s; all 256 extrapolated {c5['all_256_s_extrapolated']:.0f} s")

    from fcbga_amd import screening as scr
    out["screening"] = screening_timing(hs, scr)
    log(f"screening: {out['screening']['joints']} joints x {out['screening']['steps']} steps in "
        f"{out['screening']['ms']:.1f} ms")

    if not args.no_cpu:
        log("CPU oracle sample ...")
        base = cpu_baseline(spec, mesh)
        base.update(host_facts())
        # the six-temperature sweep as the reference's worker farm runs it: one
        # process per temperature, min(6, floor(nproc / 1.5)) at once, each the
        # oracle's assembly + set-up + 5 PCG iterations
        import multiprocessing as mp
        nw = max(1, min(len(temps), int(base["nproc"] // 1.5)))
        t0 = time.perf_counter()
        with mp.get_context("spawn").Pool(nw) as pool:
            res = pool.map(_cpu_sample_child, [(args.config, T, 5) for T in temps])
        dts = time.perf_counter() - t0
        base["sweep"] = {"processes": nw, "temperatures": len(temps), "seconds": dts,
                         "value": sum(n * it for n, it, _ in res) / dts, "unit": "DOF*iter/s",
                         "sample": "per temperature: oracle assembly + multigrid set-up + 5 PCG iterations, "
                                   "FP64, 1 thread per process"}
        # the same screening loop in the oracle (C, 1 thread) on every 97th joint
        from oracle import oracle as orc
        steps, table, i25 = screening_case(scr)
        sub = table[::97]
        t0 = time.perf_counter()
        orc.screen_joints([5.87e6, 7460.0, 2.0, 0.0942, 58.3, 0.015, 9350.0, 1.5, 45.9, 0.4], steps, sub,
                          table.shape[1], i25)
        base["screening_s_full_estimate"] = (time.perf_counter() - t0) * table.shape[0] / len(sub)
        base["screening_sample"] = f"{len(sub)} of {table.shape[0]} joints, 1 thread"
        out["cpu_baseline"] = base
        log(f"CPU: {base['value']:.3e} DOF*iter/s over {base['seconds']:.1f} s; screening "
            f"{base['screening_s_full_estimate']:.1f} s for all joints")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
