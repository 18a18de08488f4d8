"""Thermal path: the plan of steps and evaluation temperatures, the step
solves and their summation into displacement fields (temperatureSamples,
thermalPathPoints, Analysis.evaluationTemperatures / buildPath / run /
finishPath: src/core/11-analysis.js:29-261).

The plan is host logic (it depends only on the configuration's process
temperatures and the Tg values of its materials); the steps run on the device
one after the other with a warm-start basis shared along the path, and the
fields are summed in device memory."""
from __future__ import annotations

import math

import numpy as np

INC_DT_MAX = 30.0        # src/core/00-constants.js:29
INC_DT_TG = 10.0         # :30
INC_TG_BAND = 15.0       # :31
CYCLE_DT_STEP = 10.0     # :106
CYCLE_DT_FINE = 5.0      # :107
CYCLE_TG_BAND = 15.0     # :108
REFLOW_DT = 15.0         # :109
REFLOW_DT_FINE = 10.0    # :110
STAGE_DDI, STAGE_WCUF, STAGE_MOLD = -3, -2, -1   # :27


def _js_round(x):
    """Math.round: halves go up."""
    return math.floor(x + 0.5)


def _num(v):
    return int(v) if float(v).is_integer() else float(v)


def temperature_samples(Tmin, Tmax, tgs, extras=()):
    """Every CYCLE_DT_STEP, refined to CYCLE_DT_FINE within +-CYCLE_TG_BAND of a
    Tg, plus extras, within [Tmin, Tmax], rounded to 1e-3 (temperatureSamples,
    src/core/11-analysis.js:29-40)."""
    out = {}

    def add(T):
        if Tmin - 1e-9 <= T <= Tmax + 1e-9:
            out.setdefault(_js_round(T * 1000) / 1000, None)

    add(Tmin)
    add(Tmax)
    T = math.ceil(Tmin / CYCLE_DT_STEP) * CYCLE_DT_STEP
    while T <= Tmax:
        add(T)
        T += CYCLE_DT_STEP
    for tg in tgs:
        T = math.ceil((tg - CYCLE_TG_BAND) / CYCLE_DT_FINE) * CYCLE_DT_FINE
        while T <= tg + CYCLE_TG_BAND:
            add(T)
            T += CYCLE_DT_FINE
    for T in extras or ():
        add(T)
    return sorted(out)


def thermal_path_points(Ta, Tb, tgs, stops=()):
    """Points of the path from Ta to Tb: steps of at most INC_DT_MAX, at most
    INC_DT_TG where a step meets the INC_TG_BAND window of a Tg, and every stop
    strictly between (thermalPathPoints, src/core/11-analysis.js:51-68)."""
    d = Tb - Ta
    direction = (d > 0) - (d < 0)
    pts = [Ta]
    if direction == 0:
        return pts
    inner = sorted((T for T in (stops or ()) if (T - Ta) * direction > 1e-9 and (Tb - T) * direction > 1e-9),
                   key=lambda T: direction * T)
    inner.append(Tb)
    T = Ta
    for S in inner:
        while (S - T) * direction > 1e-9:
            lo, hi = min(T, T + direction * INC_DT_MAX), max(T, T + direction * INC_DT_MAX)
            h = INC_DT_TG if any(tg + INC_TG_BAND > lo and tg - INC_TG_BAND < hi for tg in tgs) else INC_DT_MAX
            nxt = _js_round((T + direction * h) * 1e6) / 1e6
            T = nxt if (S - nxt) * direction > 1e-9 else S
            pts.append(T)
    return pts


def evaluation_temperatures(cfg, tgs, solidi, evaluations="full"):
    """{all, cycling, reflow} (Analysis.evaluationTemperatures, src/core/11-analysis.js:135-148)."""
    if evaluations == "preview":
        return {"all": [25], "cycling": [], "reflow": [25]}
    cyc = temperature_samples(cfg["cycling"]["Tmin"], cfg["cycling"]["Tmax"], tgs, [25])
    peak = cfg["reflow"]["peak"]
    cand = {25: None, peak: None}
    for s in solidi:
        cand.setdefault(s, None)
        cand.setdefault(s + 1, None)
    for T in cfg["reflow"]["reportTemps"]:
        cand.setdefault(T, None)
    T = math.ceil(cfg["cycling"]["Tmax"] / REFLOW_DT) * REFLOW_DT
    while T <= peak:
        cand.setdefault(T, None)
        T += REFLOW_DT
    for tg in tgs:
        T = math.ceil((tg - CYCLE_TG_BAND) / REFLOW_DT_FINE) * REFLOW_DT_FINE
        while T <= tg + CYCLE_TG_BAND:
            cand.setdefault(T, None)
            T += REFLOW_DT_FINE
    ref = sorted(T for T in cand if 25 <= T <= peak)
    return {"all": sorted(set(cyc) | set(ref)), "cycling": cyc, "reflow": ref}


def build_path(cfg, tgs, stage_tgs, solidi, evaluations="full"):
    """Steps [{Ta, Tb, stage, branch}] and the path's step lists (chain,
    chainJoin, cpi, up, down, ddiJoin, ddi) with the evaluation temperatures
    (Analysis.buildPath, src/core/11-analysis.js:151-187).  stage_tgs: Tg values
    of the materials born by each stage."""
    p = cfg["process"]
    fin = 2 if cfg["packageType"] == "bare" else 3
    has_modules = bool(cfg.get("modules"))
    temps = evaluation_temperatures(cfg, tgs, solidi, evaluations)
    steps = []

    def seg(Ta, Tb, stage, branch, stops=()):
        pts = thermal_path_points(Ta, Tb, stage_tgs[stage], list(stops) + list(solidi))
        ids = []
        for k in range(1, len(pts)):
            steps.append({"Ta": _num(pts[k - 1]), "Tb": _num(pts[k]), "stage": stage, "branch": branch})
            ids.append(len(steps) - 1)
        return ids

    sm, sm_join, ddi = [], [], []
    if has_modules:
        mp = cfg["stack"]["process"]
        wcuf = cfg["stack"]["fill"] == "wcuf"
        t_fill = mp["Twcuf"] if wcuf else mp["Tmold"]
        sm_join = seg(mp["Tddi"], t_fill, STAGE_DDI, "chain")
        if evaluations == "full":
            ddi = seg(t_fill, 25, STAGE_DDI, "ddi")
        mid = seg(mp["Twcuf"], mp["Tmold"], STAGE_WCUF, "chain") if wcuf else []
        sm = sm_join + mid + seg(mp["Tmold"], p["Tsub"], STAGE_MOLD, "chain")
    s0 = seg(p["Tsub"], p["Tjoin"], 0, "chain") if abs(p["Tsub"] - p["Tjoin"]) > 1e-9 else []
    s1 = seg(p["Tjoin"], p["Tuf"], 1, "chain")
    cpi = seg(p["Tuf"], 25, 1, "cpi") if evaluations == "full" else []
    if fin == 3:
        rest = seg(p["Tuf"], p["Tattach"], 2, "chain")
        rest = rest + seg(p["Tattach"], 25, 3, "chain")
    else:
        rest = seg(p["Tuf"], 25, 2, "chain")
    ev = temps["all"]
    down = seg(25, min([25] + ev), fin, "down", ev)
    up = seg(25, max([25] + ev), fin, "up", ev)
    path = {"chain": sm + s0 + s1 + rest, "chainJoin": sm + s0 + s1, "cpi": cpi, "up": up, "down": down,
            "hasCPI": evaluations == "full", "ddiJoin": sm_join, "ddi": ddi, "hasDDI": len(ddi) > 0}
    return steps, path, temps


def finish_path(steps, path, du, temps_all):
    """Fields at every evaluation temperature and the chip-join / chiplet-join
    states from host step increments du[k] (Analysis.finishPath,
    src/core/11-analysis.js:229-261); run_path does the same on the device."""
    zeros = np.zeros
    copy = np.copy
    n = len(du[0])
    want = set(temps_all)
    u = zeros(n)
    for k in path["chain"]:
        u += du[k]
    fields = {25: copy(u)}
    for br in (path["up"], path["down"]):
        v = copy(u)
        for k in br:
            v += du[k]
            if steps[k]["Tb"] in want:
                fields[steps[k]["Tb"]] = copy(v)
    u_cpi = u_ddi = None
    if path["hasCPI"]:
        u_cpi = zeros(n)
        for k in path["chainJoin"] + path["cpi"]:
            u_cpi += du[k]
    if path["hasDDI"]:
        u_ddi = zeros(n)
        for k in path["ddiJoin"] + path["ddi"]:
            u_ddi += du[k]
    return fields, u_cpi, u_ddi


def run_path(model, steps, path, temps_all, step_of, tol=1e-8, precision=None, warm=20):
    """Solve every step in plan order on the device with one warm-start basis
    along the path (Analysis.run, src/core/11-analysis.js:212-216); the
    increments stay in device memory and every field is summed there from the
    increments that lead to it, added in the reference's order
    (finishPath).  step_of(step) gives the hip_solver.Step of a plan step.
    Returns (fields, u_cpi, u_ddi, records)."""
    from . import hip_solver as hs
    precision = hs.PREC_MIXED if precision is None else precision
    recs = []
    model.warm_reset()
    model.path_reset()
    for k, s in enumerate(steps):
        rec = model.path_solve(k, step_of(s), tol=tol, precision=precision, warm=warm)
        if not rec["converged"]:
            raise hs.SolverError(f"Linear solver did not converge (step {k}: {s['Ta']} -> {s['Tb']} C): "
                                 f"relative residual {rec['relres']:.2e}")
        recs.append(rec)
    fields, u_cpi, u_ddi = finish_path_device(model, steps, path, temps_all)
    return fields, u_cpi, u_ddi, recs


def finish_path_device(model, steps, path, temps_all):
    """finishPath (src/core/11-analysis.js:236-255) over the increments kept on
    the device: every field is the sum of the increments that lead to it, added
    in the reference's order.  Returns (fields, u_cpi, u_ddi)."""
    want = set(temps_all)
    fields = {25: model.path_sum(path["chain"])}
    for br in (path["up"], path["down"]):
        for q, k in enumerate(br):
            if steps[k]["Tb"] in want:
                fields[steps[k]["Tb"]] = model.path_sum(path["chain"] + br[:q + 1])
    u_cpi = model.path_sum(path["chainJoin"] + path["cpi"]) if path["hasCPI"] else None
    u_ddi = model.path_sum(path["ddiJoin"] + path["ddi"]) if path["hasDDI"] else None
    return fields, u_cpi, u_ddi


SWEEP_BLOCKS_PER_WORKER = 3  # src/core/00-constants.js:65
SWEEP_BLOCK_MIN = 2          # :66


def path_blocks(n, workers):
    """Blocks of consecutive steps of a plan of n steps for the given number
    of workers (runSweep, src/worker.js:114-117)."""
    size = max(SWEEP_BLOCK_MIN, math.ceil(n / (SWEEP_BLOCKS_PER_WORKER * workers)))
    return [list(range(k, min(n, k + size))) for k in range(0, n, size)]


def run_path_blocks(models, steps, path, temps_all, step_of, tol=1e-8, precision=None, warm=20):
    """The plan's steps over several model instances of the same mesh, the
    reference's worker farm (runSweep / evalSteps, src/worker.js:76-133):
    blocks of consecutive steps are taken from one queue by one host thread
    per instance, each instance solving its blocks in order with its own
    warm-start basis (its own stream, so the solves overlap on the device);
    the increments then meet in the first instance (device to device) and
    the fields are summed there in finishPath's order.  With one instance
    this is run_path.  Returns (fields, u_cpi, u_ddi, records)."""
    import threading
    from concurrent.futures import ThreadPoolExecutor

    from . import hip_solver as hs
    precision = hs.PREC_MIXED if precision is None else precision
    blocks = path_blocks(len(steps), len(models))
    lock = threading.Lock()
    nxt = [0]
    recs = [None] * len(steps)
    owner = [0] * len(steps)

    def worker(mi):
        m = models[mi]
        m.warm_reset()
        m.path_reset()
        while True:
            with lock:
                if nxt[0] >= len(blocks):
                    return
                b = blocks[nxt[0]]
                nxt[0] += 1
            for k in b:
                rec = m.path_solve(k, step_of(steps[k]), tol=tol, precision=precision, warm=warm)
                if not rec["converged"]:
                    raise hs.SolverError(f"Linear solver did not converge (step {k}: {steps[k]['Ta']} -> "
                                         f"{steps[k]['Tb']} C): relative residual {rec['relres']:.2e}")
                recs[k], owner[k] = rec, mi

    with ThreadPoolExecutor(len(models)) as ex:
        list(ex.map(worker, range(len(models))))
    for k, mi in enumerate(owner):
        if mi:
            models[0].path_copy_from(models[mi], k)
    fields, u_cpi, u_ddi = finish_path_device(models[0], steps, path, temps_all)
    return fields, u_cpi, u_ddi, recs
