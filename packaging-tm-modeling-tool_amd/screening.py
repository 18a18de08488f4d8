"""Bump screening through the temperature cycle (screenJoints, cycleHistory,
screeningSites: src/core/13-fatigue.js:275-455).

Host side of the batched device kernel: the cycle history, the sites kept for
screening, the per-step material data (modulus and thermal strain on the
reference's 1 degC property grid, interpolation weights into the bump-layer
strain table) and the ranking.  The per-site loop over the cycle steps
(AnandPoint.stepColumn / stepSeriesColumn) runs on the device, one thread per
joint, through fcbga_screen_joints."""
from __future__ import annotations

import math

import numpy as np

TGRID_MIN, TGRID_MAX = -80, 300      # 1 degC property grid (src/core/00-constants.js:12-13)
TGRID_N = TGRID_MAX - TGRID_MIN + 1
T_STRAIN_ORIGIN = 25.0               # origin of the cumulative thermal strain F(T)
RAMP_MAX_DT = 5.0                    # degC per ramp step (src/core/00-constants.js:111)
DWELL_MIN_STEPS = 10                 # steps per dwell (:112)
SCREEN_MAX_BUMPS = 20000             # sites screened per die at most (:102)
SCREEN_MAX_DDI = 5000                # DDI microbumps per chiplet at most (:103)
SCREEN_KEEP_ROWS = 3                 # outer rows always kept (:104)
SCREEN_CORNER_BLOCK = 10             # corner blocks always kept (:105)


def find_interval(a, x):
    """Index i with a[i] <= x < a[i+1], clamped to [0, n-2] (findInterval, src/core/01-util.js:12-23)."""
    n = len(a)
    if x <= a[0]:
        return 0
    if x >= a[n - 1]:
        return n - 2
    lo, hi = 0, n - 1
    while hi - lo > 1:
        mid = (lo + hi) >> 1
        if a[mid] <= x:
            lo = mid
        else:
            hi = mid
    return lo


def interp_table(xs, ys, x, extrap="linear"):
    """Piecewise-linear table (interpTable, src/core/01-util.js:25-33)."""
    n = len(xs)
    if n == 1:
        return ys[0]
    if x <= xs[0] and extrap != "linear":
        return ys[0]
    if x >= xs[n - 1] and extrap != "linear":
        return ys[n - 1]
    i = find_interval(xs, x)
    t = (x - xs[i]) / (xs[i + 1] - xs[i])
    return ys[i] + t * (ys[i + 1] - ys[i])


class MaterialTables:
    """E(T) (MPa) and the thermal strain F(T) = alpha (T - 25) of an isotropic
    material with a tabulated modulus and a constant CTE on the 1 degC grid
    (isoTables / gridAt, src/core/03-materials.js:347-355, 424-433)."""

    def __init__(self, E_T, E_v, cte_ppm, nu, scale=1000.0):
        """E_v in GPa times scale MPa (the reference's tables are in GPa)."""
        Ts = [TGRID_MIN + i for i in range(TGRID_N)]
        self.E = np.array([interp_table(E_T, E_v, float(T)) * scale for T in Ts])
        self.F = np.array([cte_ppm * 1e-6 * (T - T_STRAIN_ORIGIN) for T in Ts])
        self.nu = nu

    @classmethod
    def from_grids(cls, E, F, nu):
        """Tables given on the 1 degC grid (E in MPa, F the thermal strain), as a
        material's own tables hold them (materialTables, src/core/03-materials.js)."""
        t = cls.__new__(cls)
        t.E, t.F, t.nu = np.asarray(E, np.float64), np.asarray(F, np.float64), nu
        if len(t.E) != TGRID_N or len(t.F) != TGRID_N:
            raise ValueError("property tables must have one value per degree C of the grid")
        return t

    @staticmethod
    def grid_at(arr, T):
        x = T - TGRID_MIN
        if x <= 0:
            return float(arr[0])
        if x >= TGRID_N - 1:
            return float(arr[TGRID_N - 1])
        i = int(math.floor(x))
        t = x - i
        a, b = arr[i], arr[i + 1]
        return float(a + t * (b - a))

    def E_at(self, T):
        return self.grid_at(self.E, T)

    def F_at(self, T):
        return self.grid_at(self.F, T)


def cycle_history(cyc):
    """Time/temperature history of the cycling test (cycleHistory,
    src/core/13-fatigue.js:275-293): from 25 C, per cycle a ramp to Tmax, a hot
    dwell, a ramp to Tmin and a cold dwell; ramps in steps of at most 5 C at
    cyc['ramp'] C/min, dwells of cyc['dwell'] min in 10 steps."""
    Tmin, Tmax, rate, dwell = cyc["Tmin"], cyc["Tmax"], cyc["ramp"] / 60, cyc["dwell"] * 60
    t, T, cycle, phase = [0.0], [25.0], [0], ["start"]
    time = 0.0

    def ramp(frm, to, c):
        nonlocal time
        n = max(1, math.ceil(abs(to - frm) / RAMP_MAX_DT))
        dur = abs(to - frm) / rate
        for k in range(1, n + 1):
            time += dur / n
            t.append(time); T.append(frm + (to - frm) * k / n); cycle.append(c)
            phase.append("ramp up" if to > frm else "ramp down")

    def hold(Tv, c, label):
        nonlocal time
        for _ in range(DWELL_MIN_STEPS):
            time += dwell / DWELL_MIN_STEPS
            t.append(time); T.append(Tv); cycle.append(c); phase.append(label)

    for c in range(1, cyc["cycles"] + 1):
        ramp(T[-1], Tmax, c); hold(Tmax, c, "dwell hot")
        ramp(Tmax, Tmin, c); hold(Tmin, c, "dwell cold")
    return {"t": np.array(t), "T": np.array(T), "cycle": np.array(cycle, np.int32), "phase": phase,
            "cycles": cyc["cycles"]}


def bump_sites(d):
    """All sites of a bump field with their row index (dieBumpSites, src/core/05-geometry.js:188-196)."""
    x, y, ix, iy, row = [], [], [], [], []
    for i in range(d["nbx"]):
        for j in range(d["nby"]):
            r = min(i, j, d["nbx"] - 1 - i, d["nby"] - 1 - j)
            if d.get("peripheral") and r >= d["rows"]:
                continue
            x.append(d["bx0"] + i * d["px"]); y.append(d["by0"] + j * d["py"])
            ix.append(i); iy.append(j); row.append(r)
    return {"x": np.array(x), "y": np.array(y), "ix": np.array(ix, np.int32), "iy": np.array(iy, np.int32),
            "row": np.array(row, np.int32)}


def screening_sites(d, cap=None):
    """Sites kept for screening (screeningSites, src/core/13-fatigue.js:296-312):
    all of them up to the cap, else the outer rows and corner blocks plus an even
    stride through the rest.  Returns (sites, kept indices or None)."""
    allv = bump_sites(d)
    n = len(allv["x"])
    cap = cap or SCREEN_MAX_BUMPS
    if n <= cap:
        return allv, None
    nbx, nby = d["nbx"], d["nby"]

    def must(q):
        return bool(allv["row"][q] < SCREEN_KEEP_ROWS or
                    ((allv["ix"][q] < SCREEN_CORNER_BLOCK or allv["ix"][q] >= nbx - SCREEN_CORNER_BLOCK) and
                     (allv["iy"][q] < SCREEN_CORNER_BLOCK or allv["iy"][q] >= nby - SCREEN_CORNER_BLOCK)))

    keep, rest = [], []
    for q in range(n):
        (keep if must(q) else rest).append(q)
    budget = max(0, cap - len(keep))
    stride = len(rest) / max(budget, 1)
    k = 0
    while k < budget and k * stride < len(rest):
        keep.append(rest[int(math.floor(k * stride))])
        k += 1
    keep.sort()
    idx = np.array(keep, np.int32)
    return {key: v[idx] for key, v in allv.items()}, idx


def step_data(hist, temps, solder: MaterialTables, pillar: MaterialTables | None = None, fs=1.0):
    """Per-step inputs of the device loop (the precomputation of screenJoints,
    src/core/13-fatigue.js:394-415): interval and weight of each step's
    temperature in the strain table, dt, the solder's E(T) and thermal strain
    relative to 25 C, whether the step continues the phase of the one before,
    the cycle index; with a pillar its compliances (1/E, 1/G, 1/G) and thermal
    strain."""
    T, t, cyc, ph = hist["T"], hist["t"], hist["cycle"], hist["phase"]
    n = len(T)
    temps = [float(v) for v in temps]
    wi = np.zeros(n, np.int32)
    wt = np.zeros(n)
    Fth = np.zeros(n)
    E = np.zeros(n)
    dt = np.zeros(n)
    same = np.zeros(n, np.uint8)
    F25 = solder.F_at(25.0)
    for k in range(n):
        i = find_interval(temps, T[k])
        wi[k] = i
        wt[k] = min(max((T[k] - temps[i]) / (temps[i + 1] - temps[i]), 0.0), 1.0)
        Fth[k] = solder.F_at(T[k]) - F25
        E[k] = solder.E_at(T[k])
        dt[k] = t[k] - t[k - 1] if k > 0 else 0.0
        same[k] = 1 if k > 0 and ph[k] == ph[k - 1] else 0
    out = {"nsteps": n, "cycles": hist["cycles"], "T": np.ascontiguousarray(T, np.float64), "dt": dt, "E": E,
           "Fth": Fth, "wi": wi, "wt": wt, "same": same, "cycle": np.ascontiguousarray(cyc, np.int32),
           "Cp": None, "Fp": None, "fs": 1.0}
    if pillar is not None:
        Cp = np.zeros(3 * n)
        Fp = np.zeros(n)
        Fp25 = pillar.F_at(25.0)
        for k in range(n):
            Ep = pillar.E_at(T[k])
            Gp = Ep / (2 * (1 + pillar.nu))
            Cp[3 * k] = 1 / Ep
            Cp[3 * k + 1] = Cp[3 * k + 2] = 1 / Gp
            Fp[k] = pillar.F_at(T[k]) - Fp25
        out.update(Cp=Cp, Fp=Fp, fs=float(fs))
    return out


def rank(dW, n_top=20):
    """Indices of the finite dW, largest first (the top list of screenJoints)."""
    dW = np.asarray(dW)
    q = np.nonzero(np.isfinite(dW))[0]
    order = q[np.argsort(-dW[q], kind="stable")]
    return order[:n_top]


def site_strain_table(model, chain, up, down, temps, sites, zmid, part):
    """Bump-layer total strain (engineering) at every site for every cycling
    temperature (siteStrainTable + Analysis.qAlong, src/core/13-fatigue.js:
    318-345, src/core/11-analysis.js:298-316): the element DOFs q (24 nodal +
    9 incompatible) of each site's bump element are accumulated on the device
    (q only) along `chain`, the (Step, du) pairs from the stress-free state to
    25 C, then along the branches `up` and `down` from 25 C towards the highest
    and the lowest of `temps`, recorded at every step whose end temperature is
    in `temps`; the strain is the box strain of q at the site's natural
    coordinates in the layer at height zmid.  Returns (nsites, ntemps, 6), NaN
    for sites without an element of `part` and temperatures not reached."""
    from .post import _interval, element_in_cell
    mesh = model.mesh
    xs, ys, zs = (np.asarray(a, np.float64) for a in (mesh.xs, mesh.ys, mesh.zs))
    sx, sy = np.asarray(sites["x"]), np.asarray(sites["y"])
    n, nt = len(sx), len(temps)
    out = np.full((n, nt, 6), np.nan)
    kb = int(_interval(zs, zmid))
    site_e = np.full(n, -1, np.int64)
    xez = np.zeros((n, 3))
    for q in range(n):
        i, j = int(_interval(xs, sx[q])), int(_interval(ys, sy[q]))
        e = element_in_cell(mesh, i, j, kb, part)
        if e < 0:
            continue
        site_e[q] = e
        ei, ej, ek = (int(v) for v in mesh.e_ijk[e])
        # naturalCoords (src/core/12-post.js:4-11)
        xez[q] = [np.clip(2 * (sx[q] - xs[ei]) / (xs[ei + 1] - xs[ei]) - 1, -1, 1),
                  np.clip(2 * (sy[q] - ys[ej]) / (ys[ej + 1] - ys[ej]) - 1, -1, 1),
                  np.clip(2 * (zmid - zs[ek]) / (zs[ek + 1] - zs[ek]) - 1, -1, 1)]
    has = np.nonzero(site_e >= 0)[0]
    if not len(has):
        return out
    elems, rows = np.unique(site_e[has], return_inverse=True)
    elems = elems.astype(np.int32)
    base = np.zeros((len(elems), 81))
    for step, du in chain:
        base = model.state_accumulate(step, du, elems, base, q_only=True)
    want = [float(T) for T in temps]
    qs = {}
    if 25.0 in want:
        qs[25.0] = base
    hi, lo = max(want), min(want)
    for br, end in ((up, hi), (down, lo)):
        if end == 25.0:
            continue
        st = base.copy()
        for step, du in br:
            st = model.state_accumulate(step, du, elems, st, q_only=True)
            Tb = float(step.c.Tb)
            if Tb in want:
                qs[Tb] = st.copy()
            if Tb == end:
                break
    for ti, T in enumerate(want):
        if T not in qs:
            continue
        _, eps = model.state_stress(elems, qs[T], rows.astype(np.int32), xez[has])
        out[has, ti] = eps
    return out


def screen_joints(model, chain, up, down, temps, d, zmid, part, solder_params, solder: MaterialTables, cycling,
                  pillar: MaterialTables | None = None, fs=1.0, variant="hip"):
    """screenJoints (src/core/13-fatigue.js:375-455) from a solved thermal
    path: sites kept for screening, their bump-layer strain table, the cycle
    history and its per-step data on the host, the per-joint loop on the
    device; dW, dE, stab per site and the top 20 by dW."""
    from . import hip_solver as hs
    sites, idx = screening_sites(d, SCREEN_MAX_DDI if pillar is not None else None)
    table = site_strain_table(model, chain, up, down, temps, sites, zmid, part)
    hist = cycle_history(cycling)
    steps = step_data(hist, temps, solder, pillar, fs)
    want = [float(T) for T in temps]
    i25 = want.index(25.0) if 25.0 in want else -1
    dW, dE, stab = hs.screen_joints(solder_params, steps, table, i25, variant=variant)
    return {"x": sites["x"], "y": sites["y"], "ix": sites["ix"], "iy": sites["iy"], "dW": dW, "dE": dE,
            "stab": stab, "subsampled": idx is not None, "nSites": len(sites["x"]), "top": rank(dW),
            "cycles": cycling["cycles"], "table": table}
