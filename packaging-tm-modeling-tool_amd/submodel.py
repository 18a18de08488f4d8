"""Bump submodel through the temperature cycle (BumpSubmodel.run,
src/core/13-fatigue.js:963-1090) over the device engine (fcbga_submodel_*).

The engine holds the submodel (B-bar hexahedra on a structured grid, Anand
solder points, elastic substrate slice, underfill and silicon) and does each
time step on the device (Newton with IC(0)-PCG inner solves, commit, cutback
snapshot, layer averages).  This host side keeps what the reference's run()
does around stepTo: the cycle history, the adaptive sub-stepping with
cutbacks, the step's material data from the 1 degC property grids (_elastic,
:694-742), the prescribed boundary displacement (prescribe, :680-692), the
hysteresis samples, the per-cycle work densities and the result record."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import hip_solver as hs
from .meshspec import iso_d
from .screening import TGRID_MIN, TGRID_N, cycle_history, find_interval

SUB, SOLDER, UF, SI = 0, 1, 2, 3
NU_MAX_SUBMODEL = 0.49          # src/core/00-constants.js:20
INELASTIC_STEP_MAX = 2e-4       # :113
SUB_T_AVG = 0.025               # :114
STABILIZATION_WARN = 0.05       # :120
SYED_CW, SYED_CSTRAIN = 0.00165, 0.0513


# ---- 1 degC property grids (src/core/03-materials.js:347-420) ----------------------

def grid_at(arr, T, stride=1, comp=0):
    x = T - TGRID_MIN
    if x <= 0:
        return float(arr[comp])
    if x >= TGRID_N - 1:
        return float(arr[(TGRID_N - 1) * stride + comp])
    i = int(math.floor(x))
    t = x - i
    a, b = float(arr[i * stride + comp]), float(arr[(i + 1) * stride + comp])
    return a + t * (b - a)


def grid_add_at(grid, T, stride, n, w, out):
    x = min(max(T - TGRID_MIN, 0.0), TGRID_N - 1)
    i = min(int(math.floor(x)), TGRID_N - 2)
    t = x - i
    a, b = i * stride, i * stride + stride
    for c in range(n):
        ga, gb = float(grid[a + c]), float(grid[b + c])
        out[c] += w * (ga + t * (gb - ga))


def for_grid_segments(Ta, Tb, fn):
    lo, hi = min(Ta, Tb), max(Ta, Tb)
    t0 = lo
    while t0 < hi - 1e-12:
        t1 = min(hi, TGRID_MIN + math.floor(t0 - TGRID_MIN + 1e-9) + 1)
        fn(0.5 * (t0 + t1), t0, t1)
        t0 = t1


def grid_mean(grid, stride, n, Ta, Tb):
    out = [0.0] * n
    L = abs(Tb - Ta)
    if L < 1e-12:
        grid_add_at(grid, Ta, stride, n, 1.0, out)
        return out
    for_grid_segments(Ta, Tb, lambda tm, t0, t1: grid_add_at(grid, tm, stride, n, (t1 - t0) / L, out))
    return out


def grid_stress_increment(Dg, Fg, Ta, Tb):
    out = [0.0] * 6
    sg = 1 if Tb >= Ta else -1

    def seg(tm, t0, t1):
        D = [0.0] * 36
        grid_add_at(Dg, tm, 36, 36, 1.0, D)
        f = [grid_at(Fg, t1, 3, c) - grid_at(Fg, t0, 3, c) for c in range(3)]
        for r in range(6):
            out[r] += sg * (D[6 * r] * f[0] + D[6 * r + 1] * f[1] + D[6 * r + 2] * f[2])

    for_grid_segments(Ta, Tb, seg)
    return out


def grid_modulus_strain_increment(E, F, Ta, Tb):
    sg = 1 if Tb >= Ta else -1
    acc = [0.0]

    def seg(tm, t0, t1):
        acc[0] += grid_at(E, tm) * (grid_at(F, t1) - grid_at(F, t0))

    for_grid_segments(Ta, Tb, seg)
    return sg * acc[0]


class Tables:
    """E (MPa) and F grids of an isotropic material and its Poisson ratio."""

    def __init__(self, E, F, nu):
        self.E, self.F, self.nu = np.asarray(E, np.float64), np.asarray(F, np.float64), nu


# ---- C ABI ---------------------------------------------------------------------------

_i8p, _u8p = C.POINTER(C.c_int8), C.POINTER(C.c_uint8)
_f64p = C.POINTER(C.c_double)


class FcbgaSubmodelMesh(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("xs", _f64p), ("ys", _f64p),
                ("zs", _f64p), ("cell_mat", _i8p), ("cell_layer", _u8p)]


class FcbgaSubmodelStepIn(C.Structure):
    _fields_ = [("T", C.c_double), ("dt", C.c_double), ("E", C.c_double), ("fsolder", C.c_double),
                ("elas_D", _f64p), ("elas_sth", _f64p), ("ub", _f64p)]


class FcbgaSubmodelStepOut(C.Structure):
    _fields_ = [("converged", C.c_int32), ("newton", C.c_int32), ("pcg_iters", C.c_int32),
                ("ic_builds", C.c_int32), ("dmax", C.c_double)]


def _bind(L):
    if getattr(L, "_sub_bound", False):
        return L
    L.fcbga_submodel_create.argtypes = [C.POINTER(FcbgaSubmodelMesh), C.POINTER(hs.FcbgaAnandParams),
                                        C.POINTER(C.c_void_p)]
    L.fcbga_submodel_create.restype = C.c_int
    L.fcbga_submodel_destroy.argtypes = [C.c_void_p]
    L.fcbga_submodel_nboundary.argtypes = [C.c_void_p]
    L.fcbga_submodel_nboundary.restype = C.c_int32
    L.fcbga_submodel_step.argtypes = [C.c_void_p, C.POINTER(FcbgaSubmodelStepIn), C.POINTER(FcbgaSubmodelStepOut)]
    L.fcbga_submodel_step.restype = C.c_int
    for f in ("fcbga_submodel_save", "fcbga_submodel_restore"):
        getattr(L, f).argtypes = [C.c_void_p]
        getattr(L, f).restype = C.c_int
    L.fcbga_submodel_layer_means.argtypes = [C.c_void_p, _f64p]
    L.fcbga_submodel_layer_means.restype = C.c_int
    L.fcbga_submodel_loop_sample.argtypes = [C.c_void_p] + [C.c_double] * 5 + [_f64p]
    L.fcbga_submodel_loop_sample.restype = C.c_int
    L.fcbga_submodel_section.argtypes = [C.c_void_p, C.c_double, C.c_double, C.POINTER(C.c_float)]
    L.fcbga_submodel_section.restype = C.c_int
    L.fcbga_submodel_displacement.argtypes = [C.c_void_p, _f64p]
    L.fcbga_submodel_displacement.restype = C.c_int
    L._sub_bound = True
    return L


class Submodel:
    """One bump submodel.  setup: xs, ys, zs (planes), cell materials (z
    fastest), zt / zaf (substrate top, die active face), site {x, y}, die
    {cx, cy, rt, rb}, solder Anand parameters (dict with nu), fatigue
    constants, cycling {Tmin, Tmax, ramp, dwell, cycles}; tables: solder and
    underfill Tables, subD (TGRID_N x 36), subF (TGRID_N x 3), siD (36), siF
    (TGRID_N); boundary: (temps, values) with values[i] the prescribed
    u(T) - u(25) of the face nodes (grid order) at temps[i] and temps_all the
    evaluation temperatures the global field is sampled at."""

    def __init__(self, setup, cell_mat, solder_tab, uf_tab, subD, subF, siD, siF, btemps, bvals, temps_all,
                 variant="hip"):
        self.setup = setup
        self.L = _bind(hs.load(variant))
        self.xs, self.ys, self.zs = (np.ascontiguousarray(setup[k], np.float64) for k in ("xs", "ys", "zs"))
        nx, ny, nz = len(self.xs), len(self.ys), len(self.zs)
        self.cell_mat = np.ascontiguousarray(cell_mat, np.int8)
        assert self.cell_mat.shape == ((nx - 1) * (ny - 1) * (nz - 1),)
        # averaging layers of the solder points: die side (bit 0), substrate side (bit 1)
        zc = 0.5 * (self.zs[:-1] + self.zs[1:])
        kz = np.arange(len(self.cell_mat)) % (nz - 1)
        lay = np.zeros(len(self.cell_mat), np.uint8)
        sol = self.cell_mat == SOLDER
        lay[sol & (zc[kz] > setup["zaf"] - SUB_T_AVG)] |= 1
        lay[sol & (zc[kz] < setup["zt"] + SUB_T_AVG)] |= 2
        self.cell_layer = lay
        self.solder, self.uf = solder_tab, uf_tab
        self.subD, self.subF = np.asarray(subD, np.float64), np.asarray(subF, np.float64)
        self.siD, self.siF = np.asarray(siD, np.float64), np.asarray(siF, np.float64)
        self.btemps = [float(t) for t in btemps]
        self.bvals = np.asarray(bvals, np.float64).reshape(len(self.btemps), -1)
        self.temps_all = [float(t) for t in temps_all]
        sp = setup["solder"]
        self.params = hs.anand_params(sp["A"], sp["QR"], sp["xi"], sp["m"], sp["shat"], sp["n"], sp["h0"], sp["a"],
                                      sp["s0"], sp["nu"])
        m = FcbgaSubmodelMesh()
        m.nx, m.ny, m.nz = nx, ny, nz
        m.xs, m.ys, m.zs = (a.ctypes.data_as(_f64p) for a in (self.xs, self.ys, self.zs))
        m.cell_mat = self.cell_mat.ctypes.data_as(_i8p)
        m.cell_layer = self.cell_layer.ctypes.data_as(_u8p)
        h = C.c_void_p()
        rc = self.L.fcbga_submodel_create(C.byref(m), C.byref(self.params), C.byref(h))
        if rc != 0:
            raise hs.SolverError(hs.last_error(self.L))
        self.h = h
        self.nb = self.L.fcbga_submodel_nboundary(h)
        assert self.bvals.shape[1] == 3 * self.nb
        self.Tn = 25.0

    def close(self):
        if getattr(self, "h", None):
            self.L.fcbga_submodel_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # -- step data ----------------------------------------------------------------------
    def _iso_step(self, tab, Ta, Tb):
        Em = grid_mean(tab.E, 1, 1, Ta, Tb)[0]
        ef = grid_modulus_strain_increment(tab.E, tab.F, Ta, Tb) / Em
        D = iso_d(Em, tab.nu, NU_MAX_SUBMODEL).reshape(-1)
        return D, [(D[6 * r] + D[6 * r + 1] + D[6 * r + 2]) * ef for r in range(6)]

    def elastic(self, Ta, Tb):
        """4 x 36 mean stiffness and 4 x 6 thermal stress increment of the step
        (BumpSubmodel._elastic, src/core/13-fatigue.js:694-742, separate phases)."""
        D = np.zeros((4, 36))
        sth = np.zeros((4, 6))
        D[SUB] = grid_mean(self.subD, 36, 36, Ta, Tb)
        sth[SUB] = grid_stress_increment(self.subD, self.subF, Ta, Tb)
        for mat, tab in ((SOLDER, self.solder), (UF, self.uf)):
            d, s = self._iso_step(tab, Ta, Tb)
            D[mat], sth[mat] = d, s
        D[SI] = self.siD
        f = grid_at(self.siF, Tb) - grid_at(self.siF, Ta)
        sth[SI] = [(self.siD[6 * r] + self.siD[6 * r + 1] + self.siD[6 * r + 2]) * f for r in range(6)]
        return D, sth

    def prescribed(self, T):
        """u(T) - u(25) at the face nodes: the global field interpolated in T
        between its sample temperatures (Analysis.fieldAt)."""
        ts = self.temps_all
        i = find_interval(ts, T)
        t = min(max((T - ts[i]) / (ts[i + 1] - ts[i]), 0.0), 1.0)
        a = self.bvals[self.btemps.index(ts[i])]
        if t == 0:
            return a.copy()
        b = self.bvals[self.btemps.index(ts[i + 1])]
        return a + t * (b - a)

    def step(self, T, dt):
        D, sth = self.elastic(self.Tn, T)
        D = np.ascontiguousarray(D)
        sth = np.ascontiguousarray(sth)
        ub = np.ascontiguousarray(self.prescribed(T))
        si = FcbgaSubmodelStepIn()
        si.T, si.dt = T, dt
        si.E = grid_at(self.solder.E, T)
        si.fsolder = grid_at(self.solder.F, T) - grid_at(self.solder.F, 25.0)
        si.elas_D, si.elas_sth, si.ub = (a.ctypes.data_as(_f64p) for a in (D, sth, ub))
        so = FcbgaSubmodelStepOut()
        rc = self.L.fcbga_submodel_step(self.h, C.byref(si), C.byref(so))
        if rc != 0:
            raise hs.SolverError(hs.last_error(self.L))
        if so.converged:
            self.Tn = T
        return {"converged": bool(so.converged), "newton": so.newton, "dMax": so.dmax, "pcg": so.pcg_iters,
                "icBuilds": so.ic_builds}

    def save(self):
        if self.L.fcbga_submodel_save(self.h) != 0:
            raise hs.SolverError(hs.last_error(self.L))
        self._saveTn = self.Tn

    def restore(self):
        if self.L.fcbga_submodel_restore(self.h) != 0:
            raise hs.SolverError(hs.last_error(self.L))
        self.Tn = self._saveTn

    def layer_means(self):
        out = np.zeros(4)
        if self.L.fcbga_submodel_layer_means(self.h, out.ctypes.data_as(_f64p)) != 0:
            raise hs.SolverError(hs.last_error(self.L))
        return out

    def loop_sample(self, T, cphi, sphi):
        out = np.zeros(4)
        E = grid_at(self.solder.E, T)
        f = grid_at(self.solder.F, T) - grid_at(self.solder.F, 25.0)
        if self.L.fcbga_submodel_loop_sample(self.h, T, E, f, cphi, sphi, out.ctypes.data_as(_f64p)) != 0:
            raise hs.SolverError(hs.last_error(self.L))
        return out

    def section(self, T):
        """Corner stresses of every element at the committed state (48 float32 per element)."""
        out = np.zeros(48 * len(self.cell_mat), np.float32)
        E = grid_at(self.solder.E, T)
        f = grid_at(self.solder.F, T) - grid_at(self.solder.F, 25.0)
        if self.L.fcbga_submodel_section(self.h, E, f, out.ctypes.data_as(C.POINTER(C.c_float))) != 0:
            raise hs.SolverError(hs.last_error(self.L))
        return out

    def displacement(self):
        u = np.zeros(3 * len(self.xs) * len(self.ys) * len(self.zs))
        if self.L.fcbga_submodel_displacement(self.h, u.ctypes.data_as(_f64p)) != 0:
            raise hs.SolverError(hs.last_error(self.L))
        return u

    # -- the cycle ----------------------------------------------------------------------
    def run(self, max_steps=None, sections=True):
        """The cycle with adaptive sub-stepping (BumpSubmodel.run): returns the
        result record, with the corner stresses at the ends of the last cycle's
        ramps and dwells when sections.  max_steps stops after that many
        history steps (tests)."""
        st = self.setup
        hist = cycle_history(st["cycling"])
        nsteps = len(hist["t"]) if max_steps is None else min(len(hist["t"]), max_steps + 1)
        phi = math.atan2(st["site"]["y"] - st["die"]["cy"], st["site"]["x"] - st["die"]["cx"])
        cphi, sphi = math.cos(phi), math.sin(phi)
        loop = {k: [] for k in ("t", "T", "tau", "gamma", "tauSub", "gammaSub")}
        nc = hist["cycles"]
        Wd, Ws, Ed, Es = (np.zeros(nc + 1) for _ in range(4))
        total_newton = cutbacks = steps = 0
        nsub_next = 1
        trace = []
        snaps = []

        def sample(T, t):
            s = self.loop_sample(T, cphi, sphi)
            loop["t"].append(t); loop["T"].append(T)
            loop["tau"].append(s[0]); loop["gamma"].append(s[1]); loop["tauSub"].append(s[2]); loop["gammaSub"].append(s[3])

        sample(25.0, 0.0)
        for k in range(1, nsteps):
            T0, T1, t0s, t1s = hist["T"][k - 1], hist["T"][k], hist["t"][k - 1], hist["t"][k]
            nsub, done, d_last = nsub_next, False, 0.0
            for _ in range(8):
                if done:
                    break
                self.save()
                ok = True
                for s in range(1, nsub + 1):
                    T = T0 + (T1 - T0) * s / nsub
                    dt = (t1s - t0s) / nsub
                    res = self.step(T, dt)
                    if len(trace) < 6:
                        trace.append(dict(res, T=T, dt=dt))
                    total_newton += res["newton"]
                    steps += 1
                    d_last = res["dMax"]
                    if not res["converged"] or res["dMax"] > INELASTIC_STEP_MAX:
                        ok = False
                        break
                if ok:
                    done = True
                else:
                    self.restore()
                    nsub *= 2
                    cutbacks += 1
            nsub_next = (min(nsub * 2, 64) if d_last > 0.5 * INELASTIC_STEP_MAX else
                         (max(1, nsub >> 1) if d_last < 0.2 * INELASTIC_STEP_MAX else nsub))
            if not done:
                raise hs.SolverError(f"Submodel step failed to converge after cutbacks at T = {T1:.1f} C")
            sample(T1, t1s)
            c = int(hist["cycle"][k])
            if sections and c == nc and (k == len(hist["t"]) - 1 or hist["phase"][k + 1] != hist["phase"][k]):
                snaps.append({"T": T1, "t": t1s, "cycle": c, "phase": hist["phase"][k], "sig": self.section(T1)})
            m = self.layer_means()
            Wd[c], Ed[c], Ws[c], Es[c] = m[0], m[1], m[2], m[3]
        dWdie, dWsub = Wd[nc] - Wd[nc - 1], Ws[nc] - Ws[nc - 1]
        dEdie, dEsub = Ed[nc] - Ed[nc - 1], Es[nc] - Es[nc - 1]
        dW = max(dWdie, dWsub)
        dE = dEdie if dWdie >= dWsub else dEsub
        prev = max(Wd[nc - 1] - Wd[nc - 2], Ws[nc - 1] - Ws[nc - 2]) if nc >= 2 else math.nan
        stab = abs(dW - prev) / prev if math.isfinite(prev) and prev > 0 else math.nan
        die = st["die"]
        d_joint = 2 * die["rt"] if dWdie >= dWsub else 2 * die["rb"]
        fat = st.get("fatigue") or {}
        life = {"syedEnergy": 1 / ((fat.get("CW") or SYED_CW) * dW) if dW > 0 else math.inf,
                "syedStrain": 1 / ((fat.get("Cstrain") or SYED_CSTRAIN) * dE) if dE > 0 else math.inf}
        return {"dW": dW, "dWdie": dWdie, "dWsub": dWsub, "dE": dE, "dEdie": dEdie, "dEsub": dEsub,
                "critical": "die side" if dWdie >= dWsub else "substrate side", "stabChange": stab,
                "stabWarn": bool(math.isfinite(stab) and stab > STABILIZATION_WARN),
                "perCycle": [{"cycle": c, "dWdie": Wd[c] - Wd[c - 1], "dWsub": Ws[c] - Ws[c - 1]}
                             for c in range(1, nc + 1)],
                "loop": loop, "life": life, "dJoint": d_joint, "steps": steps, "cutbacks": cutbacks,
                "newtonPerStep": total_newton / max(steps, 1), "first": trace,
                "section": {"xs": self.xs, "ys": self.ys, "zs": self.zs, "grp": self.cell_mat, "snaps": snaps}}


def run_batch(models, workers=4, **run_kw):
    """Several submodels at once (the batch form submodelBatch(jobs) -> records
    of the many-bump workload; the reference runs one BumpSubmodel per helper
    worker, src/worker.js:96-104, 230, 291).  Each Submodel owns a device
    stream, and each run's host steps go in a thread of their own (the library
    calls release the interpreter lock), so the runs overlap on the device.
    How many overlap is bounded by the process's hardware queues (HIP's
    GPU_MAX_HW_QUEUES, 4 by default, read when the process starts HIP).
    Returns the records in the order of models."""
    from concurrent.futures import ThreadPoolExecutor
    if not models:
        return []
    with ThreadPoolExecutor(max(1, min(int(workers), len(models)))) as ex:
        return list(ex.map(lambda m: m.run(**run_kw), models))
