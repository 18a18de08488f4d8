"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_i32p = C.POINTER(C.c_int32)
_i8p = C.POINTER(C.c_int8)
_u8p = C.POINTER(C.c_uint8)
_f64p = C.POINTER(C.c_double)


class OrModel(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("nn", C.c_int32), ("ne", C.c_int32),
        ("xs", _f64p), ("ys", _f64p), ("zs", _f64p),
        ("node_id", _i32p), ("cell_elems", _i32p), ("e_group", _i32p), ("e_part", _i8p),
        ("e_scale", _f64p), ("e_ijk", _i32p), ("conn", _i32p), ("coords", _f64p), ("col_start", _i32p),
        ("ngroups", C.c_int32), ("D", _f64p), ("s0", _f64p), ("born", _u8p),
        ("e_section", _i32p), ("nsec", C.c_int32), ("secW", _f64p), ("secV", _f64p), ("sec_born", _u8p),
        ("nconstraints", C.c_int32), ("constraints", _i32p),
        ("nfx", C.c_int32), ("nfy", C.c_int32), ("feature_x", _i32p), ("feature_y", _i32p),
    ]


class OrStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("converged", C.c_int32), ("relres", C.c_double), ("r0", C.c_double),
                ("ms_total", C.c_double), ("ms_assemble", C.c_double), ("ms_setup", C.c_double),
                ("ms_pcg", C.c_double), ("levels", C.c_int32), ("coarsest_dof", C.c_int32),
                ("fallback_iters", C.c_int32), ("pad", C.c_int32), ("fallback_relres", C.c_double),
                ("ic_shift", C.c_double)]


class OrWarpage(C.Structure):
    _fields_ = [("signed_mag", C.c_double), ("mag", C.c_double), ("sign", C.c_int32),
                ("c0", C.c_double), ("c1", C.c_double), ("c2", C.c_double),
                ("rmin", C.c_double), ("rmax", C.c_double)]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_build.restype = C.c_void_p
        L.or_build.argtypes = [C.POINTER(OrModel)]
        L.or_free.argtypes = [C.c_void_p]
        L.or_sys_nn.restype = C.c_int32
        L.or_sys_nn.argtypes = [C.c_void_p]
        L.or_sys_nb.restype = C.c_int64
        L.or_sys_nb.argtypes = [C.c_void_p]
        for f in ("or_sys_rowptr", "or_sys_col"):
            getattr(L, f).restype = _i32p
            getattr(L, f).argtypes = [C.c_void_p]
        for f in ("or_sys_val", "or_sys_rhs"):
            getattr(L, f).restype = _f64p
            getattr(L, f).argtypes = [C.c_void_p]
        L.or_sys_assemble_ms.restype = C.c_double
        L.or_sys_assemble_ms.argtypes = [C.c_void_p]
        L.or_sys_ntypes.restype = C.c_int32
        L.or_sys_ntypes.argtypes = [C.c_void_p]
        L.or_matvec.argtypes = [C.c_void_p, _f64p, _f64p]
        L.or_solve.argtypes = [C.c_void_p, C.c_double, C.c_int32, C.c_int32, _f64p, C.POINTER(OrStats)]
        L.or_univ.argtypes = [_f64p, _f64p, _f64p, _f64p]
        L.or_shape_matrices.argtypes = [_f64p, _f64p, _f64p]
        L.or_element_k33.argtypes = [C.c_double, C.c_double, C.c_double, _f64p, _f64p]
        L.or_element_fth.argtypes = [C.c_double, C.c_double, C.c_double, _f64p, _f64p, _f64p]
        L.or_condense33.argtypes = [_f64p, _f64p, _f64p, _f64p, _f64p, _f64p]
        L.or_box_gauss_k33.argtypes = [C.c_double, C.c_double, C.c_double, _f64p, _f64p, _f64p, _f64p]
        L.or_element_step_points.argtypes = [C.c_double, C.c_double, C.c_double, _f64p, _f64p, _f64p, C.c_int,
                                             _f64p, _f64p, C.c_int, _f64p, _f64p, _f64p]
        L.or_chol_dense.argtypes = [_f64p, C.c_int, C.c_double]
        L.or_chol_dense.restype = C.c_int
        L.or_fit_plane.argtypes = [_f64p, _f64p, _f64p, C.c_int32, _f64p]
        L.or_warpage_jeita.argtypes = [C.c_int32, C.c_int32, C.c_int32, _f64p, _f64p, _i32p, _f64p, _f64p, _f64p,
                                       C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.POINTER(OrWarpage)]
        L.or_anand_batch.argtypes = [_f64p, C.c_int64, C.c_double, C.c_double, C.c_double, _f64p, _f64p, _f64p,
                                     _f64p, _f64p, _f64p]
        L.or_anand_batch_robust.argtypes = L.or_anand_batch.argtypes
        L.or_screen_joints.argtypes = [_f64p, C.POINTER(OrScreenSteps), C.c_int64, C.c_int32, C.c_int32, _f64p,
                                       _f64p, _f64p, _f64p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---- element -----------------------------------------------------------------------

def shape_matrices(D=None, W=None):
    M = np.zeros(6 * 1089)
    D = None if D is None else f64(D)
    W = None if W is None else f64(W)
    lib().or_shape_matrices(_p(W, _f64p), _p(D, _f64p), _p(M, _f64p))
    return M


def element_k33(hx, hy, hz, M):
    K = np.zeros(1089)
    lib().or_element_k33(hx, hy, hz, _p(f64(M), _f64p), _p(K, _f64p))
    return K.reshape(33, 33)


def element_fth(hx, hy, hz, s0=None, V=None):
    f = np.zeros(33)
    s0 = None if s0 is None else f64(s0)
    V = None if V is None else f64(V)
    lib().or_element_fth(hx, hy, hz, _p(V, _f64p), _p(s0, _f64p), _p(f, _f64p))
    return f


def condense33(K, f=None):
    K = f64(K).reshape(-1)
    Kc = np.zeros(576); fc = np.zeros(24); X = np.zeros(216); xf = np.zeros(9)
    ff = None if f is None else f64(f)
    lib().or_condense33(_p(K, _f64p), _p(ff, _f64p), _p(Kc, _f64p), _p(fc, _f64p), _p(X, _f64p), _p(xf, _f64p))
    return Kc.reshape(24, 24), fc, X.reshape(9, 24), xf


def element_step_points(hx, hy, hz, M, D, s0, born, du24, q, xez, sig=None, eps=None):
    """One step of a homogeneous element's stress history at natural points
    xez (n x 3): q (33) is updated in place; returns sig, eps (n x 6)
    accumulated onto the given arrays (src/core/09-assembly.js:494-571)."""
    xez = f64(xez).reshape(-1, 3)
    n = len(xez)
    sig = np.zeros((n, 6)) if sig is None else sig
    eps = np.zeros((n, 6)) if eps is None else eps
    lib().or_element_step_points(hx, hy, hz, _p(f64(M), _f64p), _p(f64(D), _f64p), _p(f64(s0), _f64p),
                                 1 if born else 0, _p(f64(du24), _f64p), _p(q, _f64p), n, _p(xez, _f64p),
                                 _p(sig, _f64p), _p(eps, _f64p))
    return sig, eps


def box_gauss_k33(hx, hy, hz, D, s0=None):
    K = np.zeros(1089); f = np.zeros(33)
    s0 = None if s0 is None else f64(s0)
    lib().or_box_gauss_k33(hx, hy, hz, _p(f64(D), _f64p), _p(s0, _f64p), _p(K, _f64p), _p(f, _f64p))
    return K.reshape(33, 33), f


# ---- system ------------------------------------------------------------------------

class System:
    """Assembled full block-CSR system of one step of a mesh (owns the C handle)."""

    def __init__(self, mesh, D, s0, born):
        self._keep = []
        k = self._keep.append
        m = OrModel()
        m.nx, m.ny, m.nz, m.nn, m.ne = mesh.nx, mesh.ny, mesh.nz, mesh.nn, mesh.ne
        arrs = dict(
            xs=f64(mesh.xs), ys=f64(mesh.ys), zs=f64(mesh.zs),
            node_id=np.ascontiguousarray(mesh.node_id, np.int32),
            cell_elems=np.ascontiguousarray(mesh.cell_elems, np.int32),
            e_group=np.ascontiguousarray(mesh.e_group, np.int32),
            e_part=np.ascontiguousarray(mesh.e_part, np.int8),
            e_scale=f64(mesh.e_scale),
            e_ijk=np.ascontiguousarray(mesh.e_ijk, np.int32).reshape(-1),
            conn=np.ascontiguousarray(mesh.conn, np.int32).reshape(-1),
            coords=f64(mesh.coords).reshape(-1),
            col_start=np.ascontiguousarray(mesh.col_start, np.int32),
            D=f64(D).reshape(-1), s0=f64(s0).reshape(-1), born=np.ascontiguousarray(born, np.uint8),
            constraints=np.ascontiguousarray(mesh.constraint_dofs, np.int32),
            feature_x=np.ascontiguousarray(mesh.feature_x, np.int32),
            feature_y=np.ascontiguousarray(mesh.feature_y, np.int32),
        )
        for v in arrs.values():
            k(v)
        for name in ("xs", "ys", "zs", "e_scale", "coords", "D", "s0"):
            setattr(m, name, _p(arrs[name], _f64p))
        for name in ("node_id", "cell_elems", "e_group", "e_ijk", "conn", "col_start", "constraints",
                     "feature_x", "feature_y"):
            setattr(m, name, _p(arrs[name], _i32p))
        m.e_part = _p(arrs["e_part"], _i8p)
        m.born = _p(arrs["born"], _u8p)
        m.ngroups = len(arrs["born"])
        m.e_section = None
        m.nsec = 0
        m.nconstraints = len(arrs["constraints"])
        m.nfx = len(arrs["feature_x"])
        m.nfy = len(arrs["feature_y"])
        self.model = m
        self.mesh = mesh
        self.h = lib().or_build(C.byref(m))
        self.nn = lib().or_sys_nn(self.h)
        self.ndof = 3 * self.nn
        self.solved = False

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_free(self.h)
            self.h = None

    @property
    def nb(self):
        return lib().or_sys_nb(self.h)

    def rhs(self):
        return np.ctypeslib.as_array(lib().or_sys_rhs(self.h), shape=(self.ndof,)).copy()

    def csr(self):
        """(rowPtr, col, val[nb, 3, 3]) copies of the assembled block CSR."""
        nb = self.nb
        rp = np.ctypeslib.as_array(lib().or_sys_rowptr(self.h), shape=(self.nn + 1,)).copy()
        col = np.ctypeslib.as_array(lib().or_sys_col(self.h), shape=(nb,)).copy()
        val = np.ctypeslib.as_array(lib().or_sys_val(self.h), shape=(nb * 9,)).copy().reshape(nb, 3, 3)
        return rp, col, val

    def matvec(self, x):
        assert not self.solved, "values are Jacobi-scaled in place by solve()"
        x = f64(x)
        y = np.zeros(self.ndof)
        lib().or_matvec(self.h, _p(x, _f64p), _p(y, _f64p))
        return y

    def solve(self, tol=1e-10, maxit=2000, precond=0):
        u = np.zeros(self.ndof)
        st = OrStats()
        rc = lib().or_solve(self.h, tol, maxit, precond, _p(u, _f64p), C.byref(st))
        if rc != 0:
            raise RuntimeError("IC(0) failed even with the maximum diagonal shift")
        self.solved = True
        return u, {f: getattr(st, f) for f, _ in OrStats._fields_}

    @property
    def assemble_ms(self):
        return lib().or_sys_assemble_ms(self.h)

    @property
    def ntypes(self):
        return lib().or_sys_ntypes(self.h)


def warpage_jeita(mesh, u, bga):
    """JEITA warpage of the bottom plane (src/core/12-post.js:60-99)."""
    from_x, from_y, zone = bga
    out = OrWarpage()
    bx, by = f64(from_x), f64(from_y)
    nid = np.ascontiguousarray(mesh.node_id, np.int32)
    lib().or_warpage_jeita(mesh.nx, mesh.ny, mesh.nz, _p(f64(mesh.xs), _f64p), _p(f64(mesh.ys), _f64p),
                           _p(nid, _i32p), _p(f64(u), _f64p), _p(bx, _f64p), _p(by, _f64p), len(bx),
                           zone[0], zone[1], zone[2], zone[3], C.byref(out))
    return {f: getattr(out, f) for f, _ in OrWarpage._fields_}


def anand_batch(P, eps, T, dt, E, state, tangent=True, robust=False):
    """Oracle AnandPoint.step over a batch; P = (A, QR, xi, m, shat, n, h0, a, s0, nu).
    robust: the bracket-end guard of the global viscoplastic Newton."""
    Pa = f64(P)
    eps = f64(eps).reshape(-1, 6)
    n = eps.shape[0]
    assert state.shape == (n, 11) and state.flags.c_contiguous and state.dtype == np.float64
    tr = np.zeros((n, 11)); sig = np.zeros((n, 6)); dep = np.zeros(n)
    D = np.zeros((n, 36)) if tangent else None
    fn = lib().or_anand_batch_robust if robust else lib().or_anand_batch
    fn(_p(Pa, _f64p), n, E, T, dt, _p(eps, _f64p), _p(state, _f64p), _p(tr, _f64p), _p(sig, _f64p), _p(D, _f64p),
       _p(dep, _f64p))
    return tr, sig, D, dep


class OrScreenSteps(C.Structure):
    _fields_ = [("nsteps", C.c_int32), ("cycles", C.c_int32)] + \
        [(k, _f64p) for k in ("T", "dt", "E", "Fth", "wt", "Cp", "Fp")] + \
        [("wi", _i32p), ("cycle", _i32p), ("same", C.POINTER(C.c_uint8)), ("fs", C.c_double)]


def screen_joints(P, steps, table, ntemps, i25):
    """Oracle screening loop (screenJoints, src/core/13-fatigue.js:416-451) over
    per-step data `steps` (keys of the product's step_data) and a strain table
    (nsites x ntemps x 6): per-site dW, dE, stab."""
    Pa = f64(P)
    tab = f64(table).reshape(-1)
    n = len(tab) // (6 * ntemps)
    keep = {}

    def arr(k, dt):
        v = steps[k]
        if v is None:
            return None
        keep[k] = np.ascontiguousarray(v, dt)
        return keep[k]

    h = OrScreenSteps()
    h.nsteps, h.cycles, h.fs = steps["nsteps"], steps["cycles"], steps["fs"]
    for k in ("T", "dt", "E", "Fth", "wt", "Cp", "Fp"):
        setattr(h, k, _p(arr(k, np.float64), _f64p))
    h.wi = _p(arr("wi", np.int32), _i32p)
    h.cycle = _p(arr("cycle", np.int32), _i32p)
    h.same = _p(arr("same", np.uint8), C.POINTER(C.c_uint8))
    dW, dE, stab = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().or_screen_joints(_p(Pa, _f64p), C.byref(h), n, ntemps, i25, _p(tab, _f64p), _p(dW, _f64p),
                           _p(dE, _f64p), _p(stab, _f64p))
    return dW, dE, stab
