"""ctypes binding of the C ABI in include/fcbga_hip.h.

`Model(mesh)` wraps fcbga_model_create (the FEModel constructor's role,
src/core/09-assembly.js:208-233) and `Model.solve_step(step, ...)` wraps
fcbga_step_solve (assembleStep + LinearSolver.solve, src/core/11-analysis.js:189-206).
The product library is libfcbga_hip.so (HIP, gfx950); loading fails loudly
when it is missing.  `variant="hostemu"` loads the host-emulation build of the
same sources, which only the test suite uses to check indexing without a GPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {"hip": os.path.join(HERE, "libfcbga_hip.so"), "hostemu": os.path.join(HERE, "libfcbga_hostemu.so")}
# FCBGA_HOSTEMU_SANITIZE=1 swaps the host-emulation build for the same build
# under the address and undefined-behaviour sanitizers (the process must then
# preload libasan; tests/test_sanitize.py runs the host-emulation tests so)
if os.environ.get("FCBGA_HOSTEMU_SANITIZE") == "1":
    LIBS["hostemu"] = os.path.join(HERE, "libfcbga_hostemu_asan.so")

PREC_FP64, PREC_MIXED = 0, 1
SMOOTH_GS, SMOOTH_COARSE_JACOBI = 0, 1  # FCBGA_SMOOTH_*
PC_MG, PC_COLUMN, PC_JACOBI, PC_IC = 0, 1, 2, 3
PC_NAMES = {PC_MG: "mg", PC_COLUMN: "column", PC_JACOBI: "jacobi", PC_IC: "ic"}

_i32p = C.POINTER(C.c_int32)
_i8p = C.POINTER(C.c_int8)
_u8p = C.POINTER(C.c_uint8)
_f64p = C.POINTER(C.c_double)


class FcbgaMesh(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("nn", C.c_int32), ("ne", C.c_int32),
                ("xs", _f64p), ("ys", _f64p), ("zs", _f64p), ("node_id", _i32p), ("cell_elems", _i32p),
                ("e_group", _i32p), ("e_part", _i8p), ("e_scale", _f64p), ("e_section", _i32p),
                ("ngroups", C.c_int32), ("nsections", C.c_int32), ("nconstraints", C.c_int32),
                ("constraint_dofs", _i32p), ("nfeature_x", C.c_int32), ("nfeature_y", C.c_int32),
                ("feature_x", _i32p), ("feature_y", _i32p), ("nmodule_constraints", C.c_int32),
                ("module_constraint_dofs", _i32p)]


class FcbgaStep(C.Structure):
    _fields_ = [("Ta", C.c_double), ("Tb", C.c_double), ("stage", C.c_int32), ("D", _f64p), ("s0", _f64p),
                ("born", _u8p), ("secW", _f64p), ("secV", _f64p), ("sec_born", _u8p)]


class FcbgaGroupGrids(C.Structure):
    _fields_ = [("stage", C.c_int32), ("layered", C.c_int32), ("Dgrid", _f64p), ("Dconst", _f64p), ("F", _f64p)]


class FcbgaSectionGrids(C.Structure):
    _fields_ = [("ns", C.c_int32), ("group", C.c_int32), ("W", _f64p), ("D", _f64p), ("F", _f64p), ("Ls", _f64p)]


class FcbgaSolveOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxit", C.c_int32), ("precision", C.c_int32), ("precond", C.c_int32),
                ("sweeps", C.c_int32), ("warm", C.c_int32), ("smoother", C.c_int32)]


class FcbgaSolveStats(C.Structure):
    _fields_ = [("iters", C.c_int32), ("converged", C.c_int32), ("relres", C.c_double), ("r0", C.c_double),
                ("ms_total", C.c_double), ("ms_assemble", C.c_double), ("ms_setup", C.c_double),
                ("ms_pcg", C.c_double), ("levels", C.c_int32), ("coarsest_dof", C.c_int32), ("ndof", C.c_int64),
                ("ntypes", C.c_int32), ("sym_tables", C.c_int32), ("warm", C.c_int32), ("precond", C.c_int32),
                ("fallback_first", C.c_int32), ("fallback_iters", C.c_int32), ("fallback_relres", C.c_double),
                ("ic_shift", C.c_double)]


def _record(st):
    """The solve record: the stats fields, precond as the reference's name and
    fallback None or {first, iters, relres} (LinearSolver.solve,
    src/core/10-solver.js:965-993)."""
    rec = {f: getattr(st, f) for f, _ in FcbgaSolveStats._fields_
           if f not in ("fallback_first", "fallback_iters", "fallback_relres")}
    rec["precond"] = PC_NAMES.get(st.precond, str(st.precond))
    rec["fallback"] = None if st.fallback_first < 0 else {
        "first": PC_NAMES.get(st.fallback_first, str(st.fallback_first)), "iters": st.fallback_iters,
        "relres": st.fallback_relres}
    return rec


class FcbgaScreenSteps(C.Structure):
    _fields_ = [("nsteps", C.c_int32), ("cycles", C.c_int32)] + \
        [(k, _f64p) for k in ("T", "dt", "E", "Fth", "wt", "Cp", "Fp")] + \
        [("wi", _i32p), ("cycle", _i32p), ("same", C.POINTER(C.c_uint8)), ("fs", C.c_double)]


EXPORTS = ["fcbga_model_create", "fcbga_model_destroy", "fcbga_model_ndof", "fcbga_warm_reset", "fcbga_step_solve",
           "fcbga_step_solve_device", "fcbga_step_load", "fcbga_step_apply", "fcbga_set_kernel_timing",
           "fcbga_kernel_times", "fcbga_anand_update", "fcbga_anand_update_device", "fcbga_state_accumulate",
           "fcbga_state_stress", "fcbga_adjoint_forward", "fcbga_adjoint_backward", "fcbga_screen_joints",
           "fcbga_submodel_create", "fcbga_submodel_destroy", "fcbga_submodel_nboundary", "fcbga_submodel_step",
           "fcbga_submodel_save", "fcbga_submodel_restore", "fcbga_submodel_layer_means",
           "fcbga_submodel_loop_sample", "fcbga_submodel_section", "fcbga_submodel_displacement",
           "fcbga_submodel_points", "fcbga_screen_joints_noalloc",
           "fcbga_vp_create", "fcbga_vp_destroy", "fcbga_vp_nsolder", "fcbga_vp_step", "fcbga_vp_save",
           "fcbga_vp_restore", "fcbga_vp_points", "fcbga_vp_displacement", "fcbga_vp_set_displacement",
           "fcbga_vp_residual", "fcbga_set_progress_cb", "fcbga_vp_set_prescribed",
           "fcbga_path_reset", "fcbga_path_solve", "fcbga_path_sum", "fcbga_path_du", "fcbga_path_load",
           "fcbga_path_load64", "fcbga_path_copy",
           "fcbga_mesh_fingerprint", "fcbga_prepare_step", "fcbga_adjoint_backward_u", "fcbga_last_error"]

ANAND_STATE = 11
STATE = 81  # stress-history state of one element (FCBGA_STATE)


class FcbgaObjective(C.Structure):
    _fields_ = [("k_top", C.c_int32), ("n_vm", C.c_int32), ("vm_elems", _i32p), ("p", C.c_double)]


class FcbgaAnandParams(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("A", "QR", "xi", "m", "shat", "n", "h0", "a", "s0", "nu")]


class FcbgaProgress(C.Structure):
    _fields_ = [("event", C.c_int32), ("iters", C.c_int32), ("maxit", C.c_int32), ("precond", C.c_int32),
                ("relres", C.c_double), ("r0", C.c_double), ("tol", C.c_double), ("ms", C.c_double)]


PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.POINTER(FcbgaProgress), C.c_void_p)
ERR_CANCELLED = -5


class FcbgaVpSolder(C.Structure):
    _fields_ = [("group", C.c_int32), ("anand", FcbgaAnandParams), ("nE", C.c_int32), ("E_T", _f64p),
                ("E_v", _f64p), ("cte", C.c_double), ("Tref", C.c_double)]


class FcbgaVpOpts(C.Structure):
    _fields_ = [("inner", FcbgaSolveOpts), ("maxit", C.c_int32), ("res_tol", C.c_double), ("du_tol", C.c_double),
                ("inner_floor", C.c_double)]


class FcbgaVpStats(C.Structure):
    _fields_ = [("converged", C.c_int32), ("newton", C.c_int32), ("pcg_iters", C.c_int32), ("nonfinite", C.c_int32),
                ("dmax", C.c_double), ("res", C.c_double), ("ms", C.c_double), ("ms_setup", C.c_double),
                ("ms_pcg", C.c_double)]

# kernel classes of fcbga_kernel_times (include/fcbga_hip.h)
KCLASS = {"apply": 0, "fine_sweep": 1, "fine_resid": 2, "shell1_sweep": 3, "shell_sweep": 4, "coarsest": 5,
          "ic_apply": 6, "ic_factor": 7}
KCLASS_COUNT = 8

_loaded = {}


class SolverError(RuntimeError):
    pass


def load(variant: str = "hip"):
    if variant in _loaded:
        return _loaded[variant]
    path = LIBS[variant]
    if not os.path.exists(path):
        raise SolverError(f"{os.path.basename(path)} is not built: run __graft_entry__.build() first")
    L = C.CDLL(path)
    L.fcbga_model_create.argtypes = [C.POINTER(FcbgaMesh), C.POINTER(C.c_void_p)]
    L.fcbga_model_create.restype = C.c_int
    L.fcbga_model_destroy.argtypes = [C.c_void_p]
    L.fcbga_model_ndof.argtypes = [C.c_void_p]
    L.fcbga_model_ndof.restype = C.c_int64
    for f in ("fcbga_step_solve", "fcbga_step_solve_device"):
        getattr(L, f).argtypes = [C.c_void_p, C.POINTER(FcbgaStep), C.POINTER(FcbgaSolveOpts), C.c_void_p,
                                  C.POINTER(FcbgaSolveStats)]
        getattr(L, f).restype = C.c_int
    L.fcbga_step_load.argtypes = [C.c_void_p, C.POINTER(FcbgaStep), _f64p]
    L.fcbga_step_load.restype = C.c_int
    L.fcbga_step_apply.argtypes = [C.c_void_p, C.POINTER(FcbgaStep), _f64p, _f64p]
    L.fcbga_step_apply.restype = C.c_int
    L.fcbga_set_kernel_timing.argtypes = [C.c_void_p, C.c_int]
    L.fcbga_set_kernel_timing.restype = C.c_int
    L.fcbga_kernel_times.argtypes = [C.c_void_p, _f64p, _i32p]
    L.fcbga_kernel_times.restype = C.c_int
    L.fcbga_anand_update.argtypes = [C.POINTER(FcbgaAnandParams), C.c_int64, C.c_double, C.c_double, C.c_double,
                                     _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.fcbga_anand_update.restype = C.c_int
    L.fcbga_anand_update_device.argtypes = [C.POINTER(FcbgaAnandParams), C.c_int64, C.c_double, C.c_double,
                                            C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_void_p, C.c_void_p]
    L.fcbga_anand_update_device.restype = C.c_int
    L.fcbga_state_accumulate.argtypes = [C.c_void_p, C.POINTER(FcbgaStep), _f64p, C.c_int64, _i32p, _f64p,
                                         C.c_int32]
    L.fcbga_state_accumulate.restype = C.c_int
    L.fcbga_state_stress.argtypes = [C.c_void_p, C.c_int64, _i32p, _f64p, C.c_int64, _i32p, _f64p, _f64p, _f64p]
    L.fcbga_state_stress.restype = C.c_int
    L.fcbga_adjoint_forward.argtypes = [C.c_void_p, C.POINTER(FcbgaStep), C.POINTER(FcbgaSolveOpts),
                                        C.POINTER(FcbgaObjective), _f64p, _f64p, C.POINTER(FcbgaSolveStats)]
    L.fcbga_adjoint_forward.restype = C.c_int
    L.fcbga_adjoint_backward.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p, C.POINTER(FcbgaSolveStats)]
    L.fcbga_adjoint_backward.restype = C.c_int
    L.fcbga_screen_joints.argtypes = [C.POINTER(FcbgaAnandParams), C.POINTER(FcbgaScreenSteps), C.c_int64,
                                      C.c_int32, C.c_int32, _f64p, _f64p, _f64p, _f64p]
    L.fcbga_screen_joints.restype = C.c_int
    L.fcbga_screen_joints_noalloc.argtypes = [C.POINTER(FcbgaAnandParams), C.POINTER(FcbgaScreenSteps), C.c_int64,
                                              C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p]
    L.fcbga_screen_joints_noalloc.restype = C.c_int
    L.fcbga_path_reset.argtypes = [C.c_void_p]
    L.fcbga_path_reset.restype = C.c_int
    L.fcbga_path_solve.argtypes = [C.c_void_p, C.c_int32, C.POINTER(FcbgaStep), C.POINTER(FcbgaSolveOpts),
                                   C.POINTER(FcbgaSolveStats)]
    L.fcbga_path_solve.restype = C.c_int
    L.fcbga_path_sum.argtypes = [C.c_void_p, C.c_int32, _i32p, _f64p]
    L.fcbga_path_sum.restype = C.c_int
    L.fcbga_path_du.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_float)]
    L.fcbga_path_du.restype = C.c_int
    L.fcbga_path_load.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_float)]
    L.fcbga_path_load.restype = C.c_int
    L.fcbga_path_load64.argtypes = [C.c_void_p, C.c_int32, _f64p]
    L.fcbga_path_load64.restype = C.c_int
    L.fcbga_path_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
    L.fcbga_path_copy.restype = C.c_int
    L.fcbga_mesh_fingerprint.argtypes = [C.c_int32, C.c_int32, _f64p, C.c_int32, _f64p, C.c_int32, _f64p, C.c_int32,
                                         _i32p, C.POINTER(C.c_int8), _f64p, C.c_char_p]
    L.fcbga_mesh_fingerprint.restype = C.c_int
    L.fcbga_prepare_step.argtypes = [C.c_int32, C.POINTER(FcbgaGroupGrids), C.c_int32, C.POINTER(FcbgaSectionGrids),
                                     C.c_double, C.c_double, C.c_int32, _f64p, _f64p, _u8p, _f64p, _f64p, _u8p]
    L.fcbga_prepare_step.restype = C.c_int
    L.fcbga_adjoint_backward_u.argtypes = [C.c_void_p, _f64p, _f64p, _f64p, _f64p, C.POINTER(FcbgaSolveStats)]
    L.fcbga_adjoint_backward_u.restype = C.c_int
    L.fcbga_warm_reset.argtypes = [C.c_void_p]
    L.fcbga_warm_reset.restype = C.c_int
    L.fcbga_set_progress_cb.argtypes = [C.c_void_p, PROGRESS_FN, C.c_void_p, C.c_int32]
    L.fcbga_set_progress_cb.restype = C.c_int
    L.fcbga_vp_create.argtypes = [C.POINTER(FcbgaMesh), C.POINTER(FcbgaVpSolder), C.POINTER(C.c_void_p)]
    L.fcbga_vp_create.restype = C.c_int
    L.fcbga_vp_destroy.argtypes = [C.c_void_p]
    L.fcbga_vp_destroy.restype = None
    L.fcbga_vp_nsolder.argtypes = [C.c_void_p]
    L.fcbga_vp_nsolder.restype = C.c_int32
    L.fcbga_vp_step.argtypes = [C.c_void_p, C.POINTER(FcbgaStep), C.c_double, C.c_double, C.POINTER(FcbgaVpOpts),
                                C.POINTER(FcbgaVpStats)]
    L.fcbga_vp_step.restype = C.c_int
    for f in ("fcbga_vp_save", "fcbga_vp_restore"):
        getattr(L, f).argtypes = [C.c_void_p]
        getattr(L, f).restype = C.c_int
    L.fcbga_vp_points.argtypes = [C.c_void_p, _f64p, _i32p]
    L.fcbga_vp_points.restype = C.c_int
    L.fcbga_vp_displacement.argtypes = [C.c_void_p, _f64p]
    L.fcbga_vp_displacement.restype = C.c_int
    L.fcbga_vp_set_prescribed.argtypes = [C.c_void_p, _f64p]
    L.fcbga_vp_set_prescribed.restype = C.c_int
    L.fcbga_vp_set_displacement.argtypes = [C.c_void_p, _f64p]
    L.fcbga_vp_set_displacement.restype = C.c_int
    L.fcbga_vp_residual.argtypes = [C.c_void_p, C.POINTER(FcbgaStep), C.c_double, C.c_double, _f64p, _f64p, _f64p]
    L.fcbga_vp_residual.restype = C.c_int
    L.fcbga_last_error.argtypes = [C.c_char_p, C.c_size_t]
    L.fcbga_last_error.restype = C.c_int
    _loaded[variant] = L
    return L


def last_error(L) -> str:
    buf = C.create_string_buffer(1024)
    L.fcbga_last_error(buf, 1024)
    return buf.value.decode()


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class Step:
    """Per-step group data (the prepareStep output, src/core/09-assembly.js:346-365)."""

    def __init__(self, Ta, Tb, stage, D, s0, born, secW=None, secV=None, sec_born=None):
        self.D = np.ascontiguousarray(D, np.float64).reshape(-1)
        self.s0 = np.ascontiguousarray(s0, np.float64).reshape(-1)
        self.born = np.ascontiguousarray(born, np.uint8)
        self.secW = None if secW is None else np.ascontiguousarray(secW, np.float64).reshape(-1)
        self.secV = None if secV is None else np.ascontiguousarray(secV, np.float64).reshape(-1)
        self.sec_born = None if sec_born is None else np.ascontiguousarray(sec_born, np.uint8)
        st = FcbgaStep()
        st.Ta, st.Tb, st.stage = float(Ta), float(Tb), int(stage)
        st.D = _ptr(self.D, _f64p)
        st.s0 = _ptr(self.s0, _f64p)
        st.born = _ptr(self.born, _u8p)
        st.secW = _ptr(self.secW, _f64p)
        st.secV = _ptr(self.secV, _f64p)
        st.sec_born = _ptr(self.sec_born, _u8p)
        self.c = st


def _mesh_struct(mesh):
    """FcbgaMesh over the mesh's arrays; returns (struct, arrays to keep alive)."""
    a = dict(
        xs=np.ascontiguousarray(mesh.xs, np.float64), ys=np.ascontiguousarray(mesh.ys, np.float64),
        zs=np.ascontiguousarray(mesh.zs, np.float64),
        node_id=np.ascontiguousarray(mesh.node_id, np.int32),
        cell_elems=np.ascontiguousarray(mesh.cell_elems, np.int32),
        e_group=np.ascontiguousarray(mesh.e_group, np.int32),
        e_part=np.ascontiguousarray(mesh.e_part, np.int8),
        e_scale=np.ascontiguousarray(mesh.e_scale, np.float64),
        cons=np.ascontiguousarray(mesh.constraint_dofs, np.int32),
        fx=np.ascontiguousarray(mesh.feature_x, np.int32), fy=np.ascontiguousarray(mesh.feature_y, np.int32))
    nsec = int(getattr(mesh, "nsections", 0) or 0)
    if nsec > 0:
        a["e_section"] = np.ascontiguousarray(mesh.e_section, np.int32)
    m = FcbgaMesh()
    m.nx, m.ny, m.nz, m.nn, m.ne = mesh.nx, mesh.ny, mesh.nz, mesh.nn, mesh.ne
    m.xs, m.ys, m.zs = _ptr(a["xs"], _f64p), _ptr(a["ys"], _f64p), _ptr(a["zs"], _f64p)
    m.node_id = _ptr(a["node_id"], _i32p)
    m.cell_elems = _ptr(a["cell_elems"], _i32p)
    m.e_group = _ptr(a["e_group"], _i32p)
    m.e_part = _ptr(a["e_part"], _i8p)
    m.e_scale = _ptr(a["e_scale"], _f64p)
    m.e_section = _ptr(a["e_section"], _i32p) if nsec > 0 else None
    m.ngroups = len(mesh.groups)
    m.nsections = nsec
    m.nconstraints = len(a["cons"])
    m.constraint_dofs = _ptr(a["cons"], _i32p)
    m.nfeature_x, m.nfeature_y = len(a["fx"]), len(a["fy"])
    m.feature_x, m.feature_y = _ptr(a["fx"], _i32p), _ptr(a["fy"], _i32p)
    mod = getattr(mesh, "module_constraint_dofs", None)
    if mod is not None and len(mod):
        a["mod"] = np.ascontiguousarray(mod, np.int32)
        m.nmodule_constraints = len(a["mod"])
        m.module_constraint_dofs = _ptr(a["mod"], _i32p)
    return m, a


class Model:
    def __init__(self, mesh, variant: str = "hip"):
        self.variant = variant
        self.L = load(variant)
        self.mesh = mesh
        m, self._keep = _mesh_struct(mesh)
        self._cmesh = m
        h = C.c_void_p()
        rc = self.L.fcbga_model_create(C.byref(m), C.byref(h))
        if rc != 0:
            raise SolverError(f"fcbga_model_create failed ({rc}): {last_error(self.L)}")
        self.h = h
        self.ndof = int(self.L.fcbga_model_ndof(h))

    def close(self):
        if getattr(self, "h", None):
            self.L.fcbga_model_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @staticmethod
    def _opts(tol, maxit, precision, precond, sweeps=0, warm=0, smoother=0):
        o = FcbgaSolveOpts()
        o.tol, o.maxit, o.precision, o.precond, o.sweeps, o.warm = tol, maxit, precision, precond, sweeps, warm
        o.smoother = smoother
        return o

    def warm_reset(self):
        """Empty the warm-start basis."""
        self.L.fcbga_warm_reset(self.h)

    def solve_step(self, step: Step, tol=1e-8, maxit=2000, precision=PREC_FP64, precond=PC_MG, check=True, sweeps=0,
                   warm=0, smoother=SMOOTH_GS):
        """du (3*nn, reference node order) and the solve record; warm > 0
        projects the load onto (and keeps) up to `warm` previous solutions;
        smoother: SMOOTH_GS (the reference's) or SMOOTH_COARSE_JACOBI."""
        du = np.zeros(self.ndof)
        st = FcbgaSolveStats()
        o = self._opts(tol, maxit, precision, precond, sweeps, warm, smoother)
        rc = self.L.fcbga_step_solve(self.h, C.byref(step.c), C.byref(o), du.ctypes.data_as(C.c_void_p),
                                     C.byref(st))
        rec = _record(st)
        if rc != 0 and (check or rc != -3):
            raise SolverError(f"fcbga_step_solve failed ({rc}): {last_error(self.L)}")
        return du, rec

    def set_progress(self, fn=None, every=1):
        """Live progress from inside the PCG: fn(record dict) on the calling
        thread every `every` iterations; a true return stops the solve, which
        then raises SolverError('cancelled').  fn None removes it."""
        if fn is None:
            self._progress = None
            self.L.fcbga_set_progress_cb(self.h, PROGRESS_FN(), None, 1)
            return

        def tramp(p, _user):
            rec = {k: getattr(p.contents, k) for k, _ in FcbgaProgress._fields_}
            return 1 if fn(rec) else 0

        self._progress = PROGRESS_FN(tramp)  # kept alive while installed
        if self.L.fcbga_set_progress_cb(self.h, self._progress, None, int(every)) != 0:
            raise SolverError(last_error(self.L))

    def solve_step_device(self, step: Step, du_dev_ptr: int, tol=1e-8, maxit=2000, precision=PREC_FP64,
                          precond=PC_MG, sweeps=0, warm=0, smoother=SMOOTH_GS):
        st = FcbgaSolveStats()
        o = self._opts(tol, maxit, precision, precond, sweeps, warm, smoother)
        rc = self.L.fcbga_step_solve_device(self.h, C.byref(step.c), C.byref(o), C.c_void_p(du_dev_ptr),
                                            C.byref(st))
        if rc != 0:
            raise SolverError(f"fcbga_step_solve_device failed ({rc}): {last_error(self.L)}")
        return _record(st)

    def path_reset(self):
        self.L.fcbga_path_reset(self.h)

    def path_solve(self, k, step: Step, tol=1e-8, maxit=2000, precision=PREC_MIXED, precond=PC_MG, warm=20,
                   smoother=SMOOTH_GS):
        """Solve step k of a thermal path, keeping its increment on the device."""
        st = FcbgaSolveStats()
        o = self._opts(tol, maxit, precision, precond, 0, warm, smoother)
        rc = self.L.fcbga_path_solve(self.h, int(k), C.byref(step.c), C.byref(o), C.byref(st))
        rec = _record(st)
        if rc != 0 and rc != -3:
            raise SolverError(f"fcbga_path_solve failed ({rc}): {last_error(self.L)}")
        return rec

    def path_sum(self, ids):
        """Sum of the kept increments of steps ids, added in that order (host FP64)."""
        ids = np.ascontiguousarray(ids, np.int32)
        out = np.zeros(self.ndof)
        rc = self.L.fcbga_path_sum(self.h, len(ids), ids.ctypes.data_as(_i32p), out.ctypes.data_as(_f64p))
        if rc != 0:
            raise SolverError(last_error(self.L))
        return out

    def path_du(self, k):
        """FP32 increment of step k (as the reference keeps it after finishPath)."""
        out = np.zeros(self.ndof, np.float32)
        rc = self.L.fcbga_path_du(self.h, int(k), out.ctypes.data_as(C.POINTER(C.c_float)))
        if rc != 0:
            raise SolverError(last_error(self.L))
        return out

    def path_load(self, k, du):
        """Keep a given increment as step k: FP32 from a results file
        (Analysis.restorePath) or FP64 from another model's solve (the partial
        {k, du} message of the worker farm), by the array's type."""
        du = np.asarray(du)
        if du.shape != (self.ndof,):
            raise ValueError(f"increment of {du.shape} values for a model of {self.ndof} DOF")
        if du.dtype == np.float64:
            du = np.ascontiguousarray(du)
            rc = self.L.fcbga_path_load64(self.h, int(k), du.ctypes.data_as(_f64p))
        else:
            du = np.ascontiguousarray(du, np.float32)
            rc = self.L.fcbga_path_load(self.h, int(k), du.ctypes.data_as(C.POINTER(C.c_float)))
        if rc != 0:
            raise SolverError(last_error(self.L))

    def path_copy_from(self, other, k):
        """Keep step k's increment as kept by another model of the same mesh
        (device to device)."""
        rc = self.L.fcbga_path_copy(self.h, other.h, int(k))
        if rc != 0:
            raise SolverError(last_error(self.L))

    def set_kernel_timing(self, enable: bool):
        self.L.fcbga_set_kernel_timing(self.h, 1 if enable else 0)

    def kernel_times(self):
        """{class: (total ms, launches)} accumulated since set_kernel_timing."""
        ms = np.zeros(KCLASS_COUNT)
        n = np.zeros(KCLASS_COUNT, np.int32)
        self.L.fcbga_kernel_times(self.h, ms.ctypes.data_as(_f64p), n.ctypes.data_as(_i32p))
        return {k: (float(ms[v]), int(n[v])) for k, v in KCLASS.items()}

    def load(self, step: Step):
        f = np.zeros(self.ndof)
        rc = self.L.fcbga_step_load(self.h, C.byref(step.c), f.ctypes.data_as(_f64p))
        if rc != 0:
            raise SolverError(last_error(self.L))
        return f

    def state_accumulate(self, step: Step, du, elems, states=None, q_only=False):
        """Add one step (increment du, 3*nn) to the stress-history states of
        elements `elems` (accumulateStep, src/core/09-assembly.js:494-545);
        states (len(elems) x STATE) start at zero when not given."""
        elems = np.ascontiguousarray(elems, np.int32)
        du = np.ascontiguousarray(du, np.float64)
        if du.shape != (self.ndof,):
            raise SolverError("du must have 3*nn entries")
        st = np.zeros((len(elems), STATE)) if states is None else np.ascontiguousarray(states, np.float64).copy()
        rc = self.L.fcbga_state_accumulate(self.h, C.byref(step.c), du.ctypes.data_as(_f64p), len(elems),
                                           elems.ctypes.data_as(_i32p), st.ctypes.data_as(_f64p),
                                           1 if q_only else 0)
        if rc != 0:
            raise SolverError(last_error(self.L))
        return st

    def state_stress(self, elems, states, row, xez):
        """Stress and total strain (npts x 6 each) at natural points xez
        (npts x 3) of elements elems[row] with states[row] (stateStress,
        src/core/09-assembly.js:556-571)."""
        elems = np.ascontiguousarray(elems, np.int32)
        states = np.ascontiguousarray(states, np.float64)
        row = np.ascontiguousarray(row, np.int32)
        xez = np.ascontiguousarray(xez, np.float64).reshape(-1, 3)
        if states.shape != (len(elems), STATE) or len(row) != len(xez):
            raise SolverError("state_stress: shape mismatch")
        sig = np.zeros((len(row), 6))
        eps = np.zeros((len(row), 6))
        rc = self.L.fcbga_state_stress(self.h, len(elems), elems.ctypes.data_as(_i32p),
                                       states.ctypes.data_as(_f64p), len(row), row.ctypes.data_as(_i32p),
                                       xez.ctypes.data_as(_f64p), sig.ctypes.data_as(_f64p),
                                       eps.ctypes.data_as(_f64p))
        if rc != 0:
            raise SolverError(last_error(self.L))
        return sig, eps

    def adjoint_forward(self, step: Step, k_top, vm_elems=(), p=8.0, tol=1e-10, maxit=2000, precision=PREC_MIXED,
                        precond=PC_MG, check=True):
        """Forward of the sensitivity pair: solve the step and evaluate J =
        [warpage proxy on plane k_top, von Mises p-norm over vm_elems]; returns
        (J (2), du, solve record).  The solution stays on the model for
        adjoint_backward."""
        vm = np.ascontiguousarray(vm_elems, np.int32)
        obj = FcbgaObjective()
        obj.k_top, obj.n_vm, obj.p = int(k_top), len(vm), float(p)
        obj.vm_elems = vm.ctypes.data_as(_i32p) if len(vm) else None
        J = np.zeros(2)
        du = np.zeros(self.ndof)
        st = FcbgaSolveStats()
        o = self._opts(tol, maxit, precision, precond)
        rc = self.L.fcbga_adjoint_forward(self.h, C.byref(step.c), C.byref(o), C.byref(obj), J.ctypes.data_as(_f64p),
                                          du.ctypes.data_as(_f64p), C.byref(st))
        if rc != 0 and (check or rc != -3):
            raise SolverError(f"fcbga_adjoint_forward failed ({rc}): {last_error(self.L)}")
        return J, du, _record(st)

    def adjoint_backward(self, gJ, elem_cte=True):
        """Backward of the pair: log-derivatives of gJ . J with respect to each
        group's E and CTE and (elem_cte) each element's CTE, from one adjoint
        solve; returns (dlnE, dlnA, dlnA_elem or None, solve record)."""
        g = np.ascontiguousarray(gJ, np.float64).reshape(2)
        ng = int(self.mesh.e_group.max()) + 1 if len(self.mesh.groups) == 0 else len(self.mesh.groups)
        dE = np.zeros(ng)
        dA = np.zeros(ng)
        de = np.zeros(self.mesh.ne) if elem_cte else None
        st = FcbgaSolveStats()
        rc = self.L.fcbga_adjoint_backward(self.h, g.ctypes.data_as(_f64p), dE.ctypes.data_as(_f64p),
                                           dA.ctypes.data_as(_f64p),
                                           de.ctypes.data_as(_f64p) if de is not None else None, C.byref(st))
        if rc != 0:
            raise SolverError(f"fcbga_adjoint_backward failed ({rc}): {last_error(self.L)}")
        return dE, dA, de, _record(st)

    def adjoint_backward_u(self, dJdu, elem_cte=False):
        """Backward of the pair for the caller's own objective: dJdu (3 nn) is
        dJ/du of the forward's step solution; returns (dlnE, dlnA, dlnA_elem or
        None, solve record) as adjoint_backward."""
        g = np.ascontiguousarray(dJdu, np.float64).reshape(-1)
        if len(g) != self.ndof:
            raise SolverError(f"dJdu has {len(g)} values, the model {self.ndof} DOFs")
        ng = int(self.mesh.e_group.max()) + 1 if len(self.mesh.groups) == 0 else len(self.mesh.groups)
        dE, dA = np.zeros(ng), np.zeros(ng)
        de = np.zeros(self.mesh.ne) if elem_cte else None
        st = FcbgaSolveStats()
        rc = self.L.fcbga_adjoint_backward_u(self.h, g.ctypes.data_as(_f64p), dE.ctypes.data_as(_f64p),
                                             dA.ctypes.data_as(_f64p),
                                             de.ctypes.data_as(_f64p) if de is not None else None, C.byref(st))
        if rc != 0:
            raise SolverError(f"fcbga_adjoint_backward_u failed ({rc}): {last_error(self.L)}")
        return dE, dA, de, _record(st)

    def apply(self, step: Step, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.ndof)
        rc = self.L.fcbga_step_apply(self.h, C.byref(step.c), x.ctypes.data_as(_f64p), y.ctypes.data_as(_f64p))
        if rc != 0:
            raise SolverError(last_error(self.L))
        return y


TGRID_MIN, TGRID_N = -80, 381  # FCBGA_TGRID_MIN, FCBGA_TGRID_N


def prepare_step(groups, sections, Ta, Tb, stage, variant="hip"):
    """Step data of [Ta, Tb] in `stage` from 1 degC property grids (fcbga_prepare_step:
    FEModel.prepareStep, src/core/09-assembly.js:346-365).  groups: dicts with stage,
    layered and, for the others, Dgrid (TGRID_N x 36) or Dconst (36) and F (TGRID_N x 3);
    sections: dicts with ns, group (-1: always born), W (TGRID_N x 108), D (ns x TGRID_N x 36),
    F (ns x TGRID_N x 3) and Ls (ns x 3).  Returns a Step."""
    L = load(variant)
    keep = []

    def arr(a, n):
        if a is None:
            return None
        a = np.ascontiguousarray(a, np.float64).reshape(-1)
        if len(a) != n:
            raise SolverError(f"grid of {len(a)} values where {n} are expected")
        keep.append(a)
        return _ptr(a, _f64p)
    ng, ns = len(groups), len(sections)
    G = (FcbgaGroupGrids * max(ng, 1))()
    for q, g in enumerate(groups):
        G[q].stage, G[q].layered = int(g["stage"]), 1 if g.get("layered") else 0
        G[q].Dgrid = arr(g.get("Dgrid"), TGRID_N * 36)
        G[q].Dconst = arr(g.get("Dconst"), 36)
        G[q].F = arr(g.get("F"), TGRID_N * 3)
    S = (FcbgaSectionGrids * max(ns, 1))()
    for q, c in enumerate(sections):
        n = int(c["ns"])
        S[q].ns, S[q].group = n, int(c["group"])
        S[q].W = arr(c["W"], TGRID_N * 108)
        S[q].D, S[q].F, S[q].Ls = arr(c["D"], n * TGRID_N * 36), arr(c["F"], n * TGRID_N * 3), arr(c["Ls"], n * 3)
    D, s0, born = np.zeros((ng, 36)), np.zeros((ng, 6)), np.zeros(ng, np.uint8)
    W, V, sb = np.zeros((ns, 108)), np.zeros((ns, 18)), np.zeros(ns, np.uint8)
    rc = L.fcbga_prepare_step(ng, G, ns, S, float(Ta), float(Tb), int(stage), _ptr(D, _f64p), _ptr(s0, _f64p),
                              _ptr(born, _u8p), _ptr(W, _f64p), _ptr(V, _f64p), _ptr(sb, _u8p))
    if rc != 0:
        raise SolverError(last_error(L))
    if ns:
        return Step(Ta, Tb, stage, D, s0, born, W, V, sb)
    return Step(Ta, Tb, stage, D, s0, born)


def step_for_spec(spec, Ta=None, Tb=None, stage=None):
    from .meshspec import group_step_data
    Ta = spec["Ta"] if Ta is None else Ta
    Tb = spec["Tb"] if Tb is None else Tb
    stage = spec.get("stage", 0) if stage is None else stage
    D, s0, born = group_step_data(spec, Ta, Tb, stage)
    return Step(Ta, Tb, stage, D, s0, born)


def anand_params(A, QR, xi, m, shat, n, h0, a, s0, nu):
    p = FcbgaAnandParams()
    for k, v in dict(A=A, QR=QR, xi=xi, m=m, shat=shat, n=n, h0=h0, a=a, s0=s0, nu=nu).items():
        setattr(p, k, float(v))
    return p


def anand_initial_state(params, npts):
    """Committed state of fresh points: zero plastic strain, s = s0 (AnandPoint.reset)."""
    st = np.zeros((npts, ANAND_STATE))
    st[:, 6] = params.s0
    return st


def anand_update(params, eps, T, dt, E, state, tangent=True, variant="hip"):
    """Batched AnandPoint.step: returns (trial state, sig, D or None, dep); the
    warm-start entry state[:, 10] is updated in place."""
    L = load(variant)
    eps = np.ascontiguousarray(eps, np.float64).reshape(-1, 6)
    n = eps.shape[0]
    assert state.shape == (n, ANAND_STATE) and state.dtype == np.float64 and state.flags.c_contiguous
    trial = np.zeros((n, ANAND_STATE))
    sig = np.zeros((n, 6))
    D = np.zeros((n, 36)) if tangent else None
    dep = np.zeros(n)
    rc = L.fcbga_anand_update(C.byref(params), n, float(T), float(dt), float(E), eps.ctypes.data_as(_f64p),
                              state.ctypes.data_as(_f64p), trial.ctypes.data_as(_f64p), sig.ctypes.data_as(_f64p),
                              D.ctypes.data_as(_f64p) if D is not None else None, dep.ctypes.data_as(_f64p))
    if rc != 0:
        raise SolverError(last_error(L))
    return trial, sig, D, dep


SCREEN_NOALLOC_MAX_STEPS = 4096


def _screen_steps_struct(steps, keep):
    def ptr(k, dt, t):
        v = steps[k]
        if v is None:
            return None
        keep[k] = np.ascontiguousarray(v, dt)
        return keep[k].ctypes.data_as(t)

    h = FcbgaScreenSteps()
    h.nsteps, h.cycles, h.fs = int(steps["nsteps"]), int(steps["cycles"]), float(steps["fs"])
    for k in ("T", "dt", "E", "Fth", "wt", "Cp", "Fp"):
        setattr(h, k, ptr(k, np.float64, _f64p))
    h.wi = ptr("wi", np.int32, _i32p)
    h.cycle = ptr("cycle", np.int32, _i32p)
    h.same = ptr("same", np.uint8, C.POINTER(C.c_uint8))
    return h


def screen_joints_noalloc(params, steps, nsites, ntemps, i25, table_ptr, dW_ptr, dE_ptr, stab_ptr, stream=None,
                          variant="hip"):
    """fcbga_screen_joints_noalloc: the screening loop with the strain table
    and the outputs already in device memory (addresses as ints; host
    addresses for the host-emulation build) and no device allocation; the
    history arrays of steps are staged in the library's static scratch."""
    L = load(variant)
    keep = {}
    h = _screen_steps_struct(steps, keep)
    rc = L.fcbga_screen_joints_noalloc(C.byref(params), C.byref(h), int(nsites), int(ntemps), int(i25),
                                       C.c_void_p(table_ptr), C.c_void_p(dW_ptr), C.c_void_p(dE_ptr),
                                       C.c_void_p(stab_ptr), C.c_void_p(stream) if stream else None)
    if rc != 0:
        raise SolverError(last_error(L))


def screen_joints(params, steps, table, i25, variant="hip"):
    """Bump screening loop on the device (fcbga_screen_joints): params the
    solder's FcbgaAnandParams, steps the per-step arrays of
    screening.step_data, table (nsites, ntemps, 6) bump-layer strains, i25 the
    row of 25 C (-1: none).  Returns per-site dW, dE, stab."""
    L = load(variant)
    tab = np.ascontiguousarray(table, np.float64)
    nsites, ntemps = tab.shape[0], tab.shape[1]
    keep = {}
    h = _screen_steps_struct(steps, keep)
    dW, dE, stab = np.zeros(nsites), np.zeros(nsites), np.zeros(nsites)
    rc = L.fcbga_screen_joints(C.byref(params), C.byref(h), nsites, ntemps, int(i25), tab.ctypes.data_as(_f64p),
                               dW.ctypes.data_as(_f64p), dE.ctypes.data_as(_f64p), stab.ctypes.data_as(_f64p))
    if rc != 0:
        raise SolverError(last_error(L))
    return dW, dE, stab


class ViscoplasticModel:
    """fcbga_vp_*: the global viscoplastic Newton over a mesh whose solder
    group carries Anand points (the bump-array cycling workload).  solder:
    dict with group, anand (A QR xi m shat n h0 a s0), nu, E_table {T, v},
    cte (ppm/K), Tref."""

    def __init__(self, mesh, solder: dict, variant: str = "hip"):
        self.L = load(variant)
        self.mesh = mesh
        m, self._keep = _mesh_struct(mesh)
        self._cmesh = m
        sp = FcbgaVpSolder()
        sp.group = int(solder["group"])
        for k, v in solder["anand"].items():
            setattr(sp.anand, k, float(v))
        sp.anand.nu = float(solder["nu"])
        self._ET = np.ascontiguousarray(solder["E_table"]["T"], np.float64)
        self._Ev = np.ascontiguousarray(solder["E_table"]["v"], np.float64)
        sp.nE = len(self._ET)
        sp.E_T, sp.E_v = _ptr(self._ET, _f64p), _ptr(self._Ev, _f64p)
        sp.cte = float(solder["cte"]) * 1e-6
        sp.Tref = float(solder["Tref"])
        self._sp = sp
        h = C.c_void_p()
        rc = self.L.fcbga_vp_create(C.byref(m), C.byref(sp), C.byref(h))
        if rc != 0:
            raise SolverError(f"fcbga_vp_create failed ({rc}): {last_error(self.L)}")
        self.h = h
        self.nsolder = int(self.L.fcbga_vp_nsolder(h))
        self.ndof = 3 * mesh.nn

    def close(self):
        if getattr(self, "h", None):
            self.L.fcbga_vp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, elastic: Step, T, dt, precision=PREC_MIXED, maxit=25, res_tol=1e-6, du_tol=1e-8,
             inner_floor=1e-8, inner_maxit=2000):
        o = FcbgaVpOpts()
        o.inner.tol, o.inner.maxit, o.inner.precision, o.inner.precond = 1e-8, int(inner_maxit), int(precision), PC_MG
        o.inner.sweeps, o.inner.warm = 0, 0
        o.maxit, o.res_tol, o.du_tol, o.inner_floor = int(maxit), float(res_tol), float(du_tol), float(inner_floor)
        st = FcbgaVpStats()
        rc = self.L.fcbga_vp_step(self.h, C.byref(elastic.c), float(T), float(dt), C.byref(o), C.byref(st))
        rec = {k: getattr(st, k) for k, _ in FcbgaVpStats._fields_}
        rec["converged"] = bool(rec["converged"])
        if rc not in (0, -3):
            raise SolverError(f"fcbga_vp_step failed ({rc}): {last_error(self.L)}")
        return rec

    def save(self):
        if self.L.fcbga_vp_save(self.h) != 0:
            raise SolverError(last_error(self.L))

    def restore(self):
        if self.L.fcbga_vp_restore(self.h) != 0:
            raise SolverError(last_error(self.L))

    def points(self):
        """Committed point states (nsolder, 8, 11) and the mesh element of each solder element."""
        st = np.empty(88 * self.nsolder)
        el = np.empty(self.nsolder, np.int32)
        if self.L.fcbga_vp_points(self.h, st.ctypes.data_as(_f64p), el.ctypes.data_as(_i32p)) != 0:
            raise SolverError(last_error(self.L))
        return st.reshape(self.nsolder, 8, ANAND_STATE), el

    def displacement(self):
        u = np.empty(self.ndof)
        if self.L.fcbga_vp_displacement(self.h, u.ctypes.data_as(_f64p)) != 0:
            raise SolverError(last_error(self.L))
        return u

    def set_prescribed(self, values):
        """Prescribed displacement of the constrained DOFs (mesh.constraint_dofs order) for later steps."""
        v = np.ascontiguousarray(values, np.float64)
        assert v.shape == (len(self.mesh.constraint_dofs),)
        if self.L.fcbga_vp_set_prescribed(self.h, v.ctypes.data_as(_f64p)) != 0:
            raise SolverError(last_error(self.L))

    def set_displacement(self, u):
        u = np.ascontiguousarray(u, np.float64)
        if self.L.fcbga_vp_set_displacement(self.h, u.ctypes.data_as(_f64p)) != 0:
            raise SolverError(last_error(self.L))

    def residual(self, elastic: Step, T, dt, x=None):
        """Out-of-balance force at the current displacement and, for x, K_t x."""
        r = np.empty(self.ndof)
        y = np.empty(self.ndof) if x is not None else None
        xx = None if x is None else np.ascontiguousarray(x, np.float64)
        rc = self.L.fcbga_vp_residual(self.h, C.byref(elastic.c), float(T), float(dt), r.ctypes.data_as(_f64p),
                                      _ptr(xx, _f64p), _ptr(y, _f64p))
        if rc != 0:
            raise SolverError(last_error(self.L))
        return (r, y) if x is not None else r
