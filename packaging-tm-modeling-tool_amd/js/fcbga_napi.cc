// fcbga_napi.cc -- Node-API addon over the C ABI of include/fcbga_hip.h.
//
// JS surface (the reference's step-solve seam, src/core/11-analysis.js:189-206):
//   const hip = require('fcbga_hip.node');
//   const model = hip.createModel(mesh);        // FEModel's mesh contract (src/core/06-mesher.js:66-71)
//   const rec = hip.solveStep(model, step, opts);
//     step: {Ta, Tb, stage, D: Float64Array(ng*36), s0: Float64Array(ng*6), born: Uint8Array(ng)}
//     opts: {tol, maxit, precision: 'mixed'|'fp64', precond: 'mg'|'ic'|'column'|'jacobi', sweeps, warm}
//     rec:  {u: Float64Array(ndof), iters, relres, converged, ms, assembleMs, pcMs, pcgMs, levels,
//            precond, rebuilt, warm, fallback, icShift}  -- the fields of LinearSolver.solve's record (src/core/10-solver.js:988-993)
//   const st = hip.stateAccumulate(model, step, du, elems[, states[, qOnly]]);  // Float64Array(n * 81)
//   const {sig, eps} = hip.stateStress(model, elems, st, rows, xez);       // stresses at natural points
//   const f = hip.adjointForward(model, step, {kTop, vmElems, p}, opts);   // {J, u, ...}: saves the context
//   const g = hip.adjointBackward(model, Float64Array.of(1, 0));          // {dlnE, dlnCte, dlnCteElem}
//   const s = hip.screenJoints(solder, steps, table, i25);                 // {dW, dE, stab} per joint
//   const h = hip.submodelCreate(mesh, solder); hip.submodelCall(h, 'step', data); ...  // bump submodel engine
//   hip.destroyModel(model);                   // or left to the garbage collector
// Argument and device errors throw Error(message); a solve that does not
// converge returns its record with converged = false (the caller throws, as
// Analysis.solveStep does).  The addon is context-aware (napi_register_module_v1) and
// keeps no global state: it can load in the compute worker's worker_thread.
#include <node_api.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/fcbga_hip.h"

namespace {

#define NAPI_OK(call)                                          \
  do {                                                         \
    if ((call) != napi_ok) {                                   \
      napi_throw_error(env, nullptr, "N-API call failed: " #call); \
      return nullptr;                                          \
    }                                                          \
  } while (0)

constexpr uint32_t MODEL_TAG = 0x4d4f444c, SUBMODEL_TAG = 0x5355424d;

struct ModelBox {
  uint32_t tag = MODEL_TAG;
  fcbga_model *m = nullptr;
  int32_t ngroups = 0, ne = 0, nsections = 0;
};

struct SubBox {
  uint32_t tag = SUBMODEL_TAG;
  fcbga_submodel *s = nullptr;
  int32_t nb = 0;
  int64_t ne = 0, nn = 0;
};

void finalize_submodel(napi_env, void *data, void *) {
  SubBox *b = static_cast<SubBox *>(data);
  if (b->s) fcbga_submodel_destroy(b->s);
  delete b;
}

void finalize_model(napi_env, void *data, void *) {
  ModelBox *b = static_cast<ModelBox *>(data);
  if (b->m) fcbga_model_destroy(b->m);
  delete b;
}

bool get_prop(napi_env env, napi_value obj, const char *name, napi_value *out) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return false;
  if (napi_get_named_property(env, obj, name, out) != napi_ok) return false;
  napi_valuetype t;
  napi_typeof(env, *out, &t);
  return t != napi_undefined && t != napi_null;
}

bool get_int(napi_env env, napi_value obj, const char *name, int32_t *v) {
  napi_value x;
  if (!get_prop(env, obj, name, &x)) return false;
  return napi_get_value_int32(env, x, v) == napi_ok;
}

bool get_double(napi_env env, napi_value obj, const char *name, double *v) {
  napi_value x;
  if (!get_prop(env, obj, name, &x)) return false;
  return napi_get_value_double(env, x, v) == napi_ok;
}

// typed array of the expected element type; returns its data and length
bool get_typed(napi_env env, napi_value obj, const char *name, napi_typedarray_type want, void **data, size_t *len) {
  napi_value x;
  if (!get_prop(env, obj, name, &x)) return false;
  bool is = false;
  napi_is_typedarray(env, x, &is);
  if (!is) return false;
  napi_typedarray_type t;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, x, &t, len, data, &ab, &off) != napi_ok) return false;
  return t == want;
}

std::string lib_error() {
  char buf[512];
  fcbga_last_error(buf, sizeof buf);
  return buf;
}

napi_value throw_msg(napi_env env, const std::string &m) {
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

napi_value CreateModel(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 1) return throw_msg(env, "createModel(mesh): mesh missing");
  napi_value mesh = argv[0];
  fcbga_mesh m;
  memset(&m, 0, sizeof m);
  size_t n;
  void *p;
  if (!get_int(env, mesh, "nx", &m.nx) || !get_int(env, mesh, "ny", &m.ny) || !get_int(env, mesh, "nz", &m.nz) ||
      !get_int(env, mesh, "nn", &m.nn) || !get_int(env, mesh, "ne", &m.ne))
    return throw_msg(env, "mesh: nx, ny, nz, nn, ne required");
  if (!get_typed(env, mesh, "xs", napi_float64_array, &p, &n) || (int)n != m.nx) return throw_msg(env, "mesh.xs: Float64Array(nx)");
  m.xs = (const double *)p;
  if (!get_typed(env, mesh, "ys", napi_float64_array, &p, &n) || (int)n != m.ny) return throw_msg(env, "mesh.ys: Float64Array(ny)");
  m.ys = (const double *)p;
  if (!get_typed(env, mesh, "zs", napi_float64_array, &p, &n) || (int)n != m.nz) return throw_msg(env, "mesh.zs: Float64Array(nz)");
  m.zs = (const double *)p;
  const size_t G = (size_t)m.nx * m.ny * m.nz, ncells = (size_t)(m.nx - 1) * (m.ny - 1) * (m.nz - 1);
  if (!get_typed(env, mesh, "nodeId", napi_int32_array, &p, &n) || n != G) return throw_msg(env, "mesh.nodeId: Int32Array(nx*ny*nz)");
  m.node_id = (const int32_t *)p;
  if (!get_typed(env, mesh, "cellElems", napi_int32_array, &p, &n) || n != 2 * ncells)
    return throw_msg(env, "mesh.cellElems: Int32Array(2*ncells)");
  m.cell_elems = (const int32_t *)p;
  if (!get_typed(env, mesh, "eGroup", napi_int32_array, &p, &n) || (int)n != m.ne) return throw_msg(env, "mesh.eGroup: Int32Array(ne)");
  m.e_group = (const int32_t *)p;
  if (!get_typed(env, mesh, "ePart", napi_int8_array, &p, &n) || (int)n != m.ne) return throw_msg(env, "mesh.ePart: Int8Array(ne)");
  m.e_part = (const int8_t *)p;
  if (!get_typed(env, mesh, "eScale", napi_float64_array, &p, &n) || (int)n != m.ne) return throw_msg(env, "mesh.eScale: Float64Array(ne)");
  m.e_scale = (const double *)p;
  napi_value groups;
  uint32_t ng = 0;
  if (get_prop(env, mesh, "groups", &groups)) napi_get_array_length(env, groups, &ng);
  int32_t ngi = 0;
  if (get_int(env, mesh, "ngroups", &ngi)) ng = (uint32_t)ngi;
  m.ngroups = (int32_t)ng;
  napi_value cons;
  if (get_prop(env, mesh, "constraints", &cons)) {
    if (!get_typed(env, cons, "dofs", napi_int32_array, &p, &n)) return throw_msg(env, "mesh.constraints.dofs: Int32Array");
    m.constraint_dofs = (const int32_t *)p;
    m.nconstraints = (int32_t)n;
  }
  // stacked-module stages: the FEModel's moduleConstraints (src/core/09-assembly.js:227)
  if (get_typed(env, mesh, "moduleConstraints", napi_int32_array, &p, &n) && n > 0) {
    m.module_constraint_dofs = (const int32_t *)p;
    m.nmodule_constraints = (int32_t)n;
  }
  // layered sections (FEModel: mesh.sections from substrateBands, src/core/09-assembly.js:217-219, 309):
  // eSection (-1 homogeneous) and the number of sections
  int32_t nsec = 0;
  if (get_int(env, mesh, "nsections", &nsec) && nsec > 0) {
    if (!get_typed(env, mesh, "eSection", napi_int32_array, &p, &n) || (int)n != m.ne)
      return throw_msg(env, "mesh.eSection: Int32Array(ne) with nsections > 0");
    m.e_section = (const int32_t *)p;
    m.nsections = nsec;
  }
  if (get_typed(env, mesh, "featureX", napi_int32_array, &p, &n)) { m.feature_x = (const int32_t *)p; m.nfeature_x = (int32_t)n; }
  if (get_typed(env, mesh, "featureY", napi_int32_array, &p, &n)) { m.feature_y = (const int32_t *)p; m.nfeature_y = (int32_t)n; }
  fcbga_model *model = nullptr;
  if (fcbga_model_create(&m, &model) != FCBGA_OK) return throw_msg(env, "createModel: " + lib_error());
  ModelBox *box = new ModelBox();
  box->m = model;
  box->ngroups = m.ngroups;
  box->ne = m.ne;
  box->nsections = m.nsections;
  napi_value ext;
  if (napi_create_external(env, box, finalize_model, nullptr, &ext) != napi_ok) {
    fcbga_model_destroy(model);
    delete box;
    return throw_msg(env, "createModel: could not wrap the model");
  }
  return ext;
}

ModelBox *unwrap(napi_env env, napi_value v) {
  void *data = nullptr;
  if (napi_get_value_external(env, v, &data) != napi_ok || !data) return nullptr;
  ModelBox *b = static_cast<ModelBox *>(data);
  return b->tag == MODEL_TAG ? b : nullptr;
}

SubBox *unwrap_sub(napi_env env, napi_value v) {
  void *data = nullptr;
  if (napi_get_value_external(env, v, &data) != napi_ok || !data) return nullptr;
  SubBox *b = static_cast<SubBox *>(data);
  return b->tag == SUBMODEL_TAG && b->s ? b : nullptr;
}

napi_value DestroyModel(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  ModelBox *b = argc ? unwrap(env, argv[0]) : nullptr;
  if (!b) return throw_msg(env, "destroyModel(model): not a model");
  if (b->m) fcbga_model_destroy(b->m);
  b->m = nullptr;
  return nullptr;
}

void set_num(napi_env env, napi_value obj, const char *k, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, obj, k, x);
}

const char *pc_name(int pc) {
  switch (pc) {
    case FCBGA_PC_MG: return "mg";
    case FCBGA_PC_COLUMN: return "column";
    case FCBGA_PC_JACOBI: return "jacobi";
    case FCBGA_PC_IC: return "ic";
  }
  return "?";
}

// opts {tol, maxit, sweeps, warm, precision: 'mixed'|'fp64', precond: 'mg'|'ic'|'column'|'jacobi',
//       smoother: 'gs'|'coarse-jacobi'}; false for an unknown precond or smoother name
bool read_opts(napi_env env, napi_value v, fcbga_solve_opts *o) {
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t != napi_object) return true;
  get_double(env, v, "tol", &o->tol);
  get_int(env, v, "maxit", &o->maxit);
  get_int(env, v, "sweeps", &o->sweeps);
  get_int(env, v, "warm", &o->warm);
  napi_value x;
  char buf[32];
  size_t len;
  if (get_prop(env, v, "precision", &x) && napi_get_value_string_utf8(env, x, buf, sizeof buf, &len) == napi_ok)
    o->precision = strcmp(buf, "fp64") == 0 ? FCBGA_PREC_FP64 : FCBGA_PREC_MIXED;
  if (get_prop(env, v, "precond", &x) && napi_get_value_string_utf8(env, x, buf, sizeof buf, &len) == napi_ok) {
    int pc = -1;
    for (int q = FCBGA_PC_MG; q <= FCBGA_PC_IC; q++)
      if (strcmp(buf, pc_name(q)) == 0) pc = q;
    if (pc < 0) return false;
    o->precond = pc;
  }
  if (get_prop(env, v, "smoother", &x) && napi_get_value_string_utf8(env, x, buf, sizeof buf, &len) == napi_ok) {
    if (strcmp(buf, "gs") == 0) o->smoother = FCBGA_SMOOTH_GS;
    else if (strcmp(buf, "coarse-jacobi") == 0) o->smoother = FCBGA_SMOOTH_COARSE_JACOBI;
    else return false;
  }
  return true;
}

// Live progress of a synchronous solve: opts.progress (a function) is called on
// this thread with {event, iters, maxit, precond, relres, r0, tol, ms} every
// opts.progressEvery iterations; a truthy return or a throw (the reference's
// checkCancel throws Error('cancelled')) stops the solve.
struct ProgressScope {
  napi_env env;
  fcbga_model *m;
  napi_ref fn = nullptr;
  bool active = false;
  static int call(const fcbga_progress *p, void *user) {
    ProgressScope *ps = static_cast<ProgressScope *>(user);
    napi_env env = ps->env;
    bool pending = false;
    napi_is_exception_pending(env, &pending);
    if (pending) return 1;
    napi_value fn, rec, res, glob;
    if (napi_get_reference_value(env, ps->fn, &fn) != napi_ok) return 0;
    napi_create_object(env, &rec);
    static const char *ev[3] = {"start", "iter", "end"};
    napi_value e;
    napi_create_string_utf8(env, ev[p->event < 0 || p->event > 2 ? 1 : p->event], NAPI_AUTO_LENGTH, &e);
    napi_set_named_property(env, rec, "event", e);
    set_num(env, rec, "iters", p->iters);
    set_num(env, rec, "maxit", p->maxit);
    set_num(env, rec, "relres", p->relres);
    set_num(env, rec, "r0", p->r0);
    set_num(env, rec, "tol", p->tol);
    set_num(env, rec, "ms", p->ms);
    napi_value pc;
    napi_create_string_utf8(env, pc_name(p->precond), NAPI_AUTO_LENGTH, &pc);
    napi_set_named_property(env, rec, "precond", pc);
    napi_get_global(env, &glob);
    if (napi_call_function(env, glob, fn, 1, &rec, &res) != napi_ok) return 1;  // threw: stop
    bool stop = false;
    napi_valuetype t;
    napi_typeof(env, res, &t);
    if (t == napi_boolean) napi_get_value_bool(env, res, &stop);
    return stop ? 1 : 0;
  }
  ProgressScope(napi_env e, fcbga_model *model, napi_value opts) : env(e), m(model) {
    napi_valuetype t;
    if (!opts || napi_typeof(env, opts, &t) != napi_ok || t != napi_object) return;
    napi_value f;
    if (!get_prop(env, opts, "progress", &f)) return;
    napi_typeof(env, f, &t);
    if (t != napi_function) return;
    int32_t every = 1;
    get_int(env, opts, "progressEvery", &every);
    if (every < 1) every = 1;
    napi_create_reference(env, f, 1, &fn);
    active = fcbga_set_progress_cb(m, &ProgressScope::call, this, every) == FCBGA_OK;
  }
  ~ProgressScope() {
    if (active) fcbga_set_progress_cb(m, nullptr, nullptr, 1);
    if (fn) napi_delete_reference(env, fn);
  }
  // after a stopped solve: the callback's own exception if it threw, else Error('cancelled')
  napi_value cancelled() {
    bool pending = false;
    napi_is_exception_pending(env, &pending);
    if (!pending) napi_throw_error(env, nullptr, "cancelled");
    return nullptr;
  }
};

// precond of the final run, fallback {first, iters, relres} or null, and the IC(0) shift
void set_pc_fields(napi_env env, napi_value rec, const fcbga_solve_stats &S) {
  napi_value pc, fb;
  napi_create_string_utf8(env, pc_name(S.precond), NAPI_AUTO_LENGTH, &pc);
  napi_set_named_property(env, rec, "precond", pc);
  if (S.fallback_first >= 0) {
    napi_create_object(env, &fb);
    napi_value first;
    napi_create_string_utf8(env, pc_name(S.fallback_first), NAPI_AUTO_LENGTH, &first);
    napi_set_named_property(env, fb, "first", first);
    set_num(env, fb, "iters", S.fallback_iters);
    set_num(env, fb, "relres", S.fallback_relres);
  } else {
    napi_get_null(env, &fb);
  }
  napi_set_named_property(env, rec, "fallback", fb);
  set_num(env, rec, "icShift", S.ic_shift);
}

napi_value solve_record(napi_env env, const fcbga_solve_stats &S);
bool read_step(napi_env env, napi_value v, const ModelBox *b, fcbga_step *st, std::string *err);

napi_value SolveStep(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 2) return throw_msg(env, "solveStep(model, step, opts)");
  ModelBox *b = unwrap(env, argv[0]);
  if (!b || !b->m) return throw_msg(env, "solveStep: not a live model");
  fcbga_step st;
  std::string err;
  if (!read_step(env, argv[1], b, &st, &err)) return throw_msg(env, "solveStep: " + err);
  fcbga_solve_opts o{1e-8, 2000, FCBGA_PREC_MIXED, FCBGA_PC_MG, 0, 0, FCBGA_SMOOTH_GS};
  if (argc >= 3 && !read_opts(env, argv[2], &o)) return throw_msg(env, "solveStep: unknown opts.precond or opts.smoother");
  const int64_t ndof = fcbga_model_ndof(b->m);
  napi_value ab;
  void *udata = nullptr;
  NAPI_OK(napi_create_arraybuffer(env, sizeof(double) * (size_t)ndof, &udata, &ab));
  fcbga_solve_stats S;
  memset(&S, 0, sizeof S);
  int rc;
  {
    ProgressScope ps(env, b->m, argc >= 3 ? argv[2] : nullptr);
    rc = fcbga_step_solve(b->m, &st, &o, (double *)udata, &S);
    if (rc == FCBGA_ERR_CANCELLED) return ps.cancelled();
  }
  if (rc != FCBGA_OK && rc != FCBGA_ERR_NOT_CONVERGED) return throw_msg(env, "solveStep: " + lib_error());
  napi_value u;
  NAPI_OK(napi_create_typedarray(env, napi_float64_array, (size_t)ndof, ab, 0, &u));
  napi_value rec = solve_record(env, S);
  napi_set_named_property(env, rec, "u", u);
  return rec;
}

// the solve record of the reference (src/core/10-solver.js:988-993) without u
napi_value solve_record(napi_env env, const fcbga_solve_stats &S) {
  napi_value rec;
  napi_create_object(env, &rec);
  set_num(env, rec, "iters", S.iters);
  set_num(env, rec, "relres", S.relres);
  napi_value cv;
  napi_get_boolean(env, S.converged != 0, &cv);
  napi_set_named_property(env, rec, "converged", cv);
  set_num(env, rec, "ms", S.ms_total);
  set_num(env, rec, "assembleMs", S.ms_assemble);
  set_num(env, rec, "pcMs", S.ms_setup);
  set_num(env, rec, "pcgMs", S.ms_pcg);
  set_num(env, rec, "levels", S.levels);
  set_num(env, rec, "iters0", S.r0);
  set_num(env, rec, "warm", S.warm);
  napi_value t;
  napi_get_boolean(env, true, &t);
  napi_set_named_property(env, rec, "rebuilt", t);
  set_pc_fields(env, rec, S);
  return rec;
}

// step record {Ta, Tb, stage, D, s0, born[, secW, secV, secBorn]} -> fcbga_step (the arrays are
// borrowed); every array's length is checked against the model's groups and sections, since the
// library reads that many values from each
bool read_step(napi_env env, napi_value v, const ModelBox *b, fcbga_step *st, std::string *err) {
  memset(st, 0, sizeof *st);
  size_t n;
  void *p;
  const size_t ng = (size_t)b->ngroups, ns = (size_t)b->nsections;
  get_double(env, v, "Ta", &st->Ta);
  get_double(env, v, "Tb", &st->Tb);
  get_int(env, v, "stage", &st->stage);
  if (!get_typed(env, v, "D", napi_float64_array, &p, &n) || n != 36 * ng) {
    *err = "step.D: Float64Array(ngroups*36)";
    return false;
  }
  st->D = (const double *)p;
  if (!get_typed(env, v, "s0", napi_float64_array, &p, &n) || n != 6 * ng) {
    *err = "step.s0: Float64Array(ngroups*6)";
    return false;
  }
  st->s0 = (const double *)p;
  if (!get_typed(env, v, "born", napi_uint8_array, &p, &n) || n != ng) {
    *err = "step.born: Uint8Array(ngroups)";
    return false;
  }
  st->born = (const uint8_t *)p;
  if (ns > 0) {
    if (!get_typed(env, v, "secW", napi_float64_array, &p, &n) || n != 108 * ns) {
      *err = "step.secW: Float64Array(nsections*108)";
      return false;
    }
    st->secW = (const double *)p;
    if (!get_typed(env, v, "secV", napi_float64_array, &p, &n) || n != 18 * ns) {
      *err = "step.secV: Float64Array(nsections*18)";
      return false;
    }
    st->secV = (const double *)p;
    if (!get_typed(env, v, "secBorn", napi_uint8_array, &p, &n) || n != ns) {
      *err = "step.secBorn: Uint8Array(nsections)";
      return false;
    }
    st->sec_born = (const uint8_t *)p;
  }
  return true;
}

// typed array argument (not a property)
bool arg_typed(napi_env env, napi_value x, napi_typedarray_type want, void **data, size_t *len) {
  bool is = false;
  napi_is_typedarray(env, x, &is);
  if (!is) return false;
  napi_typedarray_type t;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, x, &t, len, data, &ab, &off) != napi_ok) return false;
  return t == want;
}

// stateAccumulate(model, step, du, elems[, states[, qOnly]]) -> states (Float64Array(n * 81)):
// one step of the stress histories of elements (accumulateStep, src/core/09-assembly.js:494-545);
// a given states array is updated in place and returned
napi_value StateAccumulate(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 4) return throw_msg(env, "stateAccumulate(model, step, du, elems[, states[, qOnly]])");
  ModelBox *b = unwrap(env, argv[0]);
  if (!b || !b->m) return throw_msg(env, "stateAccumulate: not a live model");
  fcbga_step st;
  std::string err;
  if (!read_step(env, argv[1], b, &st, &err)) return throw_msg(env, "stateAccumulate: " + err);
  void *du, *el;
  size_t ndu, ne;
  if (!arg_typed(env, argv[2], napi_float64_array, &du, &ndu) || (int64_t)ndu != fcbga_model_ndof(b->m))
    return throw_msg(env, "stateAccumulate: du must be a Float64Array(ndof)");
  if (!arg_typed(env, argv[3], napi_int32_array, &el, &ne)) return throw_msg(env, "stateAccumulate: elems must be an Int32Array");
  napi_value states;
  void *sp = nullptr;
  size_t ns = 0;
  napi_valuetype t = napi_undefined;
  if (argc >= 5) napi_typeof(env, argv[4], &t);
  if (t != napi_undefined && t != napi_null) {
    if (!arg_typed(env, argv[4], napi_float64_array, &sp, &ns) || ns != ne * FCBGA_STATE)
      return throw_msg(env, "stateAccumulate: states must be a Float64Array(elems.length * 81)");
    states = argv[4];
  } else {
    napi_value ab;
    NAPI_OK(napi_create_arraybuffer(env, sizeof(double) * ne * FCBGA_STATE, &sp, &ab));
    memset(sp, 0, sizeof(double) * ne * FCBGA_STATE);
    NAPI_OK(napi_create_typedarray(env, napi_float64_array, ne * FCBGA_STATE, ab, 0, &states));
  }
  bool qo = false;
  if (argc >= 6) napi_get_value_bool(env, argv[5], &qo);
  if (fcbga_state_accumulate(b->m, &st, (const double *)du, (int64_t)ne, (const int32_t *)el, (double *)sp,
                             qo ? 1 : 0) != FCBGA_OK)
    return throw_msg(env, "stateAccumulate: " + lib_error());
  return states;
}

// stateStress(model, elems, states, rows, xez) -> {sig, eps} (Float64Array(6 * npts) each):
// stress and strain at natural points (stateStress, src/core/09-assembly.js:556-571)
napi_value StateStress(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 5) return throw_msg(env, "stateStress(model, elems, states, rows, xez)");
  ModelBox *b = unwrap(env, argv[0]);
  if (!b || !b->m) return throw_msg(env, "stateStress: not a live model");
  void *el, *sp, *rw, *xz;
  size_t ne, ns, nr, nx;
  if (!arg_typed(env, argv[1], napi_int32_array, &el, &ne)) return throw_msg(env, "stateStress: elems must be an Int32Array");
  if (!arg_typed(env, argv[2], napi_float64_array, &sp, &ns) || ns != ne * FCBGA_STATE)
    return throw_msg(env, "stateStress: states must be a Float64Array(elems.length * 81)");
  if (!arg_typed(env, argv[3], napi_int32_array, &rw, &nr)) return throw_msg(env, "stateStress: rows must be an Int32Array");
  if (!arg_typed(env, argv[4], napi_float64_array, &xz, &nx) || nx != 3 * nr)
    return throw_msg(env, "stateStress: xez must be a Float64Array(3 * rows.length)");
  napi_value abs, abe, sig, eps, out;
  void *ps, *pe;
  NAPI_OK(napi_create_arraybuffer(env, sizeof(double) * 6 * nr, &ps, &abs));
  NAPI_OK(napi_create_arraybuffer(env, sizeof(double) * 6 * nr, &pe, &abe));
  if (fcbga_state_stress(b->m, (int64_t)ne, (const int32_t *)el, (const double *)sp, (int64_t)nr, (const int32_t *)rw,
                         (const double *)xz, (double *)ps, (double *)pe) != FCBGA_OK)
    return throw_msg(env, "stateStress: " + lib_error());
  NAPI_OK(napi_create_typedarray(env, napi_float64_array, 6 * nr, abs, 0, &sig));
  NAPI_OK(napi_create_typedarray(env, napi_float64_array, 6 * nr, abe, 0, &eps));
  NAPI_OK(napi_create_object(env, &out));
  napi_set_named_property(env, out, "sig", sig);
  napi_set_named_property(env, out, "eps", eps);
  return out;
}

napi_value make_f64(napi_env env, size_t n, double **data) {
  napi_value ab, arr;
  void *p = nullptr;
  if (napi_create_arraybuffer(env, sizeof(double) * n, &p, &ab) != napi_ok) return nullptr;
  if (napi_create_typedarray(env, napi_float64_array, n, ab, 0, &arr) != napi_ok) return nullptr;
  *data = (double *)p;
  return arr;
}


// adjointForward(model, step, objective, opts) -> {J: Float64Array(2), u, iters, relres, converged}
//   objective: {kTop, vmElems: Int32Array, p}
// the forward of the sensitivity pair (fcbga_adjoint_forward); the model keeps the context
napi_value AdjointForward(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 3) return throw_msg(env, "adjointForward(model, step, objective[, opts])");
  ModelBox *b = unwrap(env, argv[0]);
  if (!b || !b->m) return throw_msg(env, "adjointForward: not a live model");
  fcbga_step st;
  std::string err;
  if (!read_step(env, argv[1], b, &st, &err)) return throw_msg(env, "adjointForward: " + err);
  fcbga_objective obj;
  memset(&obj, 0, sizeof obj);
  obj.p = 8;
  if (!get_int(env, argv[2], "kTop", &obj.k_top)) return throw_msg(env, "adjointForward: objective.kTop required");
  void *p;
  size_t n;
  if (get_typed(env, argv[2], "vmElems", napi_int32_array, &p, &n)) {
    obj.vm_elems = (const int32_t *)p;
    obj.n_vm = (int32_t)n;
  }
  get_double(env, argv[2], "p", &obj.p);
  fcbga_solve_opts o{1e-10, 2000, FCBGA_PREC_MIXED, FCBGA_PC_MG, 0, 0, FCBGA_SMOOTH_GS};
  if (argc >= 4 && !read_opts(env, argv[3], &o)) return throw_msg(env, "adjointForward: unknown opts.precond or opts.smoother");
  const int64_t ndof = fcbga_model_ndof(b->m);
  double *J, *u;
  napi_value jv = make_f64(env, 2, &J), uv = make_f64(env, (size_t)ndof, &u);
  if (!jv || !uv) return throw_msg(env, "adjointForward: allocation failed");
  fcbga_solve_stats S;
  memset(&S, 0, sizeof S);
  const int rc = fcbga_adjoint_forward(b->m, &st, &o, &obj, J, u, &S);
  if (rc != FCBGA_OK && rc != FCBGA_ERR_NOT_CONVERGED) return throw_msg(env, "adjointForward: " + lib_error());
  napi_value rec, cv;
  NAPI_OK(napi_create_object(env, &rec));
  napi_set_named_property(env, rec, "J", jv);
  napi_set_named_property(env, rec, "u", uv);
  set_num(env, rec, "iters", S.iters);
  set_num(env, rec, "relres", S.relres);
  set_num(env, rec, "ms", S.ms_total);
  napi_get_boolean(env, S.converged != 0, &cv);
  napi_set_named_property(env, rec, "converged", cv);
  set_pc_fields(env, rec, S);
  return rec;
}

// adjointBackward(model, gJ: Float64Array(2)[, elemCte]) -> {dlnE, dlnCte, dlnCteElem, iters, relres, converged}
// with gJ a Float64Array(3 nn) instead: dJ/du of the caller's own objective (fcbga_adjoint_backward_u)
napi_value AdjointBackward(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 2) return throw_msg(env, "adjointBackward(model, gJ[, elemCte])");
  ModelBox *b = unwrap(env, argv[0]);
  if (!b || !b->m) return throw_msg(env, "adjointBackward: not a live model");
  void *g;
  size_t n;
  const size_t ndof = (size_t)fcbga_model_ndof(b->m);
  if (!arg_typed(env, argv[1], napi_float64_array, &g, &n) || (n != 2 && n != ndof))
    return throw_msg(env, "adjointBackward: gJ must be a Float64Array(2) or dJ/du, a Float64Array(3 nn)");
  bool elem = true;
  if (argc >= 3) napi_get_value_bool(env, argv[2], &elem);
  const int32_t ng = b->ngroups, ne = b->ne;
  double *dE, *dA, *de = nullptr;
  napi_value ev = make_f64(env, (size_t)ng, &dE), av = make_f64(env, (size_t)ng, &dA), elv = nullptr;
  if (elem) elv = make_f64(env, (size_t)ne, &de);
  fcbga_solve_stats S;
  memset(&S, 0, sizeof S);
  const int rc = n == 2 ? fcbga_adjoint_backward(b->m, (const double *)g, dE, dA, de, &S)
                        : fcbga_adjoint_backward_u(b->m, (const double *)g, dE, dA, de, &S);
  if (rc != FCBGA_OK && rc != FCBGA_ERR_NOT_CONVERGED) return throw_msg(env, "adjointBackward: " + lib_error());
  napi_value rec, cv;
  NAPI_OK(napi_create_object(env, &rec));
  napi_set_named_property(env, rec, "dlnE", ev);
  napi_set_named_property(env, rec, "dlnCte", av);
  if (elv) napi_set_named_property(env, rec, "dlnCteElem", elv);
  set_num(env, rec, "iters", S.iters);
  set_num(env, rec, "relres", S.relres);
  napi_get_boolean(env, S.converged != 0, &cv);
  napi_set_named_property(env, rec, "converged", cv);
  set_pc_fields(env, rec, S);
  return rec;
}

// screenJoints(solder, steps, table, i25) -> {dW, dE, stab}  (fcbga_screen_joints)
//   solder: {A, QR, xi, m, shat, n, h0, a, s0, nu}
//   steps:  {nsteps, cycles, T, dt, E, Fth, wt: Float64Array, wi, cycle: Int32Array, same: Uint8Array,
//            Cp, Fp: Float64Array (pillar, optional), fs}
//           nsites (number of joints)
//   table:  Float64Array(nsites * ntemps * 6) bump-layer strains; i25 the row of 25 C or -1
napi_value ScreenJoints(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 4) return throw_msg(env, "screenJoints(solder, steps, table, i25)");
  fcbga_anand_params P;
  const char *names[10] = {"A", "QR", "xi", "m", "shat", "n", "h0", "a", "s0", "nu"};
  double *pv[10] = {&P.A, &P.QR, &P.xi, &P.m, &P.shat, &P.n, &P.h0, &P.a, &P.s0, &P.nu};
  for (int q = 0; q < 10; q++)
    if (!get_double(env, argv[0], names[q], pv[q])) return throw_msg(env, std::string("screenJoints: solder.") + names[q]);
  fcbga_screen_steps st;
  memset(&st, 0, sizeof st);
  int32_t nsites = 0;
  if (!get_int(env, argv[1], "nsteps", &st.nsteps) || !get_int(env, argv[1], "cycles", &st.cycles) ||
      !get_int(env, argv[1], "nsites", &nsites))
    return throw_msg(env, "screenJoints: steps.nsteps, steps.cycles and steps.nsites required");
  st.fs = 1;
  get_double(env, argv[1], "fs", &st.fs);
  const size_t n = (size_t)(st.nsteps > 0 ? st.nsteps : 0);
  struct { const char *k; const double **p; size_t len; } f64[] = {
      {"T", &st.T, n}, {"dt", &st.dt, n}, {"E", &st.E, n}, {"Fth", &st.Fth, n}, {"wt", &st.wt, n}};
  void *p;
  size_t len;
  for (auto &a : f64) {
    if (!get_typed(env, argv[1], a.k, napi_float64_array, &p, &len) || len != a.len)
      return throw_msg(env, std::string("screenJoints: steps.") + a.k + ": Float64Array(nsteps)");
    *a.p = (const double *)p;
  }
  if (!get_typed(env, argv[1], "wi", napi_int32_array, &p, &len) || len != n)
    return throw_msg(env, "screenJoints: steps.wi: Int32Array(nsteps)");
  st.wi = (const int32_t *)p;
  if (!get_typed(env, argv[1], "cycle", napi_int32_array, &p, &len) || len != n)
    return throw_msg(env, "screenJoints: steps.cycle: Int32Array(nsteps)");
  st.cycle = (const int32_t *)p;
  if (!get_typed(env, argv[1], "same", napi_uint8_array, &p, &len) || len != n)
    return throw_msg(env, "screenJoints: steps.same: Uint8Array(nsteps)");
  st.same = (const uint8_t *)p;
  if (get_typed(env, argv[1], "Cp", napi_float64_array, &p, &len)) {
    if (len != 3 * n) return throw_msg(env, "screenJoints: steps.Cp: Float64Array(3 * nsteps)");
    st.Cp = (const double *)p;
    if (!get_typed(env, argv[1], "Fp", napi_float64_array, &p, &len) || len != n)
      return throw_msg(env, "screenJoints: steps.Fp: Float64Array(nsteps) with Cp");
    st.Fp = (const double *)p;
  }
  void *tp;
  size_t tlen;
  if (!arg_typed(env, argv[2], napi_float64_array, &tp, &tlen) || nsites <= 0 || tlen % (6 * (size_t)nsites) != 0)
    return throw_msg(env, "screenJoints: table: Float64Array(nsites * ntemps * 6)");
  const int32_t ntemps = (int32_t)(tlen / (6 * (size_t)nsites));
  int32_t i25 = -1;
  napi_get_value_int32(env, argv[3], &i25);
  double *dW, *dE, *stab;
  napi_value vW = make_f64(env, nsites, &dW), vE = make_f64(env, nsites, &dE), vS = make_f64(env, nsites, &stab);
  if (!vW || !vE || !vS) return throw_msg(env, "screenJoints: out of memory");
  if (fcbga_screen_joints(&P, &st, nsites, ntemps, i25, (const double *)tp, dW, dE, stab) != FCBGA_OK)
    return throw_msg(env, "screenJoints: " + lib_error());
  napi_value rec;
  NAPI_OK(napi_create_object(env, &rec));
  napi_set_named_property(env, rec, "dW", vW);
  napi_set_named_property(env, rec, "dE", vE);
  napi_set_named_property(env, rec, "stab", vS);
  return rec;
}

// submodelCreate({nx, ny, nz, xs, ys, zs: Float64Array, cellMat: Int8Array, cellLayer: Uint8Array},
//                solder {A, QR, xi, m, shat, n, h0, a, s0, nu}) -> handle (fcbga_submodel_create)
napi_value SubmodelCreate(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 2) return throw_msg(env, "submodelCreate(mesh, solder)");
  fcbga_submodel_mesh m;
  memset(&m, 0, sizeof m);
  if (!get_int(env, argv[0], "nx", &m.nx) || !get_int(env, argv[0], "ny", &m.ny) || !get_int(env, argv[0], "nz", &m.nz))
    return throw_msg(env, "submodelCreate: mesh.nx, ny, nz required");
  void *p;
  size_t n;
  const int64_t ncell = (int64_t)(m.nx - 1) * (m.ny - 1) * (m.nz - 1);
  if (!get_typed(env, argv[0], "xs", napi_float64_array, &p, &n) || (int64_t)n != m.nx) return throw_msg(env, "submodelCreate: mesh.xs");
  m.xs = (const double *)p;
  if (!get_typed(env, argv[0], "ys", napi_float64_array, &p, &n) || (int64_t)n != m.ny) return throw_msg(env, "submodelCreate: mesh.ys");
  m.ys = (const double *)p;
  if (!get_typed(env, argv[0], "zs", napi_float64_array, &p, &n) || (int64_t)n != m.nz) return throw_msg(env, "submodelCreate: mesh.zs");
  m.zs = (const double *)p;
  if (!get_typed(env, argv[0], "cellMat", napi_int8_array, &p, &n) || (int64_t)n != ncell)
    return throw_msg(env, "submodelCreate: mesh.cellMat: Int8Array(cells)");
  m.cell_mat = (const int8_t *)p;
  if (!get_typed(env, argv[0], "cellLayer", napi_uint8_array, &p, &n) || (int64_t)n != ncell)
    return throw_msg(env, "submodelCreate: mesh.cellLayer: Uint8Array(cells)");
  m.cell_layer = (const uint8_t *)p;
  fcbga_anand_params P;
  const char *names[10] = {"A", "QR", "xi", "m", "shat", "n", "h0", "a", "s0", "nu"};
  double *pv[10] = {&P.A, &P.QR, &P.xi, &P.m, &P.shat, &P.n, &P.h0, &P.a, &P.s0, &P.nu};
  for (int q = 0; q < 10; q++)
    if (!get_double(env, argv[1], names[q], pv[q])) return throw_msg(env, std::string("submodelCreate: solder.") + names[q]);
  fcbga_submodel *sm = nullptr;
  if (fcbga_submodel_create(&m, &P, &sm) != FCBGA_OK) return throw_msg(env, "submodelCreate: " + lib_error());
  SubBox *box = new SubBox();
  box->s = sm;
  box->nb = fcbga_submodel_nboundary(sm);
  box->ne = ncell;
  box->nn = (int64_t)m.nx * m.ny * m.nz;
  napi_value ext;
  if (napi_create_external(env, box, finalize_submodel, nullptr, &ext) != napi_ok) {
    fcbga_submodel_destroy(sm);
    delete box;
    return throw_msg(env, "submodelCreate: could not wrap the submodel");
  }
  return ext;
}

// submodelCall(handle, op, ...) for the step-level engine:
//   'step', {T, dt, E, fsolder, elasD: Float64Array(144), elasSth: Float64Array(24), ub: Float64Array(3 nb)}
//        -> {converged, newton, dMax, pcgIters, icBuilds}
//   'save' | 'restore' | 'destroy' -> undefined
//   'nboundary' -> number
//   'layerMeans' -> Float64Array(4) {W die, ebar die, W sub, ebar sub}
//   'loopSample', T, E, fsolder, cphi, sphi -> Float64Array(4) {tau, gamma die; tau, gamma sub}
//   'section', E, fsolder -> Float32Array(48 per element)
//   'displacement' -> Float64Array(3 per node)
napi_value SubmodelCall(napi_env env, napi_callback_info info) {
  size_t argc = 7;
  napi_value argv[7];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 2) return throw_msg(env, "submodelCall(handle, op, ...)");
  SubBox *b = unwrap_sub(env, argv[0]);
  if (!b) return throw_msg(env, "submodelCall: not a live submodel");
  char op[32];
  size_t len;
  if (napi_get_value_string_utf8(env, argv[1], op, sizeof op, &len) != napi_ok) return throw_msg(env, "submodelCall: op");
  auto num = [&](size_t q, double *v) { return q < argc && napi_get_value_double(env, argv[q], v) == napi_ok; };
  if (!strcmp(op, "step")) {
    if (argc < 3) return throw_msg(env, "submodelCall(h, 'step', data)");
    fcbga_submodel_step_in in;
    memset(&in, 0, sizeof in);
    if (!get_double(env, argv[2], "T", &in.T) || !get_double(env, argv[2], "dt", &in.dt) ||
        !get_double(env, argv[2], "E", &in.E) || !get_double(env, argv[2], "fsolder", &in.fsolder))
      return throw_msg(env, "submodel step: T, dt, E, fsolder required");
    void *p;
    size_t n;
    if (!get_typed(env, argv[2], "elasD", napi_float64_array, &p, &n) || n != 144) return throw_msg(env, "submodel step: elasD: Float64Array(144)");
    in.elas_D = (const double *)p;
    if (!get_typed(env, argv[2], "elasSth", napi_float64_array, &p, &n) || n != 24) return throw_msg(env, "submodel step: elasSth: Float64Array(24)");
    in.elas_sth = (const double *)p;
    if (!get_typed(env, argv[2], "ub", napi_float64_array, &p, &n) || n != 3 * (size_t)b->nb)
      return throw_msg(env, "submodel step: ub: Float64Array(3 * boundary nodes)");
    in.ub = (const double *)p;
    fcbga_submodel_step_out out;
    if (fcbga_submodel_step(b->s, &in, &out) != FCBGA_OK) return throw_msg(env, "submodel step: " + lib_error());
    napi_value rec, cv;
    NAPI_OK(napi_create_object(env, &rec));
    napi_get_boolean(env, out.converged != 0, &cv);
    napi_set_named_property(env, rec, "converged", cv);
    set_num(env, rec, "newton", out.newton);
    set_num(env, rec, "dMax", out.dmax);
    set_num(env, rec, "pcgIters", out.pcg_iters);
    set_num(env, rec, "icBuilds", out.ic_builds);
    return rec;
  }
  if (!strcmp(op, "save") || !strcmp(op, "restore")) {
    const int rc = op[0] == 's' ? fcbga_submodel_save(b->s) : fcbga_submodel_restore(b->s);
    if (rc != FCBGA_OK) return throw_msg(env, std::string("submodel ") + op + ": " + lib_error());
    return nullptr;
  }
  if (!strcmp(op, "destroy")) {
    fcbga_submodel_destroy(b->s);
    b->s = nullptr;
    return nullptr;
  }
  if (!strcmp(op, "nboundary")) {
    napi_value v;
    napi_create_int32(env, b->nb, &v);
    return v;
  }
  if (!strcmp(op, "layerMeans") || !strcmp(op, "loopSample")) {
    double *o;
    napi_value arr = make_f64(env, 4, &o);
    if (!arr) return throw_msg(env, "submodel: out of memory");
    int rc;
    if (op[0] == 'l' && op[1] == 'a') rc = fcbga_submodel_layer_means(b->s, o);
    else {
      double T, E, f, c, sn;
      if (!num(2, &T) || !num(3, &E) || !num(4, &f) || !num(5, &c) || !num(6, &sn))
        return throw_msg(env, "submodelCall(h, 'loopSample', T, E, fsolder, cphi, sphi)");
      rc = fcbga_submodel_loop_sample(b->s, T, E, f, c, sn, o);
    }
    if (rc != FCBGA_OK) return throw_msg(env, std::string("submodel ") + op + ": " + lib_error());
    return arr;
  }
  if (!strcmp(op, "section")) {
    double E, f;
    if (!num(2, &E) || !num(3, &f)) return throw_msg(env, "submodelCall(h, 'section', E, fsolder)");
    napi_value ab, arr;
    void *data = nullptr;
    NAPI_OK(napi_create_arraybuffer(env, sizeof(float) * 48 * (size_t)b->ne, &data, &ab));
    NAPI_OK(napi_create_typedarray(env, napi_float32_array, 48 * (size_t)b->ne, ab, 0, &arr));
    if (fcbga_submodel_section(b->s, E, f, (float *)data) != FCBGA_OK) return throw_msg(env, "submodel section: " + lib_error());
    return arr;
  }
  if (!strcmp(op, "points")) {
    double *o;
    napi_value arr = make_f64(env, 8 * FCBGA_ANAND_STATE * (size_t)b->ne, &o);
    if (!arr) return throw_msg(env, "submodel: out of memory");
    if (fcbga_submodel_points(b->s, o) != FCBGA_OK) return throw_msg(env, "submodel points: " + lib_error());
    return arr;
  }
  if (!strcmp(op, "displacement")) {
    double *o;
    napi_value arr = make_f64(env, 3 * (size_t)b->nn, &o);
    if (!arr) return throw_msg(env, "submodel: out of memory");
    if (fcbga_submodel_displacement(b->s, o) != FCBGA_OK) return throw_msg(env, "submodel displacement: " + lib_error());
    return arr;
  }
  return throw_msg(env, std::string("submodelCall: unknown op ") + op);
}

}  // namespace

// pathCall(model, op, ...): the thermal path kept on the device (Analysis.run / restorePath /
// finishPath, src/core/11-analysis.js:212-261):
//   'reset'                  drop every kept increment
//   'solve', k, step[, opts] solve step k of the plan and keep its increment -> solve record
//   'sum', ids               sum of the kept increments of steps ids, in that order -> Float64Array
//   'du', k                  step k's increment in single precision -> Float32Array
//   'load', k, du            keep du as step k's increment: a results file's (Float32Array) or a
//                            helper's partial message (Float64Array)
napi_value PathCall(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  NAPI_OK(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  if (argc < 2) return throw_msg(env, "pathCall(model, op, ...)");
  ModelBox *b = unwrap(env, argv[0]);
  if (!b || !b->m) return throw_msg(env, "pathCall: not a live model");
  char op[16] = {0};
  size_t ol = 0;
  if (napi_get_value_string_utf8(env, argv[1], op, sizeof op, &ol) != napi_ok) return throw_msg(env, "pathCall: op");
  const std::string o(op);
  const int64_t ndof = fcbga_model_ndof(b->m);
  napi_value undef;
  napi_get_undefined(env, &undef);
  int32_t k = -1;
  if (o != "reset" && o != "sum" && (argc < 3 || napi_get_value_int32(env, argv[2], &k) != napi_ok))
    return throw_msg(env, "pathCall: step index");
  if (o == "reset") {
    if (fcbga_path_reset(b->m) != FCBGA_OK) return throw_msg(env, "pathCall reset: " + lib_error());
    return undef;
  }
  if (o == "solve") {
    if (argc < 4) return throw_msg(env, "pathCall(model, 'solve', k, step[, opts])");
    fcbga_step st;
    std::string err;
    if (!read_step(env, argv[3], b, &st, &err)) return throw_msg(env, err);
    fcbga_solve_opts so{1e-8, 2000, FCBGA_PREC_MIXED, FCBGA_PC_MG, 0, 0, FCBGA_SMOOTH_GS};
    if (argc >= 5 && !read_opts(env, argv[4], &so)) return throw_msg(env, "pathCall solve: unknown opts.precond or opts.smoother");
    fcbga_solve_stats S;
    memset(&S, 0, sizeof S);
    int rc;
    {
      ProgressScope ps(env, b->m, argc >= 5 ? argv[4] : nullptr);
      rc = fcbga_path_solve(b->m, k, &st, &so, &S);
      if (rc == FCBGA_ERR_CANCELLED) return ps.cancelled();
    }
    if (rc != FCBGA_OK && rc != FCBGA_ERR_NOT_CONVERGED) return throw_msg(env, "pathCall solve: " + lib_error());
    return solve_record(env, S);
  }
  if (o == "sum") {
    if (argc < 3) return throw_msg(env, "pathCall(model, 'sum', ids)");
    std::vector<int32_t> ids;
    bool isArr = false;
    napi_is_array(env, argv[2], &isArr);
    void *p;
    size_t n;
    if (isArr) {
      uint32_t len = 0;
      napi_get_array_length(env, argv[2], &len);
      ids.resize(len);
      for (uint32_t q = 0; q < len; q++) {
        napi_value e;
        napi_get_element(env, argv[2], q, &e);
        if (napi_get_value_int32(env, e, &ids[q]) != napi_ok) return throw_msg(env, "pathCall sum: ids");
      }
    } else if (arg_typed(env, argv[2], napi_int32_array, &p, &n)) {
      ids.assign((const int32_t *)p, (const int32_t *)p + n);
    } else {
      return throw_msg(env, "pathCall sum: ids must be an Array or Int32Array");
    }
    double *out;
    napi_value arr = make_f64(env, (size_t)ndof, &out);
    if (fcbga_path_sum(b->m, (int32_t)ids.size(), ids.data(), out) != FCBGA_OK)
      return throw_msg(env, "pathCall sum: " + lib_error());
    return arr;
  }
  if (o == "du") {
    napi_value ab, arr;
    void *data = nullptr;
    NAPI_OK(napi_create_arraybuffer(env, sizeof(float) * (size_t)ndof, &data, &ab));
    NAPI_OK(napi_create_typedarray(env, napi_float32_array, (size_t)ndof, ab, 0, &arr));
    if (fcbga_path_du(b->m, k, (float *)data) != FCBGA_OK) return throw_msg(env, "pathCall du: " + lib_error());
    return arr;
  }
  if (o == "load") {  // FP32 from a results file, FP64 from a helper's partial message
    void *p;
    size_t n;
    if (argc >= 4 && arg_typed(env, argv[3], napi_float32_array, &p, &n) && (int64_t)n == ndof) {
      if (fcbga_path_load(b->m, k, (const float *)p) != FCBGA_OK) return throw_msg(env, "pathCall load: " + lib_error());
      return undef;
    }
    if (argc >= 4 && arg_typed(env, argv[3], napi_float64_array, &p, &n) && (int64_t)n == ndof) {
      if (fcbga_path_load64(b->m, k, (const double *)p) != FCBGA_OK)
        return throw_msg(env, "pathCall load: " + lib_error());
      return undef;
    }
    return throw_msg(env, "pathCall load: du must be a Float32Array or Float64Array of ndof values");
  }
  return throw_msg(env, "pathCall: unknown op " + o);
}

static napi_value Init(napi_env env, napi_value exports) {
  struct { const char *name; napi_callback cb; } fns[] = {
      {"createModel", CreateModel}, {"destroyModel", DestroyModel}, {"solveStep", SolveStep},
      {"stateAccumulate", StateAccumulate}, {"stateStress", StateStress},
      {"adjointForward", AdjointForward}, {"adjointBackward", AdjointBackward}, {"screenJoints", ScreenJoints},
      {"submodelCreate", SubmodelCreate}, {"submodelCall", SubmodelCall}, {"pathCall", PathCall}};
  for (auto &f : fns) {
    napi_value fn;
    if (napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn) != napi_ok ||
        napi_set_named_property(env, exports, f.name, fn) != napi_ok) {
      napi_throw_error(env, nullptr, "fcbga_hip: could not register the addon functions");
      return nullptr;
    }
  }
  return exports;
}

// context-aware registration: Node calls this once per environment (main thread or worker_thread)
extern "C" NAPI_MODULE_EXPORT napi_value napi_register_module_v1(napi_env env, napi_value exports) {
  return Init(env, exports);
}
