// fcbga_hip.hip -- MI355X (gfx950) matrix-free hex8 thermoelastic step solve:
// device assembly of condensed element tables, Jacobi-scaled PCG (FP64, or
// FP64 flexible PCG around an FP32 multigrid cycle) with the shell-aggregate
// semi-coarsening multigrid, behind the C ABI of include/fcbga_hip.h.
//
// Reference behaviour followed (file:line of the reference repository):
//   element and condensation  src/core/07-element.js:19-120, src/core/09-assembly.js:26-133
//   step assembly, birth      src/core/09-assembly.js:346-424
//   constraints               src/core/09-assembly.js:445-464
//   scaling, PCG              src/core/10-solver.js:852-866, 1001-1042
//   multigrid                 src/core/10-solver.js:12-715
// Device layout: see fe_core.h.  No assembled fine matrix exists: every fine
// operator row is gathered from the condensed element tables.
#include "../../include/fcbga_hip.h"
#include "backend.h"
#include "fe_core.h"
#include "mg_setup.h"
#include "step_data.h"

#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <cstring>
#include <string>
#include <tuple>
#include <vector>

namespace {

thread_local std::string g_err;

constexpr int SHELL_MAXM = 512;
// Column descriptor of a shell level (one 48-byte record per column, read by
// every lane of the column's wavefront in one load): [0] first aggregate n0,
// [1] aggregates na, [2..3] offset of the column's inverse (int64), then the
// block-row pointers rowPtr[n0 .. n0 + min(na, SHELL_DESC_ROWS)].  The sweep
// of a column with at most SHELL_DESC_ROWS aggregates needs no other index
// load before its blocks.
constexpr int SHELL_DESC = 12;
constexpr int SHELL_DESC_ROWS = SHELL_DESC - 5;
// largest deviation (relative to the largest entry) of a condensed element
// table from its mirror-symmetric form for the compact tables to be used
constexpr double SYM_TOL = 1e-11;  // 6 DOFs x up to 85 aggregates in one shell column
constexpr int FINE_MAXNZ = 64;  // grid planes in z: one lane per plane in a fine column block
// the fine operator's products on the matrix cores unless FCBGA_ELEM_OP says otherwise
constexpr bool ELEM_OP_MFMA_DEFAULT = false;

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    p = (T *)be::dmalloc(sizeof(T) * (count ? count : 1));
    n = count;
  }
  void release() {
    be::dfree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const T *src, size_t count, be::stream_t s) {
    alloc(count);
    be::h2d(p, src, sizeof(T) * count, s);
  }
  void upload(const std::vector<T> &v, be::stream_t s) { upload(v.data(), v.size(), s); }
  void zero(be::stream_t s) {
    T *q = p;
    be::pfor((int64_t)n, s, [=] __host__ __device__(int64_t i) { q[i] = T(0); });
  }
};

// ---------------------------------------------------------------------------
// natural-coordinate integrals and shape matrices (host, per step)
// ---------------------------------------------------------------------------

struct Univ {
  double G[3 * 9 * 121], g[3 * 3 * 11], GH[9 * 121], gH[3 * 11];
};

// Integrals of d_r N_a d_s N_b over the 2x2 in-plane Gauss points at the three
// zeta stations -1, 0, 1, and their Simpson sum 1/3, 4/3, 1/3 through the
// thickness (exact for the trilinear + incompatible-mode products; the UNIV
// table of src/core/09-assembly.js:26-60).
// Built once behind a function-local static (thread-safe initialisation):
// models created on several host threads at once share the one table.
Univ make_univ() {
  Univ U;
  std::memset(&U, 0, sizeof U);
  const int XI[8] = {-1, 1, 1, -1, -1, 1, 1, -1}, ETA[8] = {-1, -1, 1, 1, -1, -1, 1, 1},
            ZE[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
  const double wz[3] = {1.0 / 3, 4.0 / 3, 1.0 / 3};
  const double gp = 1 / std::sqrt(3.0);
  for (int st = 0; st < 3; st++) {
    double zeta = st - 1;
    for (int ip = 0; ip < 4; ip++) {
      double xi = (ip & 1) ? gp : -gp, eta = (ip & 2) ? gp : -gp;
      double d[3][11];
      for (int a = 0; a < 8; a++) {
        d[0][a] = XI[a] * (1 + ETA[a] * eta) * (1 + ZE[a] * zeta) * 0.125;
        d[1][a] = ETA[a] * (1 + XI[a] * xi) * (1 + ZE[a] * zeta) * 0.125;
        d[2][a] = ZE[a] * (1 + XI[a] * xi) * (1 + ETA[a] * eta) * 0.125;
      }
      for (int r = 0; r < 3; r++)
        for (int a = 8; a < 11; a++) d[r][a] = 0;
      d[0][8] = -2 * xi; d[1][9] = -2 * eta; d[2][10] = -2 * zeta;
      for (int r = 0; r < 3; r++)
        for (int a = 0; a < 11; a++) {
          U.g[(st * 3 + r) * 11 + a] += d[r][a];
          U.gH[r * 11 + a] += wz[st] * d[r][a];
          for (int s = 0; s < 3; s++)
            for (int b = 0; b < 11; b++) {
              double v = d[r][a] * d[s][b];
              U.G[(st * 9 + r * 3 + s) * 121 + a * 11 + b] += v;
              U.GH[(r * 3 + s) * 121 + a * 11 + b] += wz[st] * v;
            }
        }
    }
  }
  return U;
}

const Univ &univ() {
  static const Univ U = make_univ();
  return U;
}

// M^m (m = xx yy zz xy yz zx) with K33 = sum_m c_m M^m from a stiffness D (one
// station, full-box integrals) or three zeta-station weights W (layered); the
// construction of shapeMatrices, src/core/09-assembly.js:69-93.
void shape_matrices(const double *W3, const double *D, double *M) {
  const Univ &U = univ();
  static const int VI[3][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}};
  static const int RS[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {1, 2}, {2, 0}};
  std::memset(M, 0, sizeof(double) * 6 * 1089);
  int nst = D ? 1 : 3;
  for (int st = 0; st < nst; st++) {
    const double *Ds = D ? D : W3 + 36 * st;
    for (int m = 0; m < 6; m++) {
      double *Mm = M + m * 1089;
      int r = RS[m][0], s = RS[m][1];
      for (int pass = 0; pass < (r == s ? 1 : 2); pass++) {
        int rr = pass ? s : r, ss = pass ? r : s;
        const double *Gb = D ? U.GH + (rr * 3 + ss) * 121 : U.G + (st * 9 + rr * 3 + ss) * 121;
        for (int p = 0; p < 3; p++)
          for (int q = 0; q < 3; q++) {
            double dv = Ds[VI[p][rr] * 6 + VI[q][ss]];
            if (dv == 0) continue;
            for (int a = 0; a < 11; a++)
              for (int b = 0; b < 11; b++) Mm[(3 * a + p) * 33 + 3 * b + q] += dv * Gb[a * 11 + b];
          }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// device multigrid level (shell)
// ---------------------------------------------------------------------------

struct DevShell {
  int nx = 0, ny = 0;
  int32_t nagg = 0, ncol = 0, nb = 0;
  bool coarsest = false;
  DBuf<int32_t> colStart, rowPtr, col, blockRow;
  DBuf<double> val64;
  DBuf<float> val32;
  DBuf<int64_t> cfPtr;
  DBuf<double> cf64, cfTmp;
  DBuf<float> cf32;
  DBuf<int32_t> colors[4];
  int32_t ncolor[4] = {0, 0, 0, 0};
  DBuf<int32_t> colDesc;  // SHELL_DESC ints per column (shell_desc_fill)
  // level 1: aggregate column and k range
  DBuf<int32_t> aggCol, k0, k1;
  // transfer to the next coarser level
  DBuf<int32_t> pPtr, pIdx;
  DBuf<double> pW64;
  DBuf<float> pW32;
  // children of this level's aggregates in the finer shell level
  DBuf<int32_t> chPtr, chIdx, chQ;
  // two-phase Galerkin product from the finer shell level (gal2): T = A_f P
  // by fine rows (tPtr: T blocks of each fine row, tsPtr/tSlot: the T block of
  // every (fine block, parent) pair in walk order), then A = P^T T by coarse
  // rows (csPtr/cSlot: the row's block of every (child, T block) pair, -1 when
  // outside the pattern)
  bool gal2 = false;
  int galMaxT = 0;
  DBuf<int32_t> tPtr, tsPtr, tSlot, csPtr, cSlot;
  DBuf<double> galT;
  // vectors
  DBuf<double> b64, x64, r64;
  DBuf<float> b32, x32, r32;
  // coarsest: dense inverse (FP64)
  DBuf<double> inv64, band, jit;
  int32_t ndense = 0;
  int64_t cfTotal = 0;
  int maxm = 0;    // largest column block (6 x aggregates)
  int maxrow = 0;  // most blocks in one row
};

}  // namespace

// saved context of the adjoint forward/backward pair
struct AdjTape {
  bool valid = false;
  uint64_t gen = 0;  // prepGen of the forward's prepared step
  double Ta = 0, Tb = 0;
  int32_t stage = 0;
  std::vector<double> D, s0, secW, secV;
  std::vector<uint8_t> born, secBorn;
  fcbga_solve_opts opts{};
  double p = 8, wmean = 0, S = 0, J[2] = {0, 0};
  int32_t ntop = 0, nvm = 0;
  DBuf<double> u, lam, rhs, rhsGrid, work, ones, states, statesTh, sig, sigTh, dJds, gvm, dE, dA, red, dJdu;
  DBuf<int32_t> topNodes, vmElems, vmSlot;
};

struct fcbga_model {
  fe::Grid g{};
  int32_t nn = 0, ne = 0, ngroups = 0, nsec = 0;
  std::vector<double> xs, ys, zs;
  std::vector<int32_t> node_id, cell_elems, e_group, e_section, constraints, fx, fy;
  std::vector<int8_t> e_part;
  std::vector<double> e_scale;
  // element types
  int32_t ntypes = 0, nslot = 0;
  std::vector<int32_t> typeSlot;  // material slot of a type
  std::vector<double> typeH;      // hx hy hz per type
  std::vector<int32_t> e_type;
  be::stream_t stream = nullptr;
  // device: mesh-constant
  DBuf<int32_t> cellType, cellElem, d_egroup, d_typeSlot;
  DBuf<double> d_escale, d_typeH;
  // free DOFs of the current step (1 free, 0 held or absent); the package's
  // set, or in a stacked-module stage the module stage's (assemble_step)
  DBuf<uint8_t> dofmask;
  std::vector<uint8_t> maskAll;      // the package constraints held
  std::vector<int32_t> modConstraints;  // held in every module stage (with the package's)
  std::vector<int64_t> modGridDofs;     // the same as grid DOFs (3 * grid node + component)
  std::vector<uint8_t> maskBorn;     // group birth flags of the module stage in dofmask (empty: maskAll)
  std::vector<uint8_t> hierBorn;     // group birth flags the hierarchy was built for (empty: every element)
  DBuf<int64_t> gridOf;
  DBuf<int32_t> fineColors[4];
  int32_t nFineColor[4] = {0, 0, 0, 0};
  // device: per step
  DBuf<double> cellScale64, Kc64, fc64, slotM, slotLoad, tabs;
  DBuf<float> cellScale32, Kc32;
  // compact mirror-symmetric element tables (fe::SYM_TABLE per type), used
  // by the fine operator when every type passes the mirror check
  DBuf<double> R64, symErr;
  // stress recovery (built on first use): grid lines, node ids, cell of each element
  DBuf<double> d_xs, d_ys, d_zs, postM, postD, postS0, postDu;
  DBuf<int32_t> d_nodeId, d_eCell;
  DBuf<uint8_t> postBorn;
  DBuf<float> R32;
  bool symTables = false;
  bool sym_allowed = true;  // FCBGA_FULL_TABLES=1 forces the full tables
  // fine operator products on the matrix cores (mop::k_elem_mfma, element
  // groups of every tile) or by node rows on the vector units
  // (k_fine_col_op); FCBGA_ELEM_OP=mfma|valu chooses, per model
  bool mfmaOp = false;
  int32_t mopTiles[3] = {0, 0, 0};  // tiles in x, in y, chunks in z
  DBuf<int32_t> mopType, mopOff;
  DBuf<int16_t> mopCell;
  DBuf<uint32_t> mopMask;
  // elements of group ownGroup (-1: none) each get a type of their own,
  // types ownType0.. in the order of ownElems; when ownK64 is set (device,
  // 576 FP64 per own type) assemble_step takes those types' tables from it
  // and gives them no thermal load (the viscoplastic Newton's tangents)
  int32_t ownGroup = -1, ownType0 = 0;
  std::vector<int32_t> ownElems;
  std::vector<int64_t> ownCells;  // cell of every own-type element
  const double *ownK64 = nullptr;
  // generation of the prepared step (operator + preconditioner), bumped by every prepare_step
  uint64_t prepGen = 0;
  int prepPrecond = -1, prepPrecision = -1;
  AdjTape tape;
  // warm start: orthonormal previous solutions (grid layout, unscaled), oldest first
  std::vector<std::unique_ptr<DBuf<double>>> basis;
  // thermal path: the increment of every solved step (FP64, reference node order)
  std::vector<std::unique_ptr<DBuf<double>>> pathDu;
  DBuf<double> pathSum;
  DBuf<double> warmTmp, warmKv;
  DBuf<uint8_t> slotHasF, slotLayered, typeHasF, d_born;
  DBuf<double> s64, w64, sinv64, bload, Db, Lb, F64, C64;
  DBuf<float> w32, sinv32, C32;
  // solver vectors (grid layout, 3 per node)
  DBuf<double> x, r, z, p, q, tmp64, dx64;
  DBuf<float> b32v, x32v, r32v, dx32;
  DBuf<double> partials, scal;
  // multigrid
  bool hierBuilt = false;
  mgs::Hierarchy H;
  std::vector<std::unique_ptr<DevShell>> lv;
  DBuf<int32_t> nodeAgg;
  DBuf<double> zp64;
  DBuf<float> zp32;
  // IC(0): upper-pattern slot bits per grid node, nodes by dependency level, factor (14 3x3 slots per grid node)
  bool icPattern = false;
  DBuf<uint16_t> icMask;
  DBuf<int32_t> icNodes, icFail;
  std::vector<int64_t> icLvPtr;
  DBuf<double> icU;
  double icShift = 0;
  bool colsFactored = false;  // fine column factors of the prepared step exist
  // kernel timing (measurement hook): event pairs per kernel class
  bool timing = false;
  // live progress and cancellation (fcbga_set_progress_cb)
  fcbga_progress_fn progFn = nullptr;
  void *progUser = nullptr;
  int32_t progEvery = 1;
  bool cancelled = false;
  double solveT0 = 0;
  int32_t itersBase = 0;
  int nsweeps = 2;     // pre- and post-smoothing sweeps per level
  int jacobiFrom = 1 << 20;  // first shell level (index into lv) smoothed by column Jacobi
  bool useGraphs = true;
#if FCBGA_GPU
  std::map<int, hipGraphExec_t> graphs;
  double *hostScal = nullptr;  // pinned
#endif
  double kt_ms[FCBGA_KCLASS_COUNT] = {};
  int32_t kt_n[FCBGA_KCLASS_COUNT] = {};
#if FCBGA_GPU
  std::vector<hipEvent_t> evPool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> evUsed;
#endif
};

namespace {

// Kernel timing: when enabled, launches of the timed classes are bracketed by
// HIP events on the model's stream; the elapsed times are summed per class
// after each solve (fcbga_kernel_times).
struct KScope {
  fcbga_model *M;
  int cls;
#if FCBGA_GPU
  hipEvent_t e0 = nullptr, e1 = nullptr;
#endif
  KScope(fcbga_model *m, int c) : M(m), cls(c) {
#if FCBGA_GPU
    if (!M->timing) return;
    for (int q = 0; q < 2; q++) {
      hipEvent_t e;
      if (M->evPool.empty()) BE_CHECK(hipEventCreate(&e));
      else { e = M->evPool.back(); M->evPool.pop_back(); }
      (q ? e1 : e0) = e;
    }
    BE_CHECK(hipEventRecord(e0, M->stream));
#endif
  }
  ~KScope() {
#if FCBGA_GPU
    if (!M->timing || !e0) return;
    (void)hipEventRecord(e1, M->stream);
    M->evUsed.push_back({cls, {e0, e1}});
#endif
  }
};

void ktimer_collect(fcbga_model *M) {
#if FCBGA_GPU
  if (M->evUsed.empty()) return;
  BE_CHECK(hipStreamSynchronize(M->stream));
  for (auto &u : M->evUsed) {
    float ms = 0;
    BE_CHECK(hipEventElapsedTime(&ms, u.second.first, u.second.second));
    M->kt_ms[u.first] += ms;
    M->kt_n[u.first] += 1;
    M->evPool.push_back(u.second.first);
    M->evPool.push_back(u.second.second);
  }
  M->evUsed.clear();
#else
  (void)M;
#endif
}

}  // namespace

namespace {

// ---------------------------------------------------------------------------
// model creation
// ---------------------------------------------------------------------------

std::vector<int32_t> size_classes(const std::vector<double> &c, std::vector<double> &rep) {
  std::vector<int32_t> cls(c.size() - 1);
  rep.clear();
  for (size_t i = 0; i + 1 < c.size(); i++) {
    double h = c[i + 1] - c[i];
    int32_t f = -1;
    for (size_t q = 0; q < rep.size(); q++)
      if (std::fabs(h - rep[q]) <= 1e-9 * std::fabs(rep[q])) { f = (int32_t)q; break; }
    if (f < 0) { f = (int32_t)rep.size(); rep.push_back(h); }
    cls[i] = f;
  }
  return cls;
}

void check_mesh(const fcbga_mesh *m) {
  if (!m) throw be::Error("mesh is NULL");
  if (m->nx < 2 || m->ny < 2 || m->nz < 2) throw be::Error("mesh needs at least 2 grid lines per axis");
  if (!m->xs || !m->ys || !m->zs || !m->node_id || !m->cell_elems || !m->e_group || !m->e_part || !m->e_scale)
    throw be::Error("mesh arrays missing");
  if (m->ngroups < 1) throw be::Error("mesh has no material groups");
  if (m->nsections > 0 && !m->e_section) throw be::Error("layered sections given without e_section");
}

void mop_setup(fcbga_model *M, const std::vector<int32_t> &cellType);

void model_create(const fcbga_mesh *mm, fcbga_model *M, int32_t ownGroup = -1) {
  check_mesh(mm);
  fe::Grid &g = M->g;
  g.nx = mm->nx; g.ny = mm->ny; g.nz = mm->nz;
  g.ncx = g.nx - 1; g.ncy = g.ny - 1; g.ncz = g.nz - 1;
  g.G = (int64_t)g.nx * g.ny * g.nz;
  const int64_t ncells = (int64_t)g.ncx * g.ncy * g.ncz;
  M->nn = mm->nn; M->ne = mm->ne; M->ngroups = mm->ngroups; M->nsec = mm->nsections;
  M->xs.assign(mm->xs, mm->xs + g.nx);
  M->ys.assign(mm->ys, mm->ys + g.ny);
  M->zs.assign(mm->zs, mm->zs + g.nz);
  M->node_id.assign(mm->node_id, mm->node_id + g.G);
  M->cell_elems.assign(mm->cell_elems, mm->cell_elems + 2 * ncells);
  M->e_group.assign(mm->e_group, mm->e_group + M->ne);
  M->e_part.assign(mm->e_part, mm->e_part + M->ne);
  M->e_scale.assign(mm->e_scale, mm->e_scale + M->ne);
  if (mm->e_section) M->e_section.assign(mm->e_section, mm->e_section + M->ne);
  else M->e_section.assign(M->ne, -1);
  M->constraints.assign(mm->constraint_dofs, mm->constraint_dofs + mm->nconstraints);
  if (mm->nmodule_constraints > 0) {
    if (!mm->module_constraint_dofs) throw be::Error("module_constraint_dofs missing");
    M->modConstraints.assign(mm->module_constraint_dofs, mm->module_constraint_dofs + mm->nmodule_constraints);
  }
  if (mm->nfeature_x) M->fx.assign(mm->feature_x, mm->feature_x + mm->nfeature_x);
  if (mm->nfeature_y) M->fy.assign(mm->feature_y, mm->feature_y + mm->nfeature_y);
  // validate
  int32_t cnt = 0;
  for (int64_t q = 0; q < g.G; q++)
    if (M->node_id[q] >= 0) {
      if (M->node_id[q] != cnt) throw be::Error("node_id must number existing grid nodes 0..nn-1 in grid order");
      cnt++;
    }
  if (cnt != M->nn) throw be::Error("node_id count differs from nn");
  for (int32_t d : M->constraints)
    if (d < 0 || d >= 3 * M->nn) throw be::Error("constraint DOF out of range");
  for (int32_t d : M->modConstraints)
    if (d < 0 || d >= 3 * M->nn) throw be::Error("module constraint DOF out of range");
  for (int32_t e = 0; e < M->ne; e++) {
    if (M->e_group[e] < 0 || M->e_group[e] >= M->ngroups) throw be::Error("element group out of range");
    if (M->e_section[e] >= M->nsec) throw be::Error("element section out of range");
  }
  // cell of every element and its type
  std::vector<double> hxr, hyr, hzr;
  std::vector<int32_t> cx = size_classes(M->xs, hxr), cy = size_classes(M->ys, hyr), cz = size_classes(M->zs, hzr);
  M->nslot = M->ngroups + M->nsec;
  std::map<std::tuple<int32_t, int32_t, int32_t, int32_t>, int32_t> tmap;
  M->e_type.assign(M->ne, -1);
  std::vector<int32_t> cellType(2 * ncells, -1), cellElem(2 * ncells, -1);
  if (ownGroup >= M->ngroups) throw be::Error("own-type group out of range");
  M->ownGroup = ownGroup;
  std::vector<int64_t> ownSlots;  // cell slot 2c+q of every own-type element, in grid order
  for (int i = 0; i < g.ncx; i++)
    for (int j = 0; j < g.ncy; j++)
      for (int k = 0; k < g.ncz; k++) {
        int64_t c = ((int64_t)i * g.ncy + j) * g.ncz + k;
        for (int q = 0; q < 2; q++) {
          int32_t e = M->cell_elems[2 * c + q];
          if (e < 0) continue;
          if (e >= M->ne) throw be::Error("cell_elems entry out of range");
          if (ownGroup >= 0 && M->e_group[e] == ownGroup) {
            if (M->e_section[e] >= 0) throw be::Error("own-type elements cannot be layered");
            ownSlots.push_back(2 * c + q);
            cellElem[2 * c + q] = e;
            continue;
          }
          int32_t slot = M->e_section[e] >= 0 ? M->ngroups + M->e_section[e] : M->e_group[e];
          auto key = std::make_tuple(slot, cx[i], cy[j], cz[k]);
          auto it = tmap.find(key);
          int32_t t;
          if (it == tmap.end()) {
            t = (int32_t)M->typeSlot.size();
            tmap[key] = t;
            M->typeSlot.push_back(slot);
            M->typeH.push_back(hxr[cx[i]]);
            M->typeH.push_back(hyr[cy[j]]);
            M->typeH.push_back(hzr[cz[k]]);
          } else t = it->second;
          M->e_type[e] = t;
          cellType[2 * c + q] = t;
          cellElem[2 * c + q] = e;
          // every corner of an element's cell must exist
          for (int a = 0; a < 8; a++) {
            int64_t gn = ((int64_t)(i + fe::corner_li(a)) * g.ny + j + fe::corner_lj(a)) * g.nz + k + fe::corner_lk(a);
            if (M->node_id[gn] < 0) throw be::Error("element corner node missing in node_id");
          }
        }
      }
  // own-type elements after the shared types, in grid order of their cells
  M->ownType0 = (int32_t)M->typeSlot.size();
  M->ownElems.clear();
  M->ownCells.clear();
  for (int64_t cs : ownSlots) {
    const int64_t c = cs / 2;
    const int k = (int)(c % g.ncz), j = (int)((c / g.ncz) % g.ncy), i = (int)(c / ((int64_t)g.ncz * g.ncy));
    const int32_t e = cellElem[cs], t = (int32_t)M->typeSlot.size();
    M->typeSlot.push_back(M->e_group[e]);
    M->typeH.push_back(M->xs[i + 1] - M->xs[i]);
    M->typeH.push_back(M->ys[j + 1] - M->ys[j]);
    M->typeH.push_back(M->zs[k + 1] - M->zs[k]);
    M->e_type[e] = t;
    cellType[cs] = t;
    M->ownElems.push_back(e);
    M->ownCells.push_back(c);
    for (int a = 0; a < 8; a++) {
      int64_t gn = ((int64_t)(i + fe::corner_li(a)) * g.ny + j + fe::corner_lj(a)) * g.nz + k + fe::corner_lk(a);
      if (M->node_id[gn] < 0) throw be::Error("element corner node missing in node_id");
    }
  }
  M->ntypes = (int32_t)M->typeSlot.size();
  if (g.nz > FINE_MAXNZ) throw be::Error("more than 64 grid planes in z");
  // DOF mask and compact -> grid map
  std::vector<uint8_t> mask(3 * g.G, 0);
  std::vector<int64_t> gridOf(M->nn);
  for (int64_t q = 0; q < g.G; q++)
    if (M->node_id[q] >= 0) {
      gridOf[M->node_id[q]] = q;
      mask[3 * q] = mask[3 * q + 1] = mask[3 * q + 2] = 1;
    }
  for (int32_t d : M->constraints) mask[3 * gridOf[d / 3] + d % 3] = 0;
  M->maskAll = mask;
  for (int32_t d : M->modConstraints) M->modGridDofs.push_back(3 * gridOf[d / 3] + d % 3);
#if FCBGA_GPU
  BE_CHECK(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking));
#endif
  be::stream_t s = M->stream;
  M->cellType.upload(cellType, s);
  M->cellElem.upload(cellElem, s);
  {
    const char *ev = getenv("FCBGA_ELEM_OP");
    M->mfmaOp = ev ? std::string(ev) == "mfma" : ELEM_OP_MFMA_DEFAULT;
    if (M->mfmaOp) mop_setup(M, cellType);
  }
  M->d_egroup.upload(M->e_group, s);
  M->d_escale.upload(M->e_scale, s);
  M->d_typeSlot.upload(M->typeSlot, s);
  M->d_typeH.upload(M->typeH, s);
  M->dofmask.upload(mask, s);
  M->gridOf.upload(gridOf, s);
  // colour classes of the (i, j) columns: (i mod 2, j mod 2)
  std::vector<int32_t> cols[4];
  for (int i = 0; i < g.nx; i++)
    for (int j = 0; j < g.ny; j++) cols[(i & 1) * 2 + (j & 1)].push_back(i * g.ny + j);
  for (int c = 0; c < 4; c++) {
    M->fineColors[c].upload(cols[c], s);
    M->nFineColor[c] = (int32_t)cols[c].size();
  }
  // natural-coordinate tables for the thermal load: g (3*3*11) then gH (3*11)
  const Univ &U = univ();
  std::vector<double> tabs(U.g, U.g + 99);
  tabs.insert(tabs.end(), U.gH, U.gH + 33);
  M->tabs.upload(tabs, s);
  be::sync(s);
}

// ---------------------------------------------------------------------------
// per-step assembly on the device
// ---------------------------------------------------------------------------

void check_step(const fcbga_model *M, const fcbga_step *st) {
  if (!st || !st->D || !st->s0 || !st->born) throw be::Error("step data missing (D, s0, born)");
  if (M->nsec > 0 && (!st->secW || !st->secV || !st->sec_born))
    throw be::Error("step data missing for layered sections (secW, secV, sec_born)");
}

// captured iteration graphs hold buffer addresses: dropped when buffers move
void drop_graphs(fcbga_model *M) {
#if FCBGA_GPU
  for (auto &gr : M->graphs) (void)hipGraphExecDestroy(gr.second);
  M->graphs.clear();
#else
  (void)M;
#endif
}

// DOF mask of a step: the package's, or in a stacked-module stage
// (FEModel.moduleStageSet, src/core/09-assembly.js:427-440) the DOFs of the
// nodes that born elements touch, less the module constraints (the package's
// 3-2-1 and each module's support).  Uploaded only when the stage's set changes.
void stage_mask(fcbga_model *M, bool modStage, const std::vector<uint8_t> &born) {
  static const std::vector<uint8_t> all;
  const std::vector<uint8_t> &key = modStage ? born : all;
  if (M->maskBorn == key) return;
  if (!modStage) {
    M->dofmask.upload(M->maskAll, M->stream);
  } else {
    const fe::Grid g = M->g;
    std::vector<uint8_t> mask(M->maskAll.size(), 0);
    for (int ci = 0; ci < g.ncx; ci++)
      for (int cj = 0; cj < g.ncy; cj++)
        for (int ck = 0; ck < g.ncz; ck++) {
          const int64_t c = ((int64_t)ci * g.ncy + cj) * g.ncz + ck;
          bool on = false;
          for (int q = 0; q < 2; q++) {
            const int32_t e = M->cell_elems[2 * c + q];
            on = on || (e >= 0 && born[M->e_group[e]]);
          }
          if (!on) continue;
          for (int a = 0; a < 8; a++) {
            const int64_t gn = ((int64_t)(ci + fe::corner_li(a)) * g.ny + cj + fe::corner_lj(a)) * g.nz + ck + fe::corner_lk(a);
            for (int p = 0; p < 3; p++) mask[3 * gn + p] = M->maskAll[3 * gn + p];
          }
        }
    for (int64_t d : M->modGridDofs) mask[d] = 0;
    M->dofmask.upload(mask, M->stream);
  }
  M->maskBorn = key;
}

void assemble_step(fcbga_model *M, const fcbga_step *st) {
  be::stream_t s = M->stream;
  const fe::Grid g = M->g;
  const int64_t ncells = (int64_t)g.ncx * g.ncy * g.ncz;
  // stacked-module stage (FEModel.assembleStep, src/core/09-assembly.js:397-403, 423):
  // unborn elements are dropped, the stage's own constraint set holds the rest,
  // and the multigrid aggregates follow the born elements (a new hierarchy
  // whenever that set changes, LinearSolver._buildPC, src/core/10-solver.js:868-883)
  const bool modStage = st->stage < 0 && !M->modConstraints.empty();
  {
    const std::vector<uint8_t> born(st->born, st->born + M->ngroups);
    stage_mask(M, modStage, born);
    const std::vector<uint8_t> hk = modStage ? born : std::vector<uint8_t>();
    if (M->hierBorn != hk) {
      if (M->hierBuilt) drop_graphs(M);
      M->hierBuilt = false;
      M->hierBorn = hk;
    }
  }
  const double unbornScale = modStage ? 0.0 : 1e-6;
  // material slots: shape matrices and loads
  std::vector<double> Mall((size_t)M->nslot * 6 * 1089), load((size_t)M->nslot * 18, 0.0);
  std::vector<uint8_t> hasF(M->nslot, 0), layered(M->nslot, 0);
  for (int32_t q = 0; q < M->ngroups; q++) {
    shape_matrices(nullptr, st->D + 36 * q, Mall.data() + (size_t)6 * 1089 * q);
    hasF[q] = st->born[q] ? 1 : 0;
    if (hasF[q])
      for (int r = 0; r < 6; r++) load[18 * q + r] = st->s0[6 * q + r];
  }
  for (int32_t q = 0; q < M->nsec; q++) {
    int32_t slot = M->ngroups + q;
    shape_matrices(st->secW + 108 * q, nullptr, Mall.data() + (size_t)6 * 1089 * slot);
    hasF[slot] = st->sec_born[q] ? 1 : 0;
    layered[slot] = 1;
    if (hasF[slot])
      for (int r = 0; r < 18; r++) load[18 * slot + r] = st->secV[18 * q + r];
  }
  M->slotM.upload(Mall, s);
  M->slotLoad.upload(load, s);
  M->slotHasF.upload(hasF, s);
  M->slotLayered.upload(layered, s);
  std::vector<uint8_t> born(st->born, st->born + M->ngroups);
  M->d_born.upload(born, s);
  // condensed element tables, one per type
  M->Kc64.alloc((size_t)M->ntypes * 576);
  M->Kc32.alloc((size_t)M->ntypes * 576);
  M->fc64.alloc((size_t)M->ntypes * 24);
  M->typeHasF.alloc(M->ntypes);
  {
    const int32_t *tslot = M->d_typeSlot.p;
    const double *tH = M->d_typeH.p, *sM = M->slotM.p, *sL = M->slotLoad.p, *tb = M->tabs.p;
    const uint8_t *sF = M->slotHasF.p, *sLay = M->slotLayered.p;
    double *Kc = M->Kc64.p, *fc = M->fc64.p;
    float *Kc32 = M->Kc32.p;
    uint8_t *tF = M->typeHasF.p;
    be::pfor(M->ntypes, s, [=] __host__ __device__(int64_t t) {
      int32_t sl = tslot[t];
      fe::element_condense(tH[3 * t], tH[3 * t + 1], tH[3 * t + 2], sM + (int64_t)sl * 6 * 1089, sF[sl], sLay[sl],
                           sL + 18 * (int64_t)sl, tb, tb + 99, Kc + 576 * t, fc + 24 * t);
      for (int q = 0; q < 576; q++) Kc32[576 * t + q] = (float)Kc[576 * t + q];
      tF[t] = sF[sl];
    }, 64);
  }
  // own-type elements: the tables come from the caller (viscoplastic tangents)
  if (M->ownK64 && M->ntypes > M->ownType0) {
    const int32_t t0 = M->ownType0;
    const double *oK = M->ownK64;
    double *Kc = M->Kc64.p, *fc = M->fc64.p;
    float *Kc32 = M->Kc32.p;
    uint8_t *tF = M->typeHasF.p;
    be::pfor((int64_t)(M->ntypes - t0) * 576, s, [=] __host__ __device__(int64_t q) {
      const int64_t t = t0 + q / 576;
      const double v = oK[q];
      Kc[576 * t0 + q] = v;
      Kc32[576 * t0 + q] = (float)v;
      if (q % 576 < 24) fc[24 * t + q % 576] = 0.0;
      if (q % 576 == 0) tF[t] = 0;
    });
  }
  // mirror-symmetric compact tables (all shared types must pass, else the
  // full tables for every type)
  M->R64.alloc((size_t)M->ntypes * fe::SYM_TABLE);
  M->R32.alloc((size_t)M->ntypes * fe::SYM_TABLE);
  M->symErr.alloc(M->ntypes);
  {
    const double *Kc = M->Kc64.p;
    double *R = M->R64.p, *err = M->symErr.p;
    float *R32 = M->R32.p;
    be::pfor(M->ntypes, s, [=] __host__ __device__(int64_t t) {
      double Rf[72];
      err[t] = fe::sym_table(Kc + 576 * t, Rf);
      fe::sym_compact(Rf, R + fe::SYM_TABLE * t);
      for (int q = 0; q < fe::SYM_TABLE; q++) R32[fe::SYM_TABLE * t + q] = (float)R[fe::SYM_TABLE * t + q];
    }, 64);
    // the own types (viscoplastic tangents) keep their full tables beside
    // the compact ones, so only the shared types must pass
    const int32_t nshared = std::min(M->ntypes, M->ownType0);
    std::vector<double> e(nshared);
    if (nshared) be::d2h(e.data(), M->symErr.p, e.size() * sizeof(double), s);
    be::sync(s);
    double emax = 0;
    for (double v : e) emax = v > emax ? v : emax;
    M->symTables = M->sym_allowed && emax <= SYM_TOL;
  }
  // element scales of the step: eScale x (born ? 1 : 1e-6), unborn 0 in a module stage
  M->cellScale64.alloc(2 * ncells);
  M->cellScale32.alloc(2 * ncells);
  {
    const int32_t *ce = M->cellElem.p, *eg = M->d_egroup.p;
    const double *es = M->d_escale.p;
    const uint8_t *bn = M->d_born.p;
    double *cs = M->cellScale64.p;
    float *cs32 = M->cellScale32.p;
    be::pfor(2 * ncells, s, [=] __host__ __device__(int64_t q) {
      int32_t e = ce[q];
      double v = e >= 0 ? es[e] * (bn[eg[e]] ? 1.0 : unbornScale) : 0.0;
      cs[q] = v;
      cs32[q] = (float)v;
    });
  }
  // load vector and diagonal / column blocks; Jacobi scaling
  const int64_t G = g.G;
  M->bload.alloc(3 * G);
  M->s64.alloc(3 * G); M->w64.alloc(3 * G); M->sinv64.alloc(3 * G);
  M->w32.alloc(3 * G); M->sinv32.alloc(3 * G);
  M->Db.alloc(9 * G); M->Lb.alloc(9 * G);
  {
    const int32_t *ct = M->cellType.p;
    const double *cs = M->cellScale64.p, *Kc = M->Kc64.p, *fc = M->fc64.p;
    const uint8_t *tF = M->typeHasF.p, *mask = M->dofmask.p;
    double *bl = M->bload.p, *sv = M->s64.p, *wv = M->w64.p, *si = M->sinv64.p, *Db = M->Db.p, *Lb = M->Lb.p;
    float *w32 = M->w32.p, *si32 = M->sinv32.p;
    be::pfor(G, s, [=] __host__ __device__(int64_t gi) {
      int k = (int)(gi % g.nz);
      int64_t cj = gi / g.nz;
      int j = (int)(cj % g.ny), i = (int)(cj / g.ny);
      double f[3], D[9], L[9];
      fe::node_load(g, i, j, k, ct, cs, fc, tF, f);
      fe::node_blocks(g, i, j, k, ct, cs, Kc, D, L);
      bool exists = D[0] != 0 || D[4] != 0 || D[8] != 0;
      double sc[3];
      for (int p = 0; p < 3; p++) {
        double d = D[4 * p];
        sc[p] = exists ? 1 / sqrt(d > 1e-300 ? d : 1e-300) : 0.0;
        uint8_t m = mask[3 * gi + p];
        bl[3 * gi + p] = m ? f[p] : 0.0;
        sv[3 * gi + p] = sc[p];
        wv[3 * gi + p] = m ? sc[p] : 0.0;
        w32[3 * gi + p] = (float)(m ? sc[p] : 0.0);
        si[3 * gi + p] = exists ? 1 / sc[p] : 0.0;
        si32[3 * gi + p] = (float)(exists ? 1 / sc[p] : 0.0);
      }
      // scaled, constraint-reduced blocks of the column smoother
      const int64_t gd = gi - 1;  // node below (same column when k > 0)
      for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) {
          uint8_t mp = mask[3 * gi + p], mq = mask[3 * gi + q];
          double v = (mp && mq) ? sc[p] * D[3 * p + q] * sc[q] : 0.0;
          if (p == q && !mp) v = 1.0;
          Db[9 * gi + 3 * p + q] = v;
          double lv = 0.0;
          if (k > 0) {
            uint8_t mb = mask[3 * gd + q];
            // the scale of the node below is applied by the next kernel
            lv = (mp && mb) ? sc[p] * L[3 * p + q] : 0.0;
          }
          Lb[9 * gi + 3 * p + q] = lv;
        }
    });
    // finish the column blocks with the scale of the node below
    be::pfor(G, s, [=] __host__ __device__(int64_t gi) {
      int k = (int)(gi % g.nz);
      if (k == 0) return;
      for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) Lb[9 * gi + 3 * p + q] *= sv[3 * (gi - 1) + q];
    });
  }
}

// column factors of the fine level (FP64) and the scan coefficients of the
// column solve formed from them (FP64, FP32 copy)
void factor_fine_columns(fcbga_model *M) {
  const fe::Grid g = M->g;
  const int64_t ncol = (int64_t)g.nx * g.ny;
  constexpr int NC = fe::SCAN_COEFS;
  M->F64.alloc(18 * g.G);
  M->C64.alloc(NC * g.G);
  M->C32.alloc(NC * g.G);
  const double *Db = M->Db.p, *Lb = M->Lb.p;
  double *F = M->F64.p, *C = M->C64.p;
  float *C32 = M->C32.p;
  const int nz = g.nz;
  be::pfor(ncol, M->stream, [=] __host__ __device__(int64_t c) {
    fe::column_factor(nz, Db + 9 * c * nz, Lb + 9 * c * nz, F + 18 * c * nz);
  }, 64);
  // the coefficients of every node at once (each needs its own factor and the
  // sub-diagonal block of the node above)
  be::pfor(g.G, M->stream, [=] __host__ __device__(int64_t gi) {
    const int k = (int)(gi % nz);
    const int64_t c = gi / nz;
    double v[NC];
    fe::node_scan_coefs_at(k, nz, F + 18 * gi, v);
    for (int q = 0; q < NC; q++) {
      C[fe::scan_coef_index<double>(c, k, nz, q)] = v[q];
      C32[fe::scan_coef_index<float>(c, k, nz, q)] = (float)v[q];
    }
  });
}

// ---------------------------------------------------------------------------
// fine-level operator pieces
// ---------------------------------------------------------------------------

#if FCBGA_GPU
// Block order across the 8 XCDs: the hardware hands out workgroups round-robin,
// so block b runs on XCD b % 8.  Giving each XCD a contiguous run of columns
// keeps neighbouring columns (whose neighbourhoods overlap) in one L2.
__device__ __forceinline__ int xcd_block(int b, int nblocks) {
  const int per = (nblocks + 7) >> 3;
  return (b & 7) * per + (b >> 3);
}
#endif

}  // namespace
#include "mfma_op.h"
namespace {

// the element groups of the matrix-core operator, from the cells' types
void mop_setup(fcbga_model *M, const std::vector<int32_t> &cellType) {
  mop::Lists L;
  mop::build_lists(M->g, cellType, L);
  M->mopTiles[0] = L.ntx; M->mopTiles[1] = L.nty; M->mopTiles[2] = L.nchunk;
  M->mopType.upload(L.type, M->stream);
  M->mopOff.upload(L.off, M->stream);
  M->mopCell.upload(L.cell, M->stream);
  M->mopMask.upload(L.mask, M->stream);
}

template <typename T>
struct FineOps {
  fe::Grid g;
  const int32_t *cellType;
  const T *cellScale, *Kc, *w, *sinv;
  const T *C;  // column scan coefficients, fe::SCAN_COEFS per node
  const int32_t *nodeAgg;
  const T *zp;
  const T *R;  // compact mirror-symmetric tables (fe::SYM_TABLE per type) when sym
  bool sym;
  // with the compact tables, the types from own0 on (the viscoplastic
  // solder's own tangents; INT32_MAX: none) take the full tables
  int own0;
  // element groups of the matrix-core operator (mop::k_elem_mfma) when mfma
  bool mfma;
  int ntx, nty, nchunk;
  const int32_t *gType, *gOff;
  const int16_t *gCell;
  const uint32_t *gMask;
};

template <typename T>
FineOps<T> fine_ops(fcbga_model *M);

// first own type of the model, INT32_MAX when it has none
int own_from(const fcbga_model *M) { return M->ownType0 < M->ntypes ? M->ownType0 : INT32_MAX; }

template <>
FineOps<double> fine_ops<double>(fcbga_model *M) {
  return FineOps<double>{M->g, M->cellType.p, M->cellScale64.p, M->Kc64.p, M->w64.p, M->sinv64.p, M->C64.p,
                         M->nodeAgg.p, M->zp64.p, M->R64.p, M->symTables, own_from(M), M->mfmaOp, M->mopTiles[0],
                         M->mopTiles[1], M->mopTiles[2], M->mopType.p, M->mopOff.p, M->mopCell.p, M->mopMask.p};
}
template <>
FineOps<float> fine_ops<float>(fcbga_model *M) {
  return FineOps<float>{M->g, M->cellType.p, M->cellScale32.p, M->Kc32.p, M->w32.p, M->sinv32.p, M->C32.p,
                        M->nodeAgg.p, M->zp32.p, M->R32.p, M->symTables, own_from(M), M->mfmaOp, M->mopTiles[0],
                        M->mopTiles[1], M->mopTiles[2], M->mopType.p, M->mopOff.p, M->mopCell.p, M->mopMask.p};
}

#if FCBGA_GPU
// FORM: 0 full tables (Kc), 1 compact tables (Kc = R), 2 compact tables with
// the full tables Kf for the types from own0 on
template <typename T, int MODE, int FORM, bool PACK>
__global__ void __launch_bounds__(64) k_fine_col_op(fe::Grid g, const int32_t *__restrict__ ct,
                                                    const T *__restrict__ cs, const T *__restrict__ Kc,
                                                    const T *__restrict__ Kf, int own0, const T *__restrict__ w,
                                                    const T *__restrict__ x, const T *__restrict__ bvec,
                                                    T *__restrict__ y);

// columns per wavefront of the column kernels: 64 / nz for grids of at most
// 32 planes, else one
inline int cols_per_wave(int nz) { return nz <= 32 ? 64 / nz : 1; }

// the table form of an operator launch: 0 full, 1 compact, 2 compact with own full tables
template <typename T>
int table_form(const FineOps<T> &o) {
  if (!o.sym) return 0;
  return o.own0 < INT32_MAX ? 2 : 1;
}

template <typename T, int MODE>
void launch_col_op(const FineOps<T> &o, const T *x, const T *b, T *y, be::stream_t s) {
  const fe::Grid g = o.g;
  if (o.mfma) {
    const int nt = o.ntx * o.nty * o.nchunk;
    hipLaunchKernelGGL((mop::k_elem_mfma<T, MODE>), dim3(8 * ((nt + 7) / 8)), dim3(mop::THREADS), 0, s, g, o.ntx,
                       o.nty, o.nchunk, o.gType, o.gCell, o.gOff, o.gMask, o.cellScale, o.Kc, o.w, x, b, y);
    BE_CHECK(hipGetLastError());
    return;
  }
  const int P = cols_per_wave(g.nz);
  const int nb = (g.nx * g.ny + P - 1) / P;
  const dim3 grid(8 * ((nb + 7) / 8));
  const size_t lds = (size_t)P * 9 * (g.nz + 2) * 3 * sizeof(T);  // the 3 x 3 column neighbourhoods
  const T *K0 = table_form(o) ? o.R : o.Kc;
  const auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, dim3(64), lds, s, g, o.cellType, o.cellScale, K0, o.Kc, o.own0, o.w, x, b, y);
  };
  if (P > 1) {
    switch (table_form(o)) {
      case 2: go(k_fine_col_op<T, MODE, 2, true>); break;
      case 1: go(k_fine_col_op<T, MODE, 1, true>); break;
      default: go(k_fine_col_op<T, MODE, 0, true>);
    }
  } else {
    switch (table_form(o)) {
      case 2: go(k_fine_col_op<T, MODE, 2, false>); break;
      case 1: go(k_fine_col_op<T, MODE, 1, false>); break;
      default: go(k_fine_col_op<T, MODE, 0, false>);
    }
  }
  BE_CHECK(hipGetLastError());
}
#endif

// y = A x on every grid node (A the scaled, constraint-reduced operator)
template <typename T>
void fine_apply(const FineOps<T> &o, const T *x, T *y, be::stream_t s) {
  const fe::Grid g = o.g;
  const int32_t *ct = o.cellType;
  const T *cs = o.cellScale, *Kc = o.Kc, *w = o.w;
#if FCBGA_GPU
  (void)ct; (void)cs; (void)Kc; (void)w;
  launch_col_op<T, 0>(o, x, (const T *)nullptr, y, s);
#else
  if (o.mfma) {
    be::sync(s);
    mop::elem_op_host<T, 0>(g, o.ntx, o.nty, o.nchunk, o.gType, o.gCell, o.gOff, o.gMask, cs, Kc, w, x,
                            (const T *)nullptr, y);
    return;
  }
  const T *R = o.R;
  const bool sym = o.sym;
  const int own0 = o.own0;
  be::pfor(g.G, s, [=] __host__ __device__(int64_t gi) {
    int k = (int)(gi % g.nz);
    int64_t cj = gi / g.nz;
    int j = (int)(cj % g.ny), i = (int)(cj / g.ny);
    T out[3];
    if (sym)
      fe::node_apply_sym<T>(g, i, j, k, ct, cs, fe::SymTabs<T>{R, Kc, own0}, w, x, out);
    else
      fe::node_apply<T, T>(g, i, j, k, ct, cs, Kc, w, x, out);
    y[3 * gi] = out[0]; y[3 * gi + 1] = out[1]; y[3 * gi + 2] = out[2];
  });
#endif
}

// res = b - A x on every grid node
template <typename T>
void fine_residual(const FineOps<T> &o, const T *b, const T *x, T *res, be::stream_t s) {
  const fe::Grid g = o.g;
  const int32_t *ct = o.cellType;
  const T *cs = o.cellScale, *Kc = o.Kc, *w = o.w;
#if FCBGA_GPU
  (void)ct; (void)cs; (void)Kc; (void)w;
  launch_col_op<T, 1>(o, x, b, res, s);
#else
  if (o.mfma) {
    be::sync(s);
    mop::elem_op_host<T, 1>(g, o.ntx, o.nty, o.nchunk, o.gType, o.gCell, o.gOff, o.gMask, cs, Kc, w, x, b, res);
    return;
  }
  const T *R = o.R;
  const bool sym = o.sym;
  const int own0 = o.own0;
  be::pfor(g.G, s, [=] __host__ __device__(int64_t gi) {
    int k = (int)(gi % g.nz);
    int64_t cj = gi / g.nz;
    int j = (int)(cj % g.ny), i = (int)(cj / g.ny);
    T out[3];
    if (sym)
      fe::node_apply_sym<T, true>(g, i, j, k, ct, cs, fe::SymTabs<T>{R, Kc, own0}, w, x, out);
    else
      fe::node_apply<T, T, true>(g, i, j, k, ct, cs, Kc, w, x, out);
    for (int p = 0; p < 3; p++) res[3 * gi + p] = b[3 * gi + p] - out[p];
  });
#endif
}

// One colour of a column block Gauss-Seidel sweep (columnSweep,
// src/core/10-solver.js:512-555): for every column of the colour, residual
// r = b - A x over its nodes, block-tridiagonal solve (solveColumn,
// src/core/10-solver.js:71-104), x += dx.  Columns of one colour share no
// element, so they update independently.  GPU: one 64-lane block per column;
// the lanes compute the residual rows and stage the column factor in LDS, one
// lane runs the substitution chain, and the lanes apply the update.
#if FCBGA_GPU
// One round of an inclusive scan of affine maps (A, b) at distance s: lane k
// composes its map with that of lane k - s (up) or k + s (down).  Every lane
// updates: a lane without a partner at distance s holds a composition that
// includes the column end's zero map (P at the bottom node, Q at the top one,
// zero coefficients past the column), so its A is exactly zero and the update
// (with the finite values the shuffle returns) leaves its b unchanged.  The
// maps' products are only needed by later rounds (needA false in the last).
template <typename T, bool UP>
__device__ __forceinline__ void scan_round(T A[9], T b[3], int s, bool needA) {
  T An[9], bn[3];
  for (int p = 0; p < 3; p++) bn[p] = UP ? __shfl_up(b[p], s, 64) : __shfl_down(b[p], s, 64);
  T rb[3];
  for (int i = 0; i < 3; i++) rb[i] = A[3 * i] * bn[0] + A[3 * i + 1] * bn[1] + A[3 * i + 2] * bn[2] + b[i];
  if (needA) {
    for (int q = 0; q < 9; q++) An[q] = UP ? __shfl_up(A[q], s, 64) : __shfl_down(A[q], s, 64);
    T R[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) R[3 * i + j] = A[3 * i] * An[j] + A[3 * i + 1] * An[3 + j] + A[3 * i + 2] * An[6 + j];
    for (int q = 0; q < 9; q++) A[q] = R[q];
  }
  for (int p = 0; p < 3; p++) b[p] = rb[p];
}

// the SCAN_COEFS coefficients of node (column c, plane k) from the device
// layout (fe::scan_coef_index): slot v of a column is one contiguous run of
// 16-byte vectors over the planes, so each load instruction of a wavefront
// reads one run
template <typename T>
__device__ __forceinline__ void load_scan_coefs(const T *__restrict__ C, int c, int k, int nz, bool valid,
                                                T out[fe::SCAN_COEFS]);
template <>
__device__ __forceinline__ void load_scan_coefs<float>(const float *__restrict__ C, int c, int k, int nz, bool valid,
                                                        float out[fe::SCAN_COEFS]) {
  constexpr int NV = fe::SCAN_COEFS / 4;
  const float4 *p = reinterpret_cast<const float4 *>(C) + (int64_t)c * NV * nz + (valid ? k : 0);
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const float4 q = p[(int64_t)v * nz];
    out[4 * v] = valid ? q.x : 0.f; out[4 * v + 1] = valid ? q.y : 0.f;
    out[4 * v + 2] = valid ? q.z : 0.f; out[4 * v + 3] = valid ? q.w : 0.f;
  }
}
template <>
__device__ __forceinline__ void load_scan_coefs<double>(const double *__restrict__ C, int c, int k, int nz, bool valid,
                                                         double out[fe::SCAN_COEFS]) {
  constexpr int NV = fe::SCAN_COEFS / 2;
  const double2 *p = reinterpret_cast<const double2 *>(C) + (int64_t)c * NV * nz + (valid ? k : 0);
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const double2 q = p[(int64_t)v * nz];
    out[2 * v] = valid ? q.x : 0.0; out[2 * v + 1] = valid ? q.y : 0.0;
  }
}

// lane k = grid plane k of the column (nz <= 64): the w*x neighbourhood of the
// column is staged in LDS, then the residual row, both substitution
// recurrences as wavefront scans in registers, then x += dx
// PACK (grids of at most 32 planes): 64 / nz columns per wavefront, lane
// q * nz + k = plane k of the block's column q, each column's neighbourhood in
// its own slice of the dynamic LDS (plane stride nz + 2); the scans stay
// within a column because its end nodes carry zero maps (P at the bottom, Q
// at the top), which the compositions across a column boundary multiply by.
template <typename T, int FORM, bool PACK>
__global__ void __launch_bounds__(64) k_fine_col_sweep(fe::Grid g, const int32_t *__restrict__ cols, int ncols,
                                                       const int32_t *__restrict__ ct, const T *__restrict__ cs,
                                                       const T *__restrict__ Kc, const T *__restrict__ Kf, int own0,
                                                       const T *__restrict__ w, const T *__restrict__ C,
                                                       const T *__restrict__ bvec, T *__restrict__ x) {
  constexpr bool SYM = FORM != 0;
  __shared__ T nb_fix[PACK ? 1 : 9 * fe::NZP * 3];  // fixed 64-plane stride
  extern __shared__ double nb_dyn[];
  const int nz = g.nz;
  const int P = PACK ? 64 / nz : 1;
  const int nzp = PACK ? nz + 2 : fe::NZP;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  if (blk * P >= ncols) return;
  const int slot = PACK ? threadIdx.x / nz : 0;
  const int k = PACK ? threadIdx.x % nz : threadIdx.x;
  const int t = blk * P + slot;
  const bool valid = PACK ? (slot < P && t < ncols) : k < nz;
  const int c = cols[valid ? t : blk * P];
  const int i = c / g.ny, j = c % g.ny;
  T *nb = PACK ? reinterpret_cast<T *>(nb_dyn) + (slot < P ? slot : 0) * 9 * nzp * 3 : nb_fix;
  const int gi = c * nz + (valid ? k : 0);
  // this node's own data first: their latency overlaps the staging below
  // (the scan coefficients are loaded after the residual, which needs the
  // registers)
  T Ck[fe::SCAN_COEFS], wv[3], xv[3], bv[3];
  for (int p = 0; p < 3; p++) {
    wv[p] = w[3 * gi + p];
    xv[p] = x[3 * gi + p];
    bv[p] = bvec[3 * gi + p];
  }
  // the types and scales of the node's four cells, with the staging loads
  fe::CellQuad<T> cq;
  if (SYM) fe::load_cell_quad<T>(g, i, j, k, valid, ct, cs, cq);
  {
    T *nb0 = PACK ? reinterpret_cast<T *>(nb_dyn) : nb_fix;
    for (int q = threadIdx.x; q < P * 9 * 2 * 3; q += 64) {  // padding levels -1 and nz of every column
      const int sl = q / 54, col = (q % 54) / 6, side = (q / 3) % 2, p = q % 3;
      nb0[((sl * 9 + col) * nzp + (side ? nz + 1 : 0)) * 3 + p] = T(0);
    }
  }
  if (valid) fe::fill_neighbourhood<T, true>(g, i, j, k, w, x, nb, nzp);
  __syncthreads();
  T r[3] = {0, 0, 0};
  T acc[3] = {0, 0, 0};
  if (SYM) {
    T up[3] = {0, 0, 0};
    const fe::SymTabs<T> tb{Kc, Kf, own0};
    // wavefronts whose nodes all see one table in their four cells load it
    // once (the scan coefficients stay behind the residual: loading them here
    // as well costs a wave per SIMD in registers, measured slower)
    if (__all(!valid || fe::cell_quad_uniform<T, FORM == 2>(cq, tb))) {
      if (valid) fe::node_rows_nb_sym_uni<T>(k, cq, tb, nb, acc, up, nzp);
    } else {
      const int pr = valid ? fe::cell_quad_pairs<T, FORM == 2>(cq, tb) : 1;
      if (__all(pr == 1)) {
        if (valid) fe::node_rows_nb_sym_pairs<T, 1>(k, cq, tb, nb, acc, up, nzp);
      } else if (__all(!valid || pr == 2)) {
        if (valid) fe::node_rows_nb_sym_pairs<T, 2>(k, cq, tb, nb, acc, up, nzp);
      } else if (valid) {
        fe::node_rows_nb_sym_q<T, FORM == 2>(k, cq, tb, nb, acc, up, nzp);
      }
    }
    for (int p = 0; p < 3; p++) {
      const T u = __shfl_up(up[p], 1, 64);
      acc[p] += k > 0 ? u : T(0);
    }
  } else if (valid) {
    fe::node_rows_nb<T>(g, i, j, k, ct, cs, Kc, nb, acc, nzp);
  }
  if (valid)
    for (int p = 0; p < 3; p++) r[p] = bv[p] - wv[p] * acc[p];
  // forward recurrence y_k = P_k y_{k-1} + Li_k r_k, then the backward one
  // dx_k = Q_k dx_{k+1} + Li_k^T y_k, each as a scan over the 64 lanes in six
  // rounds (the coefficients of lanes past the column are zero; rounds at
  // distances beyond nz change nothing)
  load_scan_coefs<T>(C, c, k, nz, valid, Ck);
  // Li is lower triangular: Ck[0..5] = Li00, Li10, Li11, Li20, Li21, Li22
  T A[9], b[3];
  for (int q = 0; q < 9; q++) A[q] = Ck[6 + q];
  b[0] = Ck[0] * r[0];
  b[1] = Ck[1] * r[0] + Ck[2] * r[1];
  b[2] = Ck[3] * r[0] + Ck[4] * r[1] + Ck[5] * r[2];
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) scan_round<T, true>(A, b, s, s < 32);
  {
    const T y0 = b[0], y1 = b[1], y2 = b[2];
    for (int q = 0; q < 9; q++) A[q] = Ck[15 + q];
    b[0] = Ck[0] * y0 + Ck[1] * y1 + Ck[3] * y2;
    b[1] = Ck[2] * y1 + Ck[4] * y2;
    b[2] = Ck[5] * y2;
  }
#pragma unroll
  for (int s = 1; s < 64; s <<= 1) scan_round<T, false>(A, b, s, s < 32);
  if (valid)
    for (int p = 0; p < 3; p++) x[3 * gi + p] = xv[p] + wv[p] * b[p];
}

// y = A x (mode 0) or y = b - A x (mode 1) over every node column, one 64-lane
// block per column with the neighbourhood staged in LDS
// (with PACK, 64 / nz columns per block as in k_fine_col_sweep)
template <typename T, int MODE, int FORM, bool PACK>
__global__ void __launch_bounds__(64) k_fine_col_op(fe::Grid g, const int32_t *__restrict__ ct,
                                                    const T *__restrict__ cs, const T *__restrict__ Kc,
                                                    const T *__restrict__ Kf, int own0, const T *__restrict__ w,
                                                    const T *__restrict__ x, const T *__restrict__ bvec,
                                                    T *__restrict__ y) {
  constexpr bool SYM = FORM != 0;
  extern __shared__ double nb_lds[];  // P * 9 * (nz + 2) * 3 values of T
  const int nz = g.nz, ncol = g.nx * g.ny;
  const int P = PACK ? 64 / nz : 1;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  if (blk * P >= ncol) return;
  const int slot = PACK ? threadIdx.x / nz : 0;
  const int k = PACK ? threadIdx.x % nz : threadIdx.x;
  const int c = blk * P + slot;
  const bool valid = PACK ? (slot < P && c < ncol) : k < nz;
  const int cc = valid ? c : blk * P;
  const int i = cc / g.ny, j = cc % g.ny;
  T *nb = reinterpret_cast<T *>(nb_lds) + (slot < P ? slot : 0) * 9 * (nz + 2) * 3;
  const int gi = cc * nz + k;
  fe::CellQuad<T> cq;
  if (SYM) fe::load_cell_quad<T>(g, i, j, k, valid, ct, cs, cq);
  for (int q = threadIdx.x; q < P * 9 * 2 * 3; q += 64) {
    const int sl = q / 54, col = (q % 54) / 6, side = (q / 3) % 2, p = q % 3;
    reinterpret_cast<T *>(nb_lds)[((sl * 9 + col) * (nz + 2) + (side ? nz + 1 : 0)) * 3 + p] = T(0);
  }
  if (valid) fe::fill_neighbourhood<T, MODE == 1>(g, i, j, k, w, x, nb, nz + 2);
  __syncthreads();
  T acc[3] = {0, 0, 0};
  if (SYM) {
    T up[3] = {0, 0, 0};
    const fe::SymTabs<T> tb{Kc, Kf, own0};
    if (__all(!valid || fe::cell_quad_uniform<T, FORM == 2>(cq, tb))) {
      if (valid) fe::node_rows_nb_sym_uni<T>(k, cq, tb, nb, acc, up, nz + 2);
    } else {
      const int pr = valid ? fe::cell_quad_pairs<T, FORM == 2>(cq, tb) : 1;
      if (__all(pr == 1)) {
        if (valid) fe::node_rows_nb_sym_pairs<T, 1>(k, cq, tb, nb, acc, up, nz + 2);
      } else if (__all(!valid || pr == 2)) {
        if (valid) fe::node_rows_nb_sym_pairs<T, 2>(k, cq, tb, nb, acc, up, nz + 2);
      } else if (valid) {
        fe::node_rows_nb_sym_q<T, FORM == 2>(k, cq, tb, nb, acc, up, nz + 2);
      }
    }
    for (int p = 0; p < 3; p++) {
      const T u = __shfl_up(up[p], 1, 64);
      acc[p] += k > 0 ? u : T(0);
    }
  } else if (valid) {
    fe::node_rows_nb<T>(g, i, j, k, ct, cs, Kc, nb, acc, nz + 2);
  }
  if (!valid) return;
  for (int p = 0; p < 3; p++) {
    const T wi = w[3 * gi + p];
    const T out = wi != T(0) ? wi * acc[p] : (MODE ? T(0) : x[3 * gi + p]);
    y[3 * gi + p] = MODE ? bvec[3 * gi + p] - out : out;
  }
}
#endif

template <typename T>
void fine_color_sweep(fcbga_model *M, const FineOps<T> &o, const int32_t *cols, int32_t ncols, const T *b, T *x, T *r,
                      T *dx, be::stream_t s) {
  const fe::Grid g = o.g;
  const int32_t *ct = o.cellType;
  const T *cs = o.cellScale, *Kc = o.Kc, *w = o.w, *C = o.C;
  const int nz = g.nz;
  if (!ncols) return;
  KScope ks(M, FCBGA_KCLASS_FINE_SWEEP);
#if FCBGA_GPU
  (void)r; (void)dx;
  const int P = cols_per_wave(nz);
  const int nblk = (ncols + P - 1) / P;
  const dim3 grid(8 * ((nblk + 7) / 8));
  const size_t lds = P > 1 ? (size_t)P * 9 * (nz + 2) * 3 * sizeof(T) : 0;
  const T *K0 = table_form(o) ? o.R : Kc;
  const auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, grid, dim3(64), lds, s, g, cols, ncols, ct, cs, K0, Kc, o.own0, w, C, b, x);
  };
  if (P > 1) {
    switch (table_form(o)) {
      case 2: go(k_fine_col_sweep<T, 2, true>); break;
      case 1: go(k_fine_col_sweep<T, 1, true>); break;
      default: go(k_fine_col_sweep<T, 0, true>);
    }
  } else {
    switch (table_form(o)) {
      case 2: go(k_fine_col_sweep<T, 2, false>); break;
      case 1: go(k_fine_col_sweep<T, 1, false>); break;
      default: go(k_fine_col_sweep<T, 0, false>);
    }
  }
  BE_CHECK(hipGetLastError());
#else
  const T *Rt = o.R;
  const bool symT = o.sym;
  const int own0 = o.own0;
  be::pfor(ncols, s, [=] __host__ __device__(int64_t t) {
    const int c = cols[t];
    const int i = c / g.ny, j = c % g.ny;
    T nb[9 * fe::NZP * 3];
    for (int q = 0; q < 9 * fe::NZP * 3; q++) nb[q] = T(0);
    for (int k = 0; k < nz; k++) fe::fill_neighbourhood<T, true>(g, i, j, k, w, x, nb, fe::NZP);
    T upPrev[3] = {0, 0, 0};
    for (int k = 0; k < nz; k++) {
      const int gi = c * nz + k;
      T acc[3];
      if (symT) {
        T up[3];
        fe::node_rows_nb_sym<T, true>(g, i, j, k, ct, cs, fe::SymTabs<T>{Rt, Kc, own0}, nb, acc, up, fe::NZP);
        for (int p = 0; p < 3; p++) { acc[p] += upPrev[p]; upPrev[p] = up[p]; }
      } else {
        fe::node_rows_nb<T>(g, i, j, k, ct, cs, Kc, nb, acc, fe::NZP);
      }
      for (int p = 0; p < 3; p++) {
        const T wi = w[3 * gi + p];
        r[3 * gi + p] = b[3 * gi + p] - wi * acc[p];
      }
    }
  });
  be::pfor(ncols, s, [=] __host__ __device__(int64_t t) {
    int64_t c = cols[t];
    fe::column_solve_scan<T, T>(nz, C, c, r + 3 * c * nz, dx + 3 * c * nz);
    for (int q = 0; q < 3 * nz; q++) x[3 * c * nz + q] += w[3 * c * nz + q] * dx[3 * c * nz + q];
  }, 64);
#endif
}

// ---------------------------------------------------------------------------
// multigrid numeric set-up
// ---------------------------------------------------------------------------

// Index lists of the two-phase Galerkin product lv[l] -> lv[l+1] (see
// DevShell::gal2).  T(f, J) = sum over the blocks (f, g) of row f and the
// parents J of g of A_fg W_g,J; A(I, J) = sum over the children f of I of
// W_f,I^T T(f, J), for J in row I's pattern.  The T blocks of a fine row are
// the distinct parents of its columns in first-seen order.
constexpr int GAL2_MAXT = 96;   // T blocks of one fine row held in LDS
constexpr int GAL_MAXROW = 64;  // blocks of one coarse row held in LDS

void build_gal2(fcbga_model *M, int l, be::stream_t s) {
  const mgs::Level &F = M->H.L[l], &Cl = M->H.L[l + 1];
  DevShell &C = *M->lv[l + 1];
  C.gal2 = false;
  if (C.maxrow > GAL_MAXROW) return;
  const int32_t nf = F.nagg, nc = Cl.nagg;
  std::vector<int32_t> tPtr(nf + 1, 0), tsPtr(nf + 1, 0), tCol, tSlot;
  std::vector<int32_t> mark(nc, -1), pos(nc, 0);
  int maxT = 0;
  for (int32_t f = 0; f < nf; f++) {
    const int32_t t0 = (int32_t)tCol.size();
    for (int32_t k = F.rowPtr[f]; k < F.rowPtr[f + 1]; k++) {
      const int32_t gg = F.col[k];
      for (int32_t pp = F.pPtr[gg]; pp < F.pPtr[gg + 1]; pp++) {
        const int32_t J = F.pIdx[pp];
        if (mark[J] != f) {
          mark[J] = f;
          pos[J] = (int32_t)tCol.size() - t0;
          tCol.push_back(J);
        }
        tSlot.push_back(pos[J]);
      }
    }
    tPtr[f + 1] = (int32_t)tCol.size();
    tsPtr[f + 1] = (int32_t)tSlot.size();
    maxT = std::max(maxT, tPtr[f + 1] - t0);
    if (tSlot.size() > (size_t)INT32_MAX / 2) return;
  }
  if (maxT > GAL2_MAXT) return;
  std::vector<int32_t> csPtr(nc + 1, 0), cSlot, rmark(nc, -1), rpos(nc, 0);
  for (int32_t I = 0; I < nc; I++) {
    for (int32_t q = Cl.rowPtr[I]; q < Cl.rowPtr[I + 1]; q++) {
      rmark[Cl.col[q]] = I;
      rpos[Cl.col[q]] = q - Cl.rowPtr[I];
    }
    for (int32_t h = Cl.chPtr[I]; h < Cl.chPtr[I + 1]; h++) {
      const int32_t f = Cl.chIdx[h];
      for (int32_t t = tPtr[f]; t < tPtr[f + 1]; t++) cSlot.push_back(rmark[tCol[t]] == I ? rpos[tCol[t]] : -1);
    }
    csPtr[I + 1] = (int32_t)cSlot.size();
    if (cSlot.size() > (size_t)INT32_MAX / 2) return;
  }
  C.tPtr.upload(tPtr, s);
  C.tsPtr.upload(tsPtr, s);
  C.tSlot.upload(tSlot, s);
  C.csPtr.upload(csPtr, s);
  C.cSlot.upload(cSlot, s);
  C.galT.alloc((size_t)36 * tPtr[nf]);
  C.galMaxT = maxT;
  C.gal2 = true;
}

void build_hierarchy(fcbga_model *M) {
  if (M->hierBuilt) return;
  be::stream_t s = M->stream;
  mgs::MeshView mv;
  mv.nx = M->g.nx; mv.ny = M->g.ny; mv.nz = M->g.nz;
  mv.xs = M->xs.data(); mv.ys = M->ys.data(); mv.zs = M->zs.data();
  mv.node_id = M->node_id.data();
  mv.cell_elems = M->cell_elems.data();
  mv.e_part = M->e_part.data();
  mv.feature_x = M->fx; mv.feature_y = M->fy;
  if (!M->hierBorn.empty()) {
    // stacked-module stage: the born elements and the nodes they touch
    mv.elem_on.assign(M->ne, 0);
    mv.node_on.assign(M->nn, 0);
    const fe::Grid &g = M->g;
    for (int ci = 0; ci < g.ncx; ci++)
      for (int cj = 0; cj < g.ncy; cj++)
        for (int ck = 0; ck < g.ncz; ck++) {
          const int64_t c = ((int64_t)ci * g.ncy + cj) * g.ncz + ck;
          for (int q = 0; q < 2; q++) {
            const int32_t e = M->cell_elems[2 * c + q];
            if (e < 0 || !M->hierBorn[M->e_group[e]]) continue;
            mv.elem_on[e] = 1;
            for (int a = 0; a < 8; a++)
              mv.node_on[M->node_id[((int64_t)(ci + fe::corner_li(a)) * g.ny + cj + fe::corner_lj(a)) * g.nz + ck +
                                    fe::corner_lk(a)]] = 1;
          }
        }
  }
  mgs::build_hierarchy(mv, M->H);
  M->nodeAgg.upload(M->H.nodeAgg, s);
  M->zp64.upload(M->H.zp, s);
  {
    std::vector<float> z32(M->H.zp.begin(), M->H.zp.end());
    M->zp32.upload(z32, s);
  }
  M->lv.clear();
  const int nl = (int)M->H.L.size();
  for (int l = 0; l < nl; l++) {
    const mgs::Level &L = M->H.L[l];
    auto D = std::make_unique<DevShell>();
    D->nx = L.nx; D->ny = L.ny; D->nagg = L.nagg; D->ncol = L.nx * L.ny;
    D->nb = L.rowPtr[L.nagg];
    D->coarsest = (l == nl - 1);
    D->colStart.upload(L.colStart, s);
    D->rowPtr.upload(L.rowPtr, s);
    D->col.upload(L.col, s);
    std::vector<int32_t> brow(D->nb);
    for (int32_t a = 0; a < L.nagg; a++) {
      for (int32_t k = L.rowPtr[a]; k < L.rowPtr[a + 1]; k++) brow[k] = a;
      D->maxrow = std::max(D->maxrow, (int)(L.rowPtr[a + 1] - L.rowPtr[a]));
    }
    D->blockRow.upload(brow, s);
    D->val64.alloc((size_t)D->nb * 36);
    D->val32.alloc((size_t)D->nb * 36);
    std::vector<int64_t> cfp(D->ncol + 1, 0);
    for (int32_t c = 0; c < D->ncol; c++) {
      int64_t m = 6 * (int64_t)(L.colStart[c + 1] - L.colStart[c]);
      cfp[c + 1] = cfp[c] + m * m;
    }
    D->cfTotal = cfp[D->ncol];
    for (int32_t c = 0; c < D->ncol; c++) {
      const int m = 6 * (L.colStart[c + 1] - L.colStart[c]);
      if (m > SHELL_MAXM) throw be::Error("too many shell aggregates in one column");
      D->maxm = m > D->maxm ? m : D->maxm;
    }
    D->cfPtr.upload(cfp, s);
    {
      std::vector<int32_t> desc((size_t)SHELL_DESC * D->ncol, 0);
      for (int32_t c = 0; c < D->ncol; c++) {
        int32_t *d = desc.data() + (size_t)SHELL_DESC * c;
        const int32_t n0 = L.colStart[c], na = L.colStart[c + 1] - n0;
        d[0] = n0; d[1] = na;
        std::memcpy(d + 2, &cfp[c], sizeof(int64_t));
        for (int r = 0; r <= std::min(na, SHELL_DESC_ROWS); r++) d[4 + r] = L.rowPtr[n0 + r];
      }
      D->colDesc.upload(desc, s);
    }
    D->cf64.alloc(D->cfTotal);
    D->cf32.alloc(D->cfTotal);
    std::vector<int32_t> cols[4];
    for (int i = 0; i < L.nx; i++)
      for (int j = 0; j < L.ny; j++) cols[(i & 1) * 2 + (j & 1)].push_back(i * L.ny + j);
    for (int c = 0; c < 4; c++) { D->colors[c].upload(cols[c], s); D->ncolor[c] = (int32_t)cols[c].size(); }
    // what the column-walking kernels (restriction, sweeps, Galerkin) assume:
    // aggregates of a column are contiguous and in order, and an aggregate of
    // level 1 spans planes k0..k1 of its column inside the grid
    for (int32_t c = 0; c < D->ncol; c++)
      if (L.colStart[c + 1] < L.colStart[c]) throw be::Error("shell level: column starts not monotone");
    if (L.colStart[0] != 0 || L.colStart[D->ncol] != L.nagg) throw be::Error("shell level: column starts do not cover the aggregates");
    if (l == 0) {
      if ((int32_t)L.k0.size() != L.nagg || (int32_t)L.k1.size() != L.nagg)
        throw be::Error("shell level 1: plane ranges missing");
      for (int32_t a = 0; a < L.nagg; a++)
        if (L.k0[a] < 0 || L.k0[a] > L.k1[a] || L.k1[a] >= M->g.nz)
          throw be::Error("shell level 1: aggregate plane range outside its column");
      std::vector<int32_t> aggCol(L.nagg);
      for (int32_t c = 0; c < D->ncol; c++)
        for (int32_t a = L.colStart[c]; a < L.colStart[c + 1]; a++) aggCol[a] = c;
      D->aggCol.upload(aggCol, s);
      D->k0.upload(L.k0, s);
      D->k1.upload(L.k1, s);
    }
    if (!D->coarsest) {
      if (l + 1 < nl) {
        D->pPtr.upload(L.pPtr, s);
        D->pIdx.upload(L.pIdx, s);
        D->pW64.upload(L.pW, s);
        std::vector<float> w32(L.pW.begin(), L.pW.end());
        D->pW32.upload(w32, s);
      }
    }
    if (l > 0) {
      D->chPtr.upload(L.chPtr, s);
      D->chIdx.upload(L.chIdx, s);
      D->chQ.upload(L.chQ, s);
    }
    size_t n = 6 * (size_t)L.nagg;
    D->b64.alloc(n); D->x64.alloc(n); D->r64.alloc(n);
    D->b32.alloc(n); D->x32.alloc(n); D->r32.alloc(n);
    M->lv.push_back(std::move(D));
  }
  for (int l = 0; l + 1 < nl; l++) build_gal2(M, l, s);
  be::sync(s);
  M->hierBuilt = true;
}

// Galerkin product of level 1 from the fine operator (galerkinProduct,
// src/core/10-solver.js:193-245, with the fine rows gathered from the
// elements): block (A, B) = sum over nodes i of A and j of B of W_i^T Ks_ij W_j.
// The term of one node pair (i at plane ka of column (iA, jA), j at plane kk
// of column (iB, jB)); t receives W_i^T Ks_ij W_j (6 x 6).
FE_HD void galerkin_pair_term(const fe::Grid &g, int iA, int jA, int ka, int iB, int jB, int kk, int64_t gi,
                              int64_t gj, const int32_t *__restrict__ ct, const double *__restrict__ cs,
                              const double *__restrict__ Kc, const uint8_t *__restrict__ mask,
                              const double *__restrict__ sv, const double *__restrict__ zp, double t[36]) {
  double Kb[9];
  fe::pair_block(g, iA, jA, ka, iB, jB, kk, ct, cs, Kc, Kb);
  double Ks[9];
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++) {
      uint8_t mp = mask[3 * gi + p], mq = mask[3 * gj + q];
      double v = (mp && mq) ? sv[3 * gi + p] * Kb[3 * p + q] * sv[3 * gj + q] : 0.0;
      if (gi == gj && p == q && !mp) v = 1.0;
      Ks[3 * p + q] = v;
    }
  // W (3x6) of a node: diag(1/s) [I | u = t + theta x z'] rows
  double Wi[18], Wj[18];
  for (int q = 0; q < 18; q++) { Wi[q] = 0; Wj[q] = 0; }
  double zi = zp[gi], zj = zp[gj];
  double si0 = 1 / sv[3 * gi], si1 = 1 / sv[3 * gi + 1], si2 = 1 / sv[3 * gi + 2];
  double sj0 = 1 / sv[3 * gj], sj1 = 1 / sv[3 * gj + 1], sj2 = 1 / sv[3 * gj + 2];
  Wi[0] = si0; Wi[4] = zi * si0; Wi[7] = si1; Wi[9] = -zi * si1; Wi[14] = si2; Wi[17] = zi * si2;
  Wj[0] = sj0; Wj[4] = zj * sj0; Wj[7] = sj1; Wj[9] = -zj * sj1; Wj[14] = sj2; Wj[17] = zj * sj2;
  double Tm[18];
  for (int p = 0; p < 3; p++)
    for (int c = 0; c < 6; c++) {
      double v = 0;
      for (int q = 0; q < 3; q++) v += Ks[3 * p + q] * Wj[6 * q + c];
      Tm[6 * p + c] = v;
    }
  for (int a = 0; a < 6; a++)
    for (int c = 0; c < 6; c++) {
      double v = 0;
      for (int p = 0; p < 3; p++) v += Wi[6 * p + a] * Tm[6 * p + c];
      t[6 * a + c] = v;
    }
}

#if FCBGA_GPU
// One wavefront per block (A, B).  The block's node pairs are taken in the
// host loop's order, pair l = (plane k0[A] + l / 3, that plane + l % 3 - 1):
// lane l of a 64-pair chunk forms its pair's term (zero where the pair is not
// in the block), and the chunk's terms pass through LDS in two halves of 18
// entries, lanes 0..17 adding each half entry by entry in pair order, so
// every entry is the host loop's sum.  (Half the terms at a time keeps the
// LDS per wavefront small enough for four wavefronts per SIMD.)
constexpr int GAL1_STRIDE = 19;  // LDS words per pair and half term (18, padded)

__global__ void __launch_bounds__(64) k_galerkin_level1(fe::Grid g, const int32_t *__restrict__ brow,
                                                        const int32_t *__restrict__ bcol,
                                                        const int32_t *__restrict__ aggCol,
                                                        const int32_t *__restrict__ k0,
                                                        const int32_t *__restrict__ k1,
                                                        const int32_t *__restrict__ nodeAgg,
                                                        const int32_t *__restrict__ ct, const double *__restrict__ cs,
                                                        const double *__restrict__ Kc,
                                                        const uint8_t *__restrict__ mask,
                                                        const double *__restrict__ sv, const double *__restrict__ zp,
                                                        double *__restrict__ val, float *__restrict__ val32) {
  __shared__ double term[64 * GAL1_STRIDE];
  const int64_t kb = blockIdx.x;
  const int lane = threadIdx.x;
  const int32_t A = brow[kb], B = bcol[kb];
  const int32_t cA = aggCol[A], cB = aggCol[B];
  const int iA = cA / g.ny, jA = cA % g.ny, iB = cB / g.ny, jB = cB % g.ny;
  const int a0 = k0[A], npair = 3 * (k1[A] - a0 + 1), b0 = k0[B], b1 = k1[B];
  double acc0 = 0, acc1 = 0;  // entries lane and lane + 18 (lanes 0..17)
  for (int base = 0; base < npair; base += 64) {
    const int l = base + lane;
    const int ka = a0 + l / 3, kk = ka + l % 3 - 1;
    const int64_t gi = (int64_t)cA * g.nz + ka, gj = (int64_t)cB * g.nz + kk;
    bool ok = l < npair && kk >= b0 && kk <= b1;
    ok = ok && nodeAgg[gi] == A && nodeAgg[gj] == B;
    double t[36];
    if (ok) galerkin_pair_term(g, iA, jA, ka, iB, jB, kk, gi, gj, ct, cs, Kc, mask, sv, zp, t);
    const int nq = npair - base < 64 ? npair - base : 64;
    for (int half = 0; half < 2; half++) {
      for (int e = 0; e < 18; e++) term[lane * GAL1_STRIDE + e] = ok ? t[18 * half + e] : 0.0;
      __syncthreads();
      if (lane < 18) {
        double a = half ? acc1 : acc0;
        for (int q = 0; q < nq; q++) a += term[q * GAL1_STRIDE + lane];
        if (half) acc1 = a; else acc0 = a;
      }
      __syncthreads();
    }
  }
  if (lane < 18) {
    val[36 * kb + lane] = acc0;
    val32[36 * kb + lane] = (float)acc0;
    val[36 * kb + 18 + lane] = acc1;
    val32[36 * kb + 18 + lane] = (float)acc1;
  }
}
#endif

void galerkin_level1(fcbga_model *M) {
  DevShell &D = *M->lv[0];
  const fe::Grid g = M->g;
  const int32_t *brow = D.blockRow.p, *bcol = D.col.p, *aggCol = D.aggCol.p, *k0 = D.k0.p, *k1 = D.k1.p;
  const int32_t *nodeAgg = M->nodeAgg.p, *ct = M->cellType.p;
  const double *cs = M->cellScale64.p, *Kc = M->Kc64.p, *sv = M->s64.p, *zp = M->zp64.p;
  const uint8_t *mask = M->dofmask.p;
  double *val = D.val64.p;
  float *val32 = D.val32.p;
#if FCBGA_GPU
  if (D.nb > 0) {
    hipLaunchKernelGGL(k_galerkin_level1, dim3((unsigned)D.nb), dim3(64), 0, M->stream, g, brow, bcol, aggCol, k0, k1,
                       nodeAgg, ct, cs, Kc, mask, sv, zp, val, val32);
    BE_CHECK(hipGetLastError());
  }
#else
  const int nz = g.nz, ny = g.ny;
  be::pfor(D.nb, M->stream, [=] __host__ __device__(int64_t kb) {
    int32_t A = brow[kb], B = bcol[kb];
    int32_t cA = aggCol[A], cB = aggCol[B];
    int iA = cA / ny, jA = cA % ny, iB = cB / ny, jB = cB % ny;
    double acc[36];
    for (int q = 0; q < 36; q++) acc[q] = 0;
    for (int ka = k0[A]; ka <= k1[A]; ka++) {
      int64_t gi = (int64_t)cA * nz + ka;
      if (nodeAgg[gi] != A) continue;
      int kb0 = k0[B] > ka - 1 ? k0[B] : ka - 1, kb1 = k1[B] < ka + 1 ? k1[B] : ka + 1;
      for (int kk = kb0; kk <= kb1; kk++) {
        int64_t gj = (int64_t)cB * nz + kk;
        if (nodeAgg[gj] != B) continue;
        double t[36];
        galerkin_pair_term(g, iA, jA, ka, iB, jB, kk, gi, gj, ct, cs, Kc, mask, sv, zp, t);
        for (int q = 0; q < 36; q++) acc[q] += t[q];
      }
    }
    for (int q = 0; q < 36; q++) { val[36 * kb + q] = acc[q]; val32[36 * kb + q] = (float)acc[q]; }
  });
#endif
}

// Galerkin product between shell levels l (fine) and l+1 (coarse)
#if FCBGA_GPU

// One 64-lane block per coarse row I: the row's blocks accumulate in LDS
// while the wave walks the (child f, fine block k, parent pp of col k)
// triples of the row once; lane (r, c) forms Tm = K_f,g W_g entry (r, c) and
// lane (a, c) adds (W_f^T Tm)(a, c) to the block of column J = parent(pp).
// Each block entry receives its terms in the same order as the per-block
// form below.
__global__ void __launch_bounds__(64) k_galerkin_shell(
    const int32_t *__restrict__ crp, const int32_t *__restrict__ ccol, const int32_t *__restrict__ chPtr,
    const int32_t *__restrict__ chIdx, const int32_t *__restrict__ chQ, const int32_t *__restrict__ frp,
    const int32_t *__restrict__ fcol, const int32_t *__restrict__ pPtr, const int32_t *__restrict__ pIdx,
    const double *__restrict__ fval, const double *__restrict__ W, double *__restrict__ val,
    float *__restrict__ val32) {
  extern __shared__ double acc[];  // maxrow x 36
  __shared__ int32_t hkey[256];
  __shared__ int32_t hslot[256];
  const int32_t I = blockIdx.x;
  const int lane = threadIdx.x;
  const int32_t r0 = crp[I], nblk = crp[I + 1] - r0;
  for (int q = lane; q < nblk * 36; q += 64) acc[q] = 0;
  for (int q = lane; q < 256; q += 64) hkey[q] = -1;
  __syncthreads();
  if (lane == 0)  // open-addressing map column J -> slot of the row
    for (int32_t sl = 0; sl < nblk; sl++) {
      const int32_t J = ccol[r0 + sl];
      int h = J & 255;
      while (hkey[h] >= 0) h = (h + 1) & 255;
      hkey[h] = J;
      hslot[h] = sl;
    }
  __syncthreads();
  const int rr = lane / 6, cc = lane % 6;  // (r, c) for Tm and (a, c) for the product, lanes < 36
  const bool act = lane < 36;
  for (int32_t h = chPtr[I]; h < chPtr[I + 1]; h++) {
    const int32_t f = chIdx[h];
    const double *Wf = W + 36 * (int64_t)chQ[h];
    double wfa[6];
    for (int r = 0; r < 6; r++) wfa[r] = act ? Wf[6 * r + rr] : 0.0;  // W_f(r, a), a = rr
    for (int32_t k = frp[f]; k < frp[f + 1]; k++) {
      const int32_t gg = fcol[k];
      const double *Kf = fval + 36 * (int64_t)k;
      double kr[6];
      for (int q = 0; q < 6; q++) kr[q] = act ? Kf[6 * rr + q] : 0.0;
      for (int32_t pp = pPtr[gg]; pp < pPtr[gg + 1]; pp++) {
        const int32_t J = pIdx[pp];
        int hh = J & 255, probes = 0;
        while (hkey[hh] != J && hkey[hh] >= 0 && probes < 256) { hh = (hh + 1) & 255; probes++; }
        if (hkey[hh] != J) continue;  // not in the coarse pattern (the per-block form drops it too)
        const int sl = hslot[hh];
        const double *Wg = W + 36 * (int64_t)pp;
        double tm = 0;
        if (act)
          for (int q = 0; q < 6; q++) tm += kr[q] * Wg[6 * q + cc];
        double t = 0;
        for (int r = 0; r < 6; r++) t += wfa[r] * __shfl(tm, 6 * r + cc, 64);
        if (act) acc[36 * sl + lane] += t;
      }
    }
  }
  __syncthreads();
  for (int q = lane; q < nblk * 36; q += 64) {
    val[36 * (int64_t)r0 + q] = acc[q];
    val32[36 * (int64_t)r0 + q] = (float)acc[q];
  }
}
#endif

// Two-phase Galerkin product (DevShell::gal2), one 64-lane block per row:
// lane (r, c), r, c < 6, owns entry (r, c) of every block of the row's
// accumulator in LDS, so the lanes never share an LDS word and need no barrier.
// Phase 1, fine row f: T(f, .) = sum over (block k, parent pp) of A_k W_pp.
// Phase 2, coarse row I: A(I, .) = sum over (child h, T block t) of W_h^T T_t.
#if FCBGA_GPU
__global__ void __launch_bounds__(64) k_gal_ap(const int32_t *__restrict__ frp, const int32_t *__restrict__ fcol,
                                               const int32_t *__restrict__ pPtr, const double *__restrict__ W,
                                               const double *__restrict__ fval, const int32_t *__restrict__ tPtr,
                                               const int32_t *__restrict__ tsPtr, const int32_t *__restrict__ tSlot,
                                               double *__restrict__ T) {
  extern __shared__ double acc[];  // galMaxT x 36
  const int32_t f = blockIdx.x;
  const int lane = threadIdx.x;
  if (lane >= 36) return;
  const int r = lane / 6, c = lane % 6;
  const int32_t t0 = tPtr[f], nt = tPtr[f + 1] - t0;
  for (int32_t q = 0; q < nt; q++) acc[36 * q + lane] = 0.0;
  int32_t idx = tsPtr[f];
  for (int32_t k = frp[f]; k < frp[f + 1]; k++) {
    const int32_t gg = fcol[k];
    const double *K = fval + 36 * (int64_t)k + 6 * r;
    double kr[6];
#pragma unroll
    for (int q = 0; q < 6; q++) kr[q] = K[q];
    for (int32_t pp = pPtr[gg]; pp < pPtr[gg + 1]; pp++) {
      const int32_t sl = tSlot[idx++];
      const double *Wg = W + 36 * (int64_t)pp + c;
      double tm = 0.0;
#pragma unroll
      for (int q = 0; q < 6; q++) tm += kr[q] * Wg[6 * q];
      acc[36 * sl + lane] += tm;
    }
  }
  for (int32_t q = 0; q < nt; q++) T[36 * ((int64_t)t0 + q) + lane] = acc[36 * q + lane];
}

__global__ void __launch_bounds__(64) k_gal_ptap(const int32_t *__restrict__ crp, const int32_t *__restrict__ chPtr,
                                                 const int32_t *__restrict__ chIdx, const int32_t *__restrict__ chQ,
                                                 const double *__restrict__ W, const int32_t *__restrict__ tPtr,
                                                 const double *__restrict__ T, const int32_t *__restrict__ csPtr,
                                                 const int32_t *__restrict__ cSlot, double *__restrict__ val,
                                                 float *__restrict__ val32) {
  extern __shared__ double acc[];  // maxrow x 36
  const int32_t I = blockIdx.x;
  const int lane = threadIdx.x;
  if (lane >= 36) return;
  const int a = lane / 6, c = lane % 6;
  const int32_t r0 = crp[I], nblk = crp[I + 1] - r0;
  for (int32_t q = 0; q < nblk; q++) acc[36 * q + lane] = 0.0;
  int32_t idx = csPtr[I];
  for (int32_t h = chPtr[I]; h < chPtr[I + 1]; h++) {
    const int32_t f = chIdx[h];
    const double *Wf = W + 36 * (int64_t)chQ[h] + a;
    double wfa[6];
#pragma unroll
    for (int q = 0; q < 6; q++) wfa[q] = Wf[6 * q];  // W_f(q, a)
    for (int32_t t = tPtr[f]; t < tPtr[f + 1]; t++) {
      const int32_t sl = cSlot[idx++];
      if (sl < 0) continue;
      const double *Tt = T + 36 * (int64_t)t + c;
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < 6; q++) v += wfa[q] * Tt[6 * q];
      acc[36 * sl + lane] += v;
    }
  }
  for (int32_t q = 0; q < nblk; q++) {
    const double v = acc[36 * q + lane];
    val[36 * ((int64_t)r0 + q) + lane] = v;
    val32[36 * ((int64_t)r0 + q) + lane] = (float)v;
  }
}
#endif

void galerkin_shell_2phase(fcbga_model *M, int l) {
  DevShell &F = *M->lv[l];
  DevShell &C = *M->lv[l + 1];
  const int32_t *frp = F.rowPtr.p, *fcol = F.col.p, *pPtr = F.pPtr.p;
  const int32_t *tPtr = C.tPtr.p, *tsPtr = C.tsPtr.p, *tSlot = C.tSlot.p, *csPtr = C.csPtr.p, *cSlot = C.cSlot.p;
  const int32_t *crp = C.rowPtr.p, *chPtr = C.chPtr.p, *chIdx = C.chIdx.p, *chQ = C.chQ.p;
  const double *W = F.pW64.p, *fval = F.val64.p;
  double *T = C.galT.p, *val = C.val64.p;
  float *val32 = C.val32.p;
#if FCBGA_GPU
  hipLaunchKernelGGL(k_gal_ap, dim3((unsigned)F.nagg), dim3(64), (size_t)C.galMaxT * 36 * sizeof(double), M->stream,
                     frp, fcol, pPtr, W, fval, tPtr, tsPtr, tSlot, T);
  BE_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_gal_ptap, dim3((unsigned)C.nagg), dim3(64),
                     (size_t)(C.maxrow > 0 ? C.maxrow : 1) * 36 * sizeof(double), M->stream, crp, chPtr, chIdx, chQ,
                     W, tPtr, T, csPtr, cSlot, val, val32);
  BE_CHECK(hipGetLastError());
#else
  // the same sums in the same order, one fine row / coarse row per task
  be::pfor(F.nagg, M->stream, [=] __host__ __device__(int64_t f) {
    const int32_t t0 = tPtr[f], nt = tPtr[f + 1] - t0;
    double *Tf = T + 36 * (int64_t)t0;
    for (int32_t q = 0; q < 36 * nt; q++) Tf[q] = 0.0;
    int32_t idx = tsPtr[f];
    for (int32_t k = frp[f]; k < frp[f + 1]; k++) {
      const int32_t gg = fcol[k];
      const double *K = fval + 36 * (int64_t)k;
      for (int32_t pp = pPtr[gg]; pp < pPtr[gg + 1]; pp++) {
        double *Ts = Tf + 36 * tSlot[idx++];
        const double *Wg = W + 36 * (int64_t)pp;
        for (int r = 0; r < 6; r++)
          for (int c = 0; c < 6; c++) {
            double tm = 0.0;
            for (int q = 0; q < 6; q++) tm += K[6 * r + q] * Wg[6 * q + c];
            Ts[6 * r + c] += tm;
          }
      }
    }
  });
  be::pfor(C.nagg, M->stream, [=] __host__ __device__(int64_t I) {
    const int32_t r0 = crp[I], nblk = crp[I + 1] - r0;
    double *A = val + 36 * (int64_t)r0;
    for (int32_t q = 0; q < 36 * nblk; q++) A[q] = 0.0;
    int32_t idx = csPtr[I];
    for (int32_t h = chPtr[I]; h < chPtr[I + 1]; h++) {
      const int32_t f = chIdx[h];
      const double *Wf = W + 36 * (int64_t)chQ[h];
      for (int32_t t = tPtr[f]; t < tPtr[f + 1]; t++) {
        const int32_t sl = cSlot[idx++];
        if (sl < 0) continue;
        const double *Tt = T + 36 * (int64_t)t;
        for (int a = 0; a < 6; a++)
          for (int c = 0; c < 6; c++) {
            double v = 0.0;
            for (int q = 0; q < 6; q++) v += Wf[6 * q + a] * Tt[6 * q + c];
            A[36 * sl + 6 * a + c] += v;
          }
      }
    }
    for (int32_t q = 0; q < 36 * nblk; q++) val32[36 * (int64_t)r0 + q] = (float)A[q];
  });
#endif
}

void galerkin_shell(fcbga_model *M, int l) {
  DevShell &F = *M->lv[l];
  DevShell &C = *M->lv[l + 1];
  if (C.gal2) {
    galerkin_shell_2phase(M, l);
    return;
  }
  const int32_t *brow = C.blockRow.p, *bcol = C.col.p, *chPtr = C.chPtr.p, *chIdx = C.chIdx.p, *chQ = C.chQ.p;
  const int32_t *frp = F.rowPtr.p, *fcol = F.col.p, *pPtr = F.pPtr.p, *pIdx = F.pIdx.p;
  const double *fval = F.val64.p, *W = F.pW64.p;
  double *val = C.val64.p;
  float *val32 = C.val32.p;
#if FCBGA_GPU
  if (C.maxrow <= GAL_MAXROW) {
    const size_t lds = (size_t)(C.maxrow > 0 ? C.maxrow : 1) * 36 * sizeof(double);
    hipLaunchKernelGGL(k_galerkin_shell, dim3(C.nagg), dim3(64), lds, M->stream, C.rowPtr.p, bcol, chPtr, chIdx,
                       chQ, frp, fcol, pPtr, pIdx, fval, W, val, val32);
    BE_CHECK(hipGetLastError());
    return;
  }
#endif
  be::pfor(C.nb, M->stream, [=] __host__ __device__(int64_t kb) {
    int32_t I = brow[kb], J = bcol[kb];
    double acc[36];
    for (int q = 0; q < 36; q++) acc[q] = 0;
    for (int32_t h = chPtr[I]; h < chPtr[I + 1]; h++) {
      int32_t f = chIdx[h];
      const double *Wf = W + 36 * (int64_t)chQ[h];
      for (int32_t k = frp[f]; k < frp[f + 1]; k++) {
        int32_t gg = fcol[k];
        for (int32_t pp = pPtr[gg]; pp < pPtr[gg + 1]; pp++) {
          if (pIdx[pp] != J) continue;
          const double *Wg = W + 36 * (int64_t)pp, *Kf = fval + 36 * (int64_t)k;
          double Tm[36];
          for (int r = 0; r < 6; r++)
            for (int c = 0; c < 6; c++) {
              double t = 0;
              for (int q = 0; q < 6; q++) t += Kf[6 * r + q] * Wg[6 * q + c];
              Tm[6 * r + c] = t;
            }
          for (int a = 0; a < 6; a++)
            for (int c = 0; c < 6; c++) {
              double t = 0;
              for (int r = 0; r < 6; r++) t += Wf[6 * r + a] * Tm[6 * r + c];
              acc[6 * a + c] += t;
            }
        }
      }
    }
    for (int q = 0; q < 36; q++) { val[36 * kb + q] = acc[q]; val32[36 * kb + q] = (float)acc[q]; }
  });
}

// Dense Cholesky of every column of a shell level (factorShellColumns,
// src/core/10-solver.js:474-494, jitter 1e-12) followed by the explicit
// inverse of the column block, so that a column solve in the cycle is one
// small dense product spread over a wavefront.
#if FCBGA_GPU
// One 64-lane block per column, m <= 48: the column block and the inverse are
// staged in LDS.  Every entry is formed with the same operations in the same
// order as chol_dense / chol_solve (jitter rule included); lanes only split
// independent rows (factor) and unit vectors (inverse).
__global__ void __launch_bounds__(64) k_shell_factor(const int32_t *__restrict__ cst, const int32_t *__restrict__ rp,
                                                     const int32_t *__restrict__ col, const double *__restrict__ val,
                                                     const int64_t *__restrict__ cfp, double *__restrict__ cf,
                                                     float *__restrict__ cf32, int mmax) {
  extern __shared__ double lds[];
  double *A = lds, *X = lds + mmax * mmax;  // A row-major m x m; X[i * mmax + q] = row i of inverse column q
  const int32_t c = blockIdx.x;
  const int32_t n0 = cst[c], n1 = cst[c + 1];
  const int m = 6 * (n1 - n0);
  if (!m) return;
  const int lane = threadIdx.x;
  for (int q = lane; q < m * m; q += 64) A[q] = 0;
  __syncthreads();
  for (int32_t n = n0; n < n1; n++)
    for (int32_t k = rp[n]; k < rp[n + 1]; k++) {
      const int32_t j = col[k];
      if (j < n0 || j >= n1) continue;
      if (lane < 36) {
        const int p = lane / 6, q = lane % 6;
        A[((n - n0) * 6 + p) * m + (j - n0) * 6 + q] = val[36 * (int64_t)k + lane];
      }
    }
  __syncthreads();
  double dmax = 0;
  for (int i = 0; i < m; i++) dmax = fmax(dmax, fabs(A[i * m + i]));
  const double jitter = 1e-12 * (dmax != 0 ? dmax : 1);
  for (int j = 0; j < m; j++) {
    const double *rj = A + j * m;
    double d = rj[j];
    for (int k = 0; k < j; k++) d -= rj[k] * rj[k];
    if (!(d > jitter)) d = fmax(fabs(d), jitter);
    const double ljj = sqrt(d);
    const double inv = 1 / ljj;
    const int i = j + 1 + lane;
    double li = 0;
    if (i < m) {
      const double *ri = A + i * m;
      double sacc = ri[j];
      for (int k = 0; k < j; k++) sacc -= ri[k] * rj[k];
      li = sacc * inv;
    }
    __syncthreads();
    if (lane == 0) A[j * m + j] = ljj;
    if (i < m) A[i * m + j] = li;
    __syncthreads();
  }
  // inverse: lane q solves L L^T x = e_q
  const int q = lane;
  if (q < m) {
    for (int i = 0; i < m; i++) {
      double sacc = (i == q) ? 1.0 : 0.0;
      for (int k = 0; k < i; k++) sacc -= A[i * m + k] * X[k * mmax + q];
      X[i * mmax + q] = sacc / A[i * m + i];
    }
    for (int i = m - 1; i >= 0; i--) {
      double sacc = X[i * mmax + q];
      for (int k = i + 1; k < m; k++) sacc -= A[k * m + i] * X[k * mmax + q];
      X[i * mmax + q] = sacc / A[i * m + i];
    }
  }
  __syncthreads();
  double *Ai = cf + cfp[c];
  float *Ai32 = cf32 + cfp[c];
  for (int e = lane; e < m * m; e += 64) {  // column-major: Ai[q * m + i]
    const double v = X[(e % m) * mmax + e / m];
    Ai[e] = v;
    Ai32[e] = (float)v;
  }
}
#endif

void factor_shell_columns(fcbga_model *M, int l) {
  DevShell &D = *M->lv[l];
  const int32_t *cst = D.colStart.p, *rp = D.rowPtr.p, *col = D.col.p;
  const int64_t *cfp = D.cfPtr.p;
  const double *val = D.val64.p;
  double *cf = D.cf64.p;
  float *cf32 = D.cf32.p;
#if FCBGA_GPU
  if (D.maxm <= 48) {
    const int mmax = D.maxm > 0 ? D.maxm : 1;
    const size_t lds = 2 * (size_t)mmax * mmax * sizeof(double);
    hipLaunchKernelGGL(k_shell_factor, dim3(D.ncol), dim3(64), lds, M->stream, cst, rp, col, val, cfp, cf, cf32, mmax);
    BE_CHECK(hipGetLastError());
    return;
  }
#endif
  D.cfTmp.alloc(D.cfTotal);
  double *L = D.cfTmp.p;
  be::pfor(D.ncol, M->stream, [=] __host__ __device__(int64_t c) {
    int32_t n0 = cst[c], n1 = cst[c + 1];
    int m = 6 * (n1 - n0);
    if (!m) return;
    double *A = L + cfp[c];
    for (int q = 0; q < m * m; q++) A[q] = 0;
    for (int32_t n = n0; n < n1; n++)
      for (int32_t k = rp[n]; k < rp[n + 1]; k++) {
        int32_t j = col[k];
        if (j < n0 || j >= n1) continue;
        for (int p = 0; p < 6; p++)
          for (int q = 0; q < 6; q++) A[((n - n0) * 6 + p) * m + (j - n0) * 6 + q] = val[36 * (int64_t)k + 6 * p + q];
      }
    fe::chol_dense(A, m, 1e-12);
    double *Ai = cf + cfp[c];
    for (int j = 0; j < m; j++) {
      double *x = Ai + (int64_t)j * m;  // column j
      for (int i = 0; i < m; i++) x[i] = (i == j) ? 1.0 : 0.0;
      fe::chol_solve<double, double>(A, m, x);
    }
    for (int q = 0; q < m * m; q++) cf32[cfp[c] + q] = (float)Ai[q];
  }, 64);
}

// acc -= sum over the entries k = k0, k0 + S, ... < k1 of row p of block k
// (val, 36 per block) times x at block column col[k], in that order.  The
// entries are taken U at a time: their column indices are loaded together,
// then their blocks and x values, so that one pass costs two memory
// latencies for U entries (an entry past k1 re-reads entry kb and adds zero).
template <typename T, int U = 4>
FE_HD void row_sub_blocks(T &acc, int32_t k0, int32_t k1, int32_t S, int p, const int32_t *__restrict__ col,
                          const T *__restrict__ val, const T *__restrict__ x) {
  for (int32_t kb = k0; kb < k1; kb += U * S) {
    int32_t kc[U], cj[U];
    for (int u = 0; u < U; u++) kc[u] = kb + u * S < k1 ? kb + u * S : kb;
    for (int u = 0; u < U; u++) cj[u] = col[kc[u]];
    for (int u = 0; u < U; u++) {
      const T *v = val + 36 * (int64_t)kc[u] + 6 * p, *xj = x + 6 * (int64_t)cj[u];
      const T d = v[0] * xj[0] + v[1] * xj[1] + v[2] * xj[2] + v[3] * xj[3] + v[4] * xj[4] + v[5] * xj[5];
      acc -= kb + u * S < k1 ? d : T(0);  // exact: a zero term changes nothing
    }
  }
}

#if FCBGA_GPU
// inverse of a banded Cholesky factor, one wavefront per unit vector e_c,
// the vector in LDS.  Both substitutions go by blocks of 64 rows, lane l on
// row l of the block: first the lane sums its row's terms over the rows
// already solved (its own sequential dot), then the block's triangle is
// solved row by row, the new entry taken from its lane (readlane; times the
// row's reciprocal pivot, so the step's critical path is one multiply and
// one multiply-add) and the triangle's coefficients read from LDS, where
// they were staged first.
constexpr int BAND_INV_B = 64;  // rows per block = lanes

__device__ __forceinline__ double readlane_f64(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

__global__ void __launch_bounds__(64) k_band_inverse(int32_t n, int bw, int w, const double *__restrict__ Ab,
                                                     double *__restrict__ inv) {
  extern __shared__ double band_lds[];
  double *tri = band_lds;                          // tri[j * 64 + l]: row l's coefficient of block row j
  double *xs = band_lds + BAND_INV_B * BAND_INV_B;  // the vector (n values)
  const int32_t c = blockIdx.x;
  const int lane = threadIdx.x;
  for (int32_t i = lane; i < n; i += 64) xs[i] = 0.0;
  __syncthreads();
  // forward, L y = e_c: rows c .. n - 1
  for (int32_t i0 = c; i0 < n; i0 += BAND_INV_B) {
    const int32_t r = i0 + lane;
    const bool live = r < n;
    const int jn = n - i0 < BAND_INV_B ? n - i0 : BAND_INV_B;
    // L(r, i0 + j) for i0 + j < r within the band; every load is issued (an
    // entry outside reads the first one) so that their latencies overlap
#pragma unroll 16
    for (int j = 0; j < jn; j++) {
      const int32_t k = i0 + j;
      const bool in = live && k < r && k >= r - bw;
      const double v = Ab[in ? (int64_t)r * w + (k - r + bw) : 0];
      tri[j * BAND_INV_B + lane] = in ? v : 0.0;
    }
    double rhs = r == c ? 1.0 : 0.0;
    const double rd = live ? 1.0 / Ab[(int64_t)r * w + bw] : 1.0;  // the row's reciprocal pivot
    if (live) {
      double part = 0;
      const double *row = Ab + (int64_t)r * w + bw - r;  // row[k] = L(r, k)
#pragma unroll 8
      for (int32_t k = (r - bw > c ? r - bw : c); k < i0; k++) part += row[k] * xs[k];
      rhs -= part;
    }
    __syncthreads();
    for (int j = 0; j < jn; j++) {
      const int32_t rj = i0 + j;
      const double xj = readlane_f64(rhs, j) * readlane_f64(rd, j);
      if (lane == j) xs[rj] = xj;
      rhs -= tri[j * BAND_INV_B + lane] * xj;
    }
    __syncthreads();
  }
  // backward, L^T x = y: rows n - 1 .. 0, lane l on row i1 - l
  for (int32_t i1 = n - 1; i1 >= 0; i1 -= BAND_INV_B) {
    const int32_t r = i1 - lane;
    const bool live = r >= 0;
    const int jn = i1 + 1 < BAND_INV_B ? i1 + 1 : BAND_INV_B;
    // L(i1 - j, r) for i1 - j > r within the band
#pragma unroll 16
    for (int j = 0; j < jn; j++) {
      const int32_t k = i1 - j;
      const bool in = live && k > r && k <= r + bw;
      const double v = Ab[in ? (int64_t)k * w + (r - k + bw) : 0];
      tri[j * BAND_INV_B + lane] = in ? v : 0.0;
    }
    double rhs = live ? xs[r] : 0.0;
    const double rd = live ? 1.0 / Ab[(int64_t)r * w + bw] : 1.0;
    if (live) {
      double part = 0;
      const int32_t kend = r + bw < n - 1 ? r + bw : n - 1;
#pragma unroll 8
      for (int32_t k = i1 + 1; k <= kend; k++) part += Ab[(int64_t)k * w + (r - k + bw)] * xs[k];
      rhs -= part;
    }
    __syncthreads();
    for (int j = 0; j < jn; j++) {
      const int32_t rj = i1 - j;
      const double xj = readlane_f64(rhs, j) * readlane_f64(rd, j);
      if (lane == j) xs[rj] = xj;
      rhs -= tri[j * BAND_INV_B + lane] * xj;
    }
    __syncthreads();
  }
  for (int32_t i = lane; i < n; i += 64) inv[(int64_t)c * n + i] = xs[i];
}

// x = inv b, one wavefront per row (coalesced row reads), FP64 accumulation
template <typename T>
__global__ void __launch_bounds__(256) k_dense_matvec(int32_t n, const double *__restrict__ A, const T *__restrict__ b,
                                                      T *__restrict__ x) {
  const int32_t row = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  double part = 0;
  const double *Ar = A + (int64_t)row * n;
  for (int32_t j = lane; j < n; j += 64) part += Ar[j] * (double)b[j];
  for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off, 64);
  if (lane == 0) x[row] = (T)part;
}

// One column of a shell-level sweep by one wavefront, m = 6 x aggregates <= 32
// (the usual case: 1..4 aggregates per column).  The lanes are split into
// S = 64/m groups: group g takes every S-th block of a row for the residual
// and every S-th column of the inverse for the correction, and the partial
// sums meet through lane shuffles (no LDS, no barrier).  Everything the lanes
// index by comes from the column's descriptor (dsc, the 12 ints of
// SHELL_DESC, read once by the whole wavefront), so the inverse, b, the block
// columns and values are all loaded after one round trip; the inverse is
// stored by columns (read coalesced).
template <typename T, int U>
__device__ __forceinline__ void shell_column_wave(const int32_t dsc[SHELL_DESC], int lane,
                                                  const int32_t *__restrict__ col, const T *__restrict__ val,
                                                  const T *__restrict__ inv, const T *__restrict__ b,
                                                  T *__restrict__ x) {
  const int32_t n0 = dsc[0];
  const int m = 6 * dsc[1];
  if (m == 0) return;
  int64_t cfo;
  __builtin_memcpy(&cfo, dsc + 2, sizeof(cfo));
  const T *A = inv + cfo;
  const int S = 64 / m;
  const int i = lane % m, grp = lane / m;
  const int r = i / 6;  // row of the column, r < na <= SHELL_DESC_ROWS
  int32_t k0 = dsc[4], k1 = dsc[5];
#pragma unroll
  for (int q = 1; q < SHELL_DESC_ROWS; q++)
    if (r == q) { k0 = dsc[4 + q]; k1 = dsc[5 + q]; }
  T ainv[16];
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const int j = grp + u * S;
    ainv[u] = (grp < S && j < m) ? A[(int64_t)j * m + i] : T(0);
  }
  const T bl = b[6 * (int64_t)n0 + i];
  T acc = 0;
  if (grp < S) row_sub_blocks<T, U>(acc, k0 + grp, k1, S, i % 6, col, val, x);
  T rr = bl;  // every lane ends with r_i of its row i
  for (int g = 0; g < S; g++) rr += __shfl(acc, g * m + i, 64);
  T dx = 0;
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const int j = grp + u * S;
    const T rj = __shfl(rr, j < m ? j : 0, 64);
    if (grp < S && j < m) dx += ainv[u] * rj;
  }
  T d = 0;
  for (int g = 0; g < S; g++) d += __shfl(dx, g * m + i, 64);
  if (lane < m) x[6 * (int64_t)n0 + lane] += d;
}

// one colour of a shell-level column sweep, one 64-lane block per column;
// colour (ci, cj) holds the columns (ci + 2u, cj + 2v), u-major
template <typename T, int U>
__global__ void __launch_bounds__(64) k_shell_sweep(int ci, int cj, int ny, const int32_t *__restrict__ desc,
                                                    const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                    const T *__restrict__ val, const T *__restrict__ inv,
                                                    const T *__restrict__ b, T *__restrict__ x) {
  __shared__ T r[SHELL_MAXM];
  const int nyc = (ny - cj + 1) >> 1;
  const int32_t c = (ci + 2 * (int)(blockIdx.x / nyc)) * ny + cj + 2 * (int)(blockIdx.x % nyc);
  int32_t dsc[SHELL_DESC];
  {
    const int4 *dp = reinterpret_cast<const int4 *>(desc + (int64_t)SHELL_DESC * c);
#pragma unroll
    for (int v = 0; v < SHELL_DESC / 4; v++) {
      const int4 q = dp[v];
      dsc[4 * v] = q.x; dsc[4 * v + 1] = q.y; dsc[4 * v + 2] = q.z; dsc[4 * v + 3] = q.w;
    }
  }
  const int32_t n0 = dsc[0];
  const int m = 6 * dsc[1];
  const int lane = threadIdx.x;
  if (m <= 32) {
    shell_column_wave<T, U>(dsc, lane, col, val, inv, b, x);
    return;
  }
  // larger columns: one lane per row
  int64_t cfo;
  __builtin_memcpy(&cfo, dsc + 2, sizeof(cfo));
  const T *A = inv + cfo;
  for (int i = lane; i < m; i += 64) {
    const int32_t n = n0 + i / 6;
    T acc = b[6 * (int64_t)n0 + i];
    row_sub_blocks<T>(acc, rp[n], rp[n + 1], 1, i % 6, col, val, x);
    r[i] = acc;
  }
  __syncthreads();
  for (int i = lane; i < m; i += 64) {
    T acc = 0;
    for (int j = 0; j < m; j++) acc += A[(int64_t)j * m + i] * r[j];
    x[6 * (int64_t)n0 + i] += acc;
  }
}

// One damped column-Jacobi sweep over every column of a shell level:
// xout = xin + omega * (column inverse) (b - A xin) on each column's rows, all
// columns at once from xin (one launch per sweep, no colours).  The column
// work is shell_column_wave's with the iterate split into input and output.
template <typename T, int U>
__global__ void __launch_bounds__(64) k_shell_jacobi(const int32_t *__restrict__ desc,
                                                     const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                     const T *__restrict__ val, const T *__restrict__ inv,
                                                     const T *__restrict__ b, const T *__restrict__ xin,
                                                     T *__restrict__ xout, T omega) {
  __shared__ T rs[SHELL_MAXM];
  const int32_t c = blockIdx.x;
  int32_t dsc[SHELL_DESC];
  {
    const int4 *dp = reinterpret_cast<const int4 *>(desc + (int64_t)SHELL_DESC * c);
#pragma unroll
    for (int v = 0; v < SHELL_DESC / 4; v++) {
      const int4 q = dp[v];
      dsc[4 * v] = q.x; dsc[4 * v + 1] = q.y; dsc[4 * v + 2] = q.z; dsc[4 * v + 3] = q.w;
    }
  }
  const int lane = threadIdx.x;
  const int32_t n0 = dsc[0];
  const int m = 6 * dsc[1];
  if (m == 0) return;
  int64_t cfo;
  __builtin_memcpy(&cfo, dsc + 2, sizeof(cfo));
  const T *A = inv + cfo;
  if (m <= 32) {
    const int S = 64 / m;
    const int i = lane % m, grp = lane / m;
    const int r = i / 6;
    int32_t k0 = dsc[4], k1 = dsc[5];
#pragma unroll
    for (int q = 1; q < SHELL_DESC_ROWS; q++)
      if (r == q) { k0 = dsc[4 + q]; k1 = dsc[5 + q]; }
    T ainv[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int j = grp + u * S;
      ainv[u] = (grp < S && j < m) ? A[(int64_t)j * m + i] : T(0);
    }
    const T bl = b[6 * (int64_t)n0 + i];
    T acc = 0;
    if (grp < S) row_sub_blocks<T, U>(acc, k0 + grp, k1, S, i % 6, col, val, xin);
    T rr = bl;
    for (int g = 0; g < S; g++) rr += __shfl(acc, g * m + i, 64);
    T dx = 0;
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int j = grp + u * S;
      const T rj = __shfl(rr, j < m ? j : 0, 64);
      if (grp < S && j < m) dx += ainv[u] * rj;
    }
    T d = 0;
    for (int g = 0; g < S; g++) d += __shfl(dx, g * m + i, 64);
    if (lane < m) xout[6 * (int64_t)n0 + lane] = xin[6 * (int64_t)n0 + lane] + omega * d;
    return;
  }
  // larger columns: one lane per row
  for (int i = lane; i < m; i += 64) {
    const int32_t n = n0 + i / 6;
    T acc = b[6 * (int64_t)n0 + i];
    row_sub_blocks<T>(acc, rp[n], rp[n + 1], 1, i % 6, col, val, xin);
    rs[i] = acc;
  }
  __syncthreads();
  for (int i = lane; i < m; i += 64) {
    T acc = 0;
    for (int j = 0; j < m; j++) acc += A[(int64_t)j * m + i] * rs[j];
    xout[6 * (int64_t)n0 + i] = xin[6 * (int64_t)n0 + i] + omega * acc;
  }
}

#endif

// host form: the same substitutions column by column
void band_inverse_host(int32_t n, int bw, int w, const double *Ab, double *inv) {
  std::vector<double> x(n);
  for (int32_t c = 0; c < n; c++) {
    std::fill(x.begin(), x.end(), 0.0);
    for (int32_t i = c; i < n; i++) {
      double part = 0;
      for (int32_t k = std::max(c, i - bw); k < i; k++) part += Ab[(int64_t)i * w + (k - i + bw)] * x[k];
      x[i] = ((i == c ? 1.0 : 0.0) - part) / Ab[(int64_t)i * w + bw];
    }
    for (int32_t i = n - 1; i >= 0; i--) {
      double part = 0;
      for (int32_t k = i + 1; k < std::min(n, i + bw + 1); k++) part += Ab[(int64_t)k * w + (i - k + bw)] * x[k];
      x[i] = (x[i] - part) / Ab[(int64_t)i * w + bw];
    }
    for (int32_t i = 0; i < n; i++) inv[(int64_t)c * n + i] = x[i];
  }
}

void band_inverse(int32_t n, int bw, int w, const double *Ab, double *inv, be::stream_t s) {
  if (n > 4096) throw be::Error("coarsest level larger than 4096 DOFs");
#if FCBGA_GPU
  const size_t lds = sizeof(double) * ((size_t)BAND_INV_B * BAND_INV_B + n);
  hipLaunchKernelGGL(k_band_inverse, dim3(n), dim3(64), lds, s, n, bw, w, Ab, inv);
  BE_CHECK(hipGetLastError());
#else
  (void)s;
  band_inverse_host(n, bw, w, Ab, inv);
#endif
}

template <typename T>
void dense_matvec(int32_t n, const double *A, const T *b, T *x, be::stream_t s) {
#if FCBGA_GPU
  hipLaunchKernelGGL(k_dense_matvec<T>, dim3((n + 3) / 4), dim3(256), 0, s, n, A, b, x);
  BE_CHECK(hipGetLastError());
#else
  (void)s;
  for (int32_t i = 0; i < n; i++) {
    double acc = 0;
    for (int32_t j = 0; j < n; j++) acc += A[(int64_t)i * n + j] * (double)b[j];
    x[i] = (T)acc;
  }
#endif
}

#if FCBGA_GPU
// Banded Cholesky on the device, right-looking by panels of BAND_P columns
// (band rows stored [i][k - i + bw]).  k_band_jitter: the jitter of
// factorCoarsest (src/core/10-solver.js:562-597), rel * max|diag|.
// k_band_panel (one workgroup): the panel's rows j0 .. j0+P+bw-1 staged in
// LDS, factored column by column (pivot with the jitter rule, scale, rank-1
// update of the panel's later columns).  k_band_trail: the rank-P update of
// the trailing window by the panel.
constexpr int BAND_P = 32;

__global__ void __launch_bounds__(64) k_band_jitter(int32_t n, int bw, int w, const double *__restrict__ A, double rel,
                                                    double *__restrict__ jit) {
  double dm = 0;
  for (int32_t i = threadIdx.x; i < n; i += 64) dm = fmax(dm, fabs(A[(int64_t)i * w + bw]));
  for (int off = 32; off > 0; off >>= 1) dm = fmax(dm, __shfl_xor(dm, off, 64));
  if (threadIdx.x == 0) *jit = rel * (dm != 0 ? dm : 1);
}

template <int P>
__global__ void __launch_bounds__(1024) k_band_panel(int32_t n, int bw, int w, double *__restrict__ A, int32_t j0,
                                                     const double *__restrict__ jit) {
  extern __shared__ double sm[];
  const int32_t r1 = j0 + P + bw < n ? j0 + P + bw : n;
  const int R = r1 - j0;
  constexpr int LD = P + 1;              // padded row stride (no bank conflicts on column reads)
  double *Lp = sm, *colS = sm + R * LD;  // panel R x P; scaled column, double-buffered
  const int tid = threadIdx.x;
  for (int e = tid; e < R * P; e += 1024) {
    const int r = e / P, c = e % P;
    const int32_t i = j0 + r, j = j0 + c;
    Lp[r * LD + c] = (j <= i && i - j <= bw) ? A[(int64_t)i * w + (j - i + bw)] : 0.0;
  }
  __syncthreads();
  const double jitter = *jit;
  // thread (row group rg, panel column c); one barrier per column: column jj
  // is scaled into colS while the later columns take the rank-1 update from
  // the unscaled entries (scaled on the fly), and column jj-1 is copied back
  const int c = tid % P, rg = tid / P;
  constexpr int NRG = 1024 / P;
  const int Pc = n - j0 < P ? n - j0 : P;  // columns of this panel (the last one may be short)
  for (int jj = 0; jj < Pc; jj++) {
    double d = Lp[jj * LD + jj];
    if (!(d > jitter)) d = fmax(fabs(d), jitter);
    const double ljj = sqrt(d), inv = 1 / ljj;
    double *cur = colS + (jj & 1) * R;
    const double *prev = colS + ((jj + 1) & 1) * R;
    const double lc = c < Pc ? Lp[c * LD + jj] * inv : 0.0;
    for (int r = rg; r < R; r += NRG) {
      if (c == jj - 1) {
        if (r >= c) Lp[r * LD + c] = prev[r];
      } else if (c == jj) {
        if (r >= jj) cur[r] = r == jj ? ljj : Lp[r * LD + jj] * inv;
      } else if (c > jj && c < Pc && r >= c) {
        Lp[r * LD + c] -= (Lp[r * LD + jj] * inv) * lc;
      }
    }
    __syncthreads();
  }
  if (c == Pc - 1)
    for (int r = rg; r < R; r += NRG)
      if (r >= c) Lp[r * LD + c] = colS[((Pc - 1) & 1) * R + r];
  __syncthreads();
  for (int e = tid; e < R * P; e += 1024) {
    const int r = e / P, cc = e % P;
    const int32_t i = j0 + r, j = j0 + cc;
    if (j <= i && i - j <= bw) A[(int64_t)i * w + (j - i + bw)] = Lp[r * LD + cc];
  }
}

__global__ void __launch_bounds__(256) k_band_trail(int32_t n, int bw, int w, double *__restrict__ A, int32_t j0,
                                                    int P) {
  const int32_t t0 = j0 + P, t1 = j0 + P + bw < n ? j0 + P + bw : n;
  const int T = t1 - t0;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (T <= 0 || e >= (int64_t)T * T) return;
  const int ri = (int)(e / T), rk = (int)(e % T);
  if (rk > ri) return;
  const int32_t i = t0 + ri, k = t0 + rk;
  if (i - k > bw) return;
  double s = 0;
  for (int jj = 0; jj < P; jj++) {
    const int32_t j = j0 + jj;
    if (i - j > bw) continue;
    s += A[(int64_t)i * w + (j - i + bw)] * A[(int64_t)k * w + (j - k + bw)];
  }
  A[(int64_t)i * w + (k - i + bw)] -= s;
}

// Banded Cholesky of the coarsest level in one launch of one workgroup
// (BCH_THREADS threads), right-looking by blocks of BCH columns:
//  1. wave 0 factors the block's diagonal triangle in registers (lane r holds
//     row j0 + r; each pivot takes the jitter rule of factorCoarsest, and the
//     pivot and the scaled column reach the other lanes through LDS),
//  2. every thread solves one row of the panel below the block (rows j1 ..
//     j1 + bw - 1) against the triangle, staged in LDS,
//  3. the threads subtract the panel's rank-BCH product from the trailing band
//     window, 4 x 4 entries per thread, the panel staged in LDS.
// The band stays in global memory (it is a few MB, L2-resident); one
// workgroup keeps it coherent with barriers alone.  The pivots and column
// entries are those of the left-looking factor up to the order in which the
// earlier blocks' terms are subtracted.
constexpr int BCH = 32;
constexpr int BCH_LD = BCH + 1;
constexpr int BCH_THREADS = 512;

__host__ __device__ inline size_t band_chol_lds(int bw) { return sizeof(double) * (size_t)BCH_LD * (BCH + bw); }

__global__ void __launch_bounds__(BCH_THREADS) k_band_chol(int32_t n, int bw, int w, double *__restrict__ A,
                                                           double rel) {
  extern __shared__ double sm[];
  double *Dl = sm;                 // BCH x BCH_LD: the block's factored triangle
  double *Pn = sm + BCH * BCH_LD;  // bw x BCH_LD: the panel rows below it
  constexpr int NW = BCH_THREADS / 64;
  __shared__ double red[NW];
  __shared__ double colb[64];  // column broadcast of the diagonal triangle's factor
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double dm = 0;
  for (int32_t i = tid; i < n; i += BCH_THREADS) dm = fmax(dm, fabs(A[(int64_t)i * w + bw]));
  for (int off = 32; off > 0; off >>= 1) dm = fmax(dm, __shfl_xor(dm, off, 64));
  if (lane == 0) red[wv] = dm;
  __syncthreads();
  double mx = 0;
  for (int q = 0; q < NW; q++) mx = fmax(mx, red[q]);
  const double jitter = rel * (mx != 0 ? mx : 1.0);
  for (int32_t j0 = 0; j0 < n; j0 += BCH) {
    const int nb = n - j0 < BCH ? (int)(n - j0) : BCH;
    const int32_t j1 = j0 + nb;
    const int R = (int)((j1 + bw < n ? j1 + bw : n) - j1);
    if (wv == 0) {
      double row[BCH];
#pragma unroll
      for (int c = 0; c < BCH; c++)
        row[c] = (lane < nb && c <= lane && lane - c <= bw) ? A[(int64_t)(j0 + lane) * w + (c - lane + bw)] : 0.0;
      // the pivot and the scaled column pass through LDS (one wave: its LDS
      // operations run in program order, which the may-alias accesses keep)
#pragma unroll
      for (int c = 0; c < BCH; c++) {
        if (c < nb) {
          colb[lane] = row[c];
          __builtin_amdgcn_wave_barrier();
          double d = colb[c];
          if (!(d > jitter)) d = fmax(fabs(d), jitter);
          const double ljj = sqrt(d), inv = 1.0 / ljj;
          const double l = lane == c ? ljj : (lane > c ? row[c] * inv : row[c]);
          row[c] = l;
          __builtin_amdgcn_wave_barrier();
          colb[lane] = l;
          __builtin_amdgcn_wave_barrier();
#pragma unroll
          for (int c2 = c + 1; c2 < BCH; c2++) {
            const double lc = colb[c2];
            if (lane >= c2) row[c2] -= l * lc;
          }
          __builtin_amdgcn_wave_barrier();
        }
      }
      if (lane < nb) {
#pragma unroll
        for (int c = 0; c < BCH; c++) {
          Dl[lane * BCH_LD + c] = c <= lane ? row[c] : 0.0;
          if (c <= lane && lane - c <= bw) A[(int64_t)(j0 + lane) * w + (c - lane + bw)] = row[c];
        }
      }
    }
    __syncthreads();
    for (int t = tid; t < R; t += BCH_THREADS) {
      const int32_t i = j1 + t;
      double *x = Pn + t * BCH_LD;  // the row's solution, built in place
      for (int c = 0; c < BCH; c++) {
        const bool in = c < nb && i - (j0 + c) <= bw;
        x[c] = in ? A[(int64_t)i * w + (j0 + c - i + bw)] : 0.0;
      }
#pragma unroll 1
      for (int c = 0; c < nb; c++) {
        double a = x[c];
#pragma unroll 4
        for (int k = 0; k < c; k++) a -= x[k] * Dl[c * BCH_LD + k];
        a /= Dl[c * BCH_LD + c];
        x[c] = a;
        if (i - (j0 + c) <= bw) A[(int64_t)i * w + (j0 + c - i + bw)] = a;
      }
    }
    __syncthreads();
    const int T = (R + 3) >> 2;
    for (int q = tid; q < T * (T + 1) / 2; q += BCH_THREADS) {
      // tile (ti, tk), tk <= ti, in row-major order of the lower triangle
      int ti = (int)((sqrt(8.0 * q + 1.0) - 1.0) * 0.5);
      while ((ti + 1) * (ti + 2) / 2 <= q) ti++;
      while (ti * (ti + 1) / 2 > q) ti--;
      const int tk = q - ti * (ti + 1) / 2;
      double acc[4][4];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = 0.0;
#pragma unroll 2
      for (int c = 0; c < nb; c++) {
        double pi[4], pk[4];
#pragma unroll
        for (int a = 0; a < 4; a++) {
          const int ri = 4 * ti + a, rk = 4 * tk + a;
          pi[a] = ri < R ? Pn[ri * BCH_LD + c] : 0.0;
          pk[a] = rk < R ? Pn[rk * BCH_LD + c] : 0.0;
        }
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
          for (int b = 0; b < 4; b++) acc[a][b] += pi[a] * pk[b];
      }
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
          const int ri = 4 * ti + a, rk = 4 * tk + b;
          if (ri < R && rk <= ri) A[(int64_t)(j1 + ri) * w + (rk - ri + bw)] -= acc[a][b];
        }
    }
    __syncthreads();
  }
}
#endif

// serial banded Cholesky (left-looking) with the jitter rule of factorCoarsest
void band_factor_host(int32_t n, int bw, int w, double *A) {
  double dmax = 0;
  for (int32_t i = 0; i < n; i++) dmax = std::max(dmax, std::fabs(A[(size_t)i * w + bw]));
  const double jitter = 1e-12 * (dmax != 0 ? dmax : 1);
  for (int32_t j = 0; j < n; j++) {
    double d = A[(size_t)j * w + bw];
    int32_t k0 = std::max(0, j - bw);
    for (int32_t k = k0; k < j; k++) { double l = A[(size_t)j * w + (k - j + bw)]; d -= l * l; }
    if (!(d > jitter)) d = std::max(std::fabs(d), jitter);
    double ljj = std::sqrt(d);
    A[(size_t)j * w + bw] = ljj;
    double inv = 1 / ljj;
    int32_t iend = std::min(n, j + bw + 1);
    for (int32_t i = j + 1; i < iend; i++) {
      double sum = A[(size_t)i * w + (j - i + bw)];
      int32_t kk0 = std::max(k0, i - bw);
      for (int32_t k = kk0; k < j; k++) sum -= A[(size_t)i * w + (k - i + bw)] * A[(size_t)j * w + (k - j + bw)];
      A[(size_t)i * w + (j - i + bw)] = sum * inv;
    }
  }
}

// Coarsest level: the level's matrix scattered into band storage, banded
// Cholesky with the jitter rule of factorCoarsest (src/core/10-solver.js:
// 562-597), then the explicit inverse (one banded solve per unit vector) so
// that each cycle applies it as a dense FP64 product.  All on the device.
void factor_coarsest(fcbga_model *M) {
  DevShell &D = *M->lv.back();
  const mgs::Level &L = M->H.L.back();
  const int32_t na = L.nagg, n = 6 * na;
  int bwb = 0;
  for (int32_t i = 0; i < na; i++)
    for (int32_t k = L.rowPtr[i]; k < L.rowPtr[i + 1]; k++) bwb = std::max(bwb, std::abs(L.col[k] - i));
  const int bw = (bwb + 1) * 6 - 1, w = bw + 1;
  be::stream_t s = M->stream;
  D.band.alloc((size_t)n * w);
  double *A = D.band.p;
  be::pfor((int64_t)n * w, s, [=] __host__ __device__(int64_t q) { A[q] = 0.0; });
  {
    const int32_t *brow = D.blockRow.p, *bcol = D.col.p;
    const double *val = D.val64.p;
    be::pfor((int64_t)D.nb * 36, s, [=] __host__ __device__(int64_t e) {
      const int64_t k = e / 36;
      const int pq = (int)(e % 36), p = pq / 6, q = pq % 6;
      const int32_t i = brow[k], j = bcol[k];
      if (j > i) return;
      const int32_t r = 6 * i + p, c = 6 * j + q;
      if (c <= r) A[(int64_t)r * w + (c - r + bw)] = val[e];
    });
  }
#if FCBGA_GPU
  // panel width: 32 columns when the panel (rows P + bw, plus two scaled
  // columns) fits 64 KB of LDS, else 16, else the host factor below
  const auto lds_of = [&](int P) { return (size_t)(P + bw) * (P + 3) * sizeof(double); };
  const int P = lds_of(32) <= 64 * 1024 ? 32 : lds_of(16) <= 64 * 1024 ? 16 : 0;
  if (band_chol_lds(bw) <= 144 * 1024) {
    // the whole factor in one workgroup (the block and its panel fit LDS)
    hipLaunchKernelGGL(k_band_chol, dim3(1), dim3(BCH_THREADS), band_chol_lds(bw), s, n, bw, w, A, 1e-12);
    BE_CHECK(hipGetLastError());
    D.inv64.alloc((size_t)n * n);
    D.ndense = n;
    band_inverse(n, bw, w, D.band.p, D.inv64.p, s);
    return;
  }
  if (P) {
    D.jit.alloc(1);
    hipLaunchKernelGGL(k_band_jitter, dim3(1), dim3(64), 0, s, n, bw, w, A, 1e-12, D.jit.p);
    BE_CHECK(hipGetLastError());
    for (int32_t j0 = 0; j0 < n; j0 += P) {
      const int R = (j0 + P + bw < n ? j0 + P + bw : n) - j0;
      const size_t lds = (size_t)R * (P + 3) * sizeof(double);
      if (P == 32)
        hipLaunchKernelGGL(k_band_panel<32>, dim3(1), dim3(1024), lds, s, n, bw, w, A, j0, D.jit.p);
      else
        hipLaunchKernelGGL(k_band_panel<16>, dim3(1), dim3(1024), lds, s, n, bw, w, A, j0, D.jit.p);
      BE_CHECK(hipGetLastError());
      const int32_t t0 = j0 + P, t1 = j0 + P + bw < n ? j0 + P + bw : n;
      if (t1 > t0) {
        const int64_t T = t1 - t0;
        hipLaunchKernelGGL(k_band_trail, dim3((unsigned)((T * T + 255) / 256)), dim3(256), 0, s, n, bw, w, A, j0, P);
        BE_CHECK(hipGetLastError());
      }
    }
    D.inv64.alloc((size_t)n * n);
    D.ndense = n;
    band_inverse(n, bw, w, D.band.p, D.inv64.p, s);
    return;
  }
  std::vector<double> Ah((size_t)n * w);
  be::d2h(Ah.data(), A, Ah.size() * sizeof(double), s);
  be::sync(s);
  band_factor_host(n, bw, w, Ah.data());
  D.band.upload(Ah, s);
#else
  band_factor_host(n, bw, w, A);
#endif
  D.inv64.alloc((size_t)n * n);
  D.ndense = n;
  band_inverse(n, bw, w, D.band.p, D.inv64.p, s);
}

double mg_setup(fcbga_model *M) {
  double t0 = now_ms();
  build_hierarchy(M);
  galerkin_level1(M);
  const int nl = (int)M->lv.size();
  for (int l = 0; l < nl; l++) {
    if (!M->lv[l]->coarsest) factor_shell_columns(M, l);
    if (l + 1 < nl) galerkin_shell(M, l);
  }
  factor_coarsest(M);
  be::sync(M->stream);
  return now_ms() - t0;
}

// ---------------------------------------------------------------------------
// V-cycle (MultigridPC._cycleOnce, src/core/10-solver.js:687-712): two
// forward colour sweeps, residual, restriction, coarse cycle, prolongation,
// two backward colour sweeps.  Colours replace the sequential column order.
// ---------------------------------------------------------------------------

template <typename T>
struct LevelVec;
template <>
struct LevelVec<double> {
  static double *b(DevShell &D) { return D.b64.p; }
  static double *x(DevShell &D) { return D.x64.p; }
  static double *r(DevShell &D) { return D.r64.p; }
  static const double *val(DevShell &D) { return D.val64.p; }
  static const double *cf(DevShell &D) { return D.cf64.p; }
  static const double *pW(DevShell &D) { return D.pW64.p; }
};
template <>
struct LevelVec<float> {
  static float *b(DevShell &D) { return D.b32.p; }
  static float *x(DevShell &D) { return D.x32.p; }
  static float *r(DevShell &D) { return D.r32.p; }
  static const float *val(DevShell &D) { return D.val32.p; }
  static const float *cf(DevShell &D) { return D.cf32.p; }
  static const float *pW(DevShell &D) { return D.pW32.p; }
};

template <typename T>
void shell_sweep(DevShell &D, int color, be::stream_t s) {
  const int32_t *cols = D.colors[color].p, *desc = D.colDesc.p, *rp = D.rowPtr.p, *col = D.col.p;
  const T *val = LevelVec<T>::val(D), *inv = LevelVec<T>::cf(D), *b = LevelVec<T>::b(D);
  T *x = LevelVec<T>::x(D);
  const int32_t ncols = D.ncolor[color];
  if (!ncols) return;
#if FCBGA_GPU
  (void)cols;
  // entries in flight per lane: two where the colour's columns fill the chip
  // (more registers would cut the waves resident per SIMD), four elsewhere
  if (ncols > 4096)
    hipLaunchKernelGGL((k_shell_sweep<T, 2>), dim3(ncols), dim3(64), 0, s, color >> 1, color & 1, D.ny, desc, rp, col,
                       val, inv, b, x);
  else
    hipLaunchKernelGGL((k_shell_sweep<T, 4>), dim3(ncols), dim3(64), 0, s, color >> 1, color & 1, D.ny, desc, rp, col,
                       val, inv, b, x);
  BE_CHECK(hipGetLastError());
#else
  (void)s;
  T r[SHELL_MAXM];
  for (int32_t t = 0; t < ncols; t++) {
    // the column's descriptor, as the device kernel reads it
    const int32_t c = cols[t], *dsc = desc + (int64_t)SHELL_DESC * c;
    const int32_t n0 = dsc[0], na = dsc[1];
    int64_t cfo;
    std::memcpy(&cfo, dsc + 2, sizeof(cfo));
    const int m = 6 * na;
    for (int i = 0; i < m; i++) {
      const int32_t n = n0 + i / 6;
      const int p = i % 6;
      const int32_t k0 = na <= SHELL_DESC_ROWS ? dsc[4 + i / 6] : rp[n];
      const int32_t k1 = na <= SHELL_DESC_ROWS ? dsc[5 + i / 6] : rp[n + 1];
      T acc = b[6 * (int64_t)n0 + i];
      for (int32_t k = k0; k < k1; k++) {
        const T *v = val + 36 * (int64_t)k + 6 * p, *xj = x + 6 * (int64_t)col[k];
        acc -= v[0] * xj[0] + v[1] * xj[1] + v[2] * xj[2] + v[3] * xj[3] + v[4] * xj[4] + v[5] * xj[5];
      }
      r[i] = acc;
    }
    const T *A = inv + cfo;
    for (int i = 0; i < m; i++) {
      T acc = 0;
      for (int j = 0; j < m; j++) acc += A[(int64_t)j * m + i] * r[j];
      x[6 * (int64_t)n0 + i] += acc;
    }
  }
#endif
}

// The coarse-level smoother of FCBGA_SMOOTH_COARSE_JACOBI: 2 x nsweeps damped
// column-Jacobi sweeps, the iterate alternating between x and the level's
// residual vector (free until the level's residual is formed), ending in x.
constexpr double JACOBI_OMEGA = 0.6;

template <typename T>
void jacobi_smooth(DevShell &D, int nsweeps, be::stream_t s) {
  T *x = LevelVec<T>::x(D), *t = LevelVec<T>::r(D);
  const int32_t *desc = D.colDesc.p, *rp = D.rowPtr.p, *col = D.col.p;
  const T *val = LevelVec<T>::val(D), *inv = LevelVec<T>::cf(D), *b = LevelVec<T>::b(D);
  for (int q = 0; q < 2 * nsweeps; q++) {
    const T *xi = (q & 1) ? t : x;
    T *xo = (q & 1) ? x : t;
#if FCBGA_GPU
    hipLaunchKernelGGL((k_shell_jacobi<T, 4>), dim3(D.ncol), dim3(64), 0, s, desc, rp, col, val, inv, b, xi, xo,
                       (T)JACOBI_OMEGA);
    BE_CHECK(hipGetLastError());
#else
    (void)s;
    T r[SHELL_MAXM];
    for (int32_t c = 0; c < D.ncol; c++) {
      const int32_t *dsc = desc + (int64_t)SHELL_DESC * c;
      const int32_t n0 = dsc[0], na = dsc[1];
      int64_t cfo;
      std::memcpy(&cfo, dsc + 2, sizeof(cfo));
      const int m = 6 * na;
      for (int i = 0; i < m; i++) {
        const int32_t n = n0 + i / 6;
        const int p = i % 6;
        T acc = b[6 * (int64_t)n0 + i];
        for (int32_t k = rp[n]; k < rp[n + 1]; k++) {
          const T *v = val + 36 * (int64_t)k + 6 * p, *xj = xi + 6 * (int64_t)col[k];
          acc -= v[0] * xj[0] + v[1] * xj[1] + v[2] * xj[2] + v[3] * xj[3] + v[4] * xj[4] + v[5] * xj[5];
        }
        r[i] = acc;
      }
      const T *A = inv + cfo;
      for (int i = 0; i < m; i++) {
        T acc = 0;
        for (int j = 0; j < m; j++) acc += A[(int64_t)j * m + i] * r[j];
        xo[6 * (int64_t)n0 + i] = xi[6 * (int64_t)n0 + i] + (T)JACOBI_OMEGA * acc;
      }
    }
#endif
  }
}

template <typename T>
void shell_residual(DevShell &D, be::stream_t s) {
  const int32_t *rp = D.rowPtr.p, *col = D.col.p;
  const T *val = LevelVec<T>::val(D), *b = LevelVec<T>::b(D), *x = LevelVec<T>::x(D);
  T *r = LevelVec<T>::r(D);
  // one thread per scalar row (aggregate n, DOF p)
  be::pfor(6 * (int64_t)D.nagg, s, [=] __host__ __device__(int64_t t) {
    int64_t n = t / 6;
    int p = (int)(t % 6);
    T acc = b[t];
    row_sub_blocks<T>(acc, rp[n], rp[n + 1], 1, p, col, val, x);
    r[t] = acc;
  });
}

template <typename T>
void zero_vec(T *v, int64_t n, be::stream_t s) {
  be::pfor(n, s, [=] __host__ __device__(int64_t i) { v[i] = T(0); });
}

// The fine x and the x of every shell level above the coarsest start a cycle
// at zero, and each level's x is first read by its own smoothing sweep: one
// launch clears them all when the cycle starts.
template <typename T>
struct ZeroSegs {
  static constexpr int MAXV = 8;
  T *p[MAXV];
  int64_t end[MAXV];  // running ends of the segments
  int n;
};

template <typename T>
void zero_cycle_vectors(fcbga_model *M, T *xf, int64_t n3, bool withCoarse, be::stream_t s) {
  ZeroSegs<T> sg{};
  int64_t tot = 0;
  const auto add = [&](T *p, int64_t n) {
    if (sg.n == ZeroSegs<T>::MAXV) {
      zero_vec<T>(p, n, s);
      return;
    }
    sg.p[sg.n] = p;
    tot += n;
    sg.end[sg.n++] = tot;
  };
  add(xf, n3);
  if (withCoarse)
    for (auto &D : M->lv)
      if (!D->coarsest) add(LevelVec<T>::x(*D), 6 * (int64_t)D->nagg);
  be::pfor(tot, s, [=] __host__ __device__(int64_t i) {
    int q = 0;
    while (i >= sg.end[q]) q++;
    sg.p[q][i - (q ? sg.end[q - 1] : 0)] = T(0);
  });
}

template <typename T>
void shell_cycle(fcbga_model *M, int l) {
  DevShell &D = *M->lv[l];
  be::stream_t s = M->stream;
  T *x = LevelVec<T>::x(D);
  const T *b = LevelVec<T>::b(D);
  if (D.coarsest) {
    KScope ks(M, FCBGA_KCLASS_COARSEST);
    dense_matvec<T>(D.ndense, D.inv64.p, b, x, s);
    return;
  }
  // x is zero here: fine_cycle clears it with the fine x when the cycle starts
  {
    KScope ks(M, l == 0 ? FCBGA_KCLASS_SHELL1_SWEEP : FCBGA_KCLASS_SHELL_SWEEP);
    if (l >= M->jacobiFrom)
      jacobi_smooth<T>(D, M->nsweeps, s);
    else
      for (int sw = 0; sw < M->nsweeps; sw++)
        for (int c = 0; c < 4; c++) shell_sweep<T>(D, c, s);
  }
  shell_residual<T>(D, s);
  DevShell &C = *M->lv[l + 1];
  {
    const int32_t *chPtr = C.chPtr.p, *chIdx = C.chIdx.p, *chQ = C.chQ.p;
    const T *W = LevelVec<T>::pW(D), *res = LevelVec<T>::r(D);
    T *bc = LevelVec<T>::b(C);
    be::pfor(6 * (int64_t)C.nagg, s, [=] __host__ __device__(int64_t t) {
      int64_t I = t / 6;
      int a = (int)(t % 6);
      T acc = 0;
      const int32_t h1 = chPtr[I + 1];
      for (int32_t hb = chPtr[I]; hb < h1; hb += 4) {  // four children per pass, loads batched
        int32_t hc[4], qv[4], fv[4];
        for (int u = 0; u < 4; u++) hc[u] = hb + u < h1 ? hb + u : hb;
        for (int u = 0; u < 4; u++) { qv[u] = chQ[hc[u]]; fv[u] = chIdx[hc[u]]; }
        for (int u = 0; u < 4; u++) {
          const T *w = W + 36 * (int64_t)qv[u] + a, *rf = res + 6 * (int64_t)fv[u];
          const T d = w[0] * rf[0] + w[6] * rf[1] + w[12] * rf[2] + w[18] * rf[3] + w[24] * rf[4] + w[30] * rf[5];
          acc += hb + u < h1 ? d : T(0);
        }
      }
      bc[t] = acc;
    });
  }
  shell_cycle<T>(M, l + 1);
  {
    const int32_t *pPtr = D.pPtr.p, *pIdx = D.pIdx.p;
    const T *W = LevelVec<T>::pW(D), *ec = LevelVec<T>::x(C);
    be::pfor(6 * (int64_t)D.nagg, s, [=] __host__ __device__(int64_t t) {
      int64_t f = t / 6;
      int p = (int)(t % 6);
      T acc = 0;
      const int32_t q1 = pPtr[f + 1];
      for (int32_t qb = pPtr[f]; qb < q1; qb += 4) {  // four coarse aggregates per pass, loads batched
        int32_t qc[4], ev[4];
        for (int u = 0; u < 4; u++) qc[u] = qb + u < q1 ? qb + u : qb;
        for (int u = 0; u < 4; u++) ev[u] = pIdx[qc[u]];
        for (int u = 0; u < 4; u++) {
          const T *w = W + 36 * (int64_t)qc[u] + 6 * p, *e = ec + 6 * (int64_t)ev[u];
          const T d = w[0] * e[0] + w[1] * e[1] + w[2] * e[2] + w[3] * e[3] + w[4] * e[4] + w[5] * e[5];
          acc += qb + u < q1 ? d : T(0);
        }
      }
      x[t] += acc;
    });
  }
  KScope ks(M, l == 0 ? FCBGA_KCLASS_SHELL1_SWEEP : FCBGA_KCLASS_SHELL_SWEEP);
  if (l >= M->jacobiFrom) {
    jacobi_smooth<T>(D, M->nsweeps, s);
    return;
  }
  for (int sw = 0; sw < M->nsweeps; sw++)
    for (int c = 3; c >= 0; c--) shell_sweep<T>(D, c, s);
}

// z = M^-1 r on the fine grid (T precision), r and z are T vectors of 3G
#if FCBGA_GPU
// Restriction of the fine residual to shell level 0, one 64-lane block per
// grid column: lane k loads plane k's scaled residual (coalesced along the
// column) into LDS, then one lane per aggregate of the column adds its nodes
// in node order, with the per-node terms of the one-thread-per-aggregate loop.
template <typename T>
__global__ void __launch_bounds__(64) k_restrict_fine(const int32_t *__restrict__ cst,
                                                      const int32_t *__restrict__ k0,
                                                      const int32_t *__restrict__ k1,
                                                      const int32_t *__restrict__ nodeAgg,
                                                      const T *__restrict__ res, const T *__restrict__ sinv,
                                                      const T *__restrict__ zp, int nz, T *__restrict__ bc) {
  __shared__ T s0[FINE_MAXNZ], s1[FINE_MAXNZ], s2[FINE_MAXNZ], sz[FINE_MAXNZ];
  __shared__ int32_t sa[FINE_MAXNZ];
  const int64_t c = blockIdx.x;
  const int k = threadIdx.x;
  if (k < nz) {
    const int64_t gi = c * nz + k;
    sa[k] = nodeAgg[gi];
    s0[k] = res[3 * gi] * sinv[3 * gi];
    s1[k] = res[3 * gi + 1] * sinv[3 * gi + 1];
    s2[k] = res[3 * gi + 2] * sinv[3 * gi + 2];
    sz[k] = zp[gi];
  }
  __syncthreads();
  const int32_t a0 = cst[c], na = cst[c + 1] - a0;
  for (int32_t q = k; q < na; q += 64) {
    const int32_t A = a0 + q;
    T acc[6] = {0, 0, 0, 0, 0, 0};
    const int ka = k0[A], kb = k1[A];
    for (int kk = ka; kk <= kb; kk++) {
      if (sa[kk] != A) continue;
      const T r0 = s0[kk], r1 = s1[kk], r2 = s2[kk], z = sz[kk];
      acc[0] += r0; acc[1] += r1; acc[2] += r2;
      acc[3] -= z * r1; acc[4] += z * r0; acc[5] += z * r2;
    }
    for (int a = 0; a < 6; a++) bc[6 * (int64_t)A + a] = acc[a];
  }
}
#endif

template <typename T>
void fine_cycle(fcbga_model *M, const T *b, T *x, T *r, T *dx, bool withCoarse) {
  be::stream_t s = M->stream;
  FineOps<T> o = fine_ops<T>(M);
  const int64_t n3 = 3 * M->g.G;
  zero_cycle_vectors<T>(M, x, n3, withCoarse, s);
  for (int sw = 0; sw < M->nsweeps; sw++)
    for (int c = 0; c < 4; c++) fine_color_sweep<T>(M, o, M->fineColors[c].p, M->nFineColor[c], b, x, r, dx, s);
  if (withCoarse) {
    {
      KScope ks(M, FCBGA_KCLASS_FINE_RESID);
      fine_residual<T>(o, b, x, r, s);
    }
    DevShell &C = *M->lv[0];
    const int32_t *aggCol = C.aggCol.p, *k0 = C.k0.p, *k1 = C.k1.p, *nodeAgg = o.nodeAgg;
    const T *sinv = o.sinv, *zp = o.zp;
    T *bc = LevelVec<T>::b(C);
    const int nz = M->g.nz;
    const T *res = r;
#if FCBGA_GPU
    if ((int64_t)C.ncol * nz == M->g.G && nz <= FINE_MAXNZ) {
      hipLaunchKernelGGL(k_restrict_fine<T>, dim3(C.ncol), dim3(64), 0, s, C.colStart.p, k0, k1, nodeAgg, res, sinv,
                         zp, nz, bc);
      BE_CHECK(hipGetLastError());
    } else
#endif
    be::pfor(C.nagg, s, [=] __host__ __device__(int64_t A) {
      T acc[6] = {0, 0, 0, 0, 0, 0};
      const int64_t c = aggCol[A];
      const int ka = k0[A], kb = k1[A];
      // the aggregate's nodes four at a time: their loads are issued together
      // (a node past kb re-reads node ka), then added in node order
      for (int k = ka; k <= kb; k += 4) {
        int32_t ag[4];
        T rv[4][3], sv[4][3], zv[4];
        for (int u = 0; u < 4; u++) {
          const int64_t gi = c * nz + (k + u <= kb ? k + u : ka);
          ag[u] = k + u <= kb ? nodeAgg[gi] : -1;
          for (int p = 0; p < 3; p++) { rv[u][p] = res[3 * gi + p]; sv[u][p] = sinv[3 * gi + p]; }
          zv[u] = zp[gi];
        }
        for (int u = 0; u < 4; u++) {
          if (ag[u] != (int32_t)A) continue;
          const T r0 = rv[u][0] * sv[u][0], r1 = rv[u][1] * sv[u][1], r2 = rv[u][2] * sv[u][2];
          const T z = zv[u];
          acc[0] += r0; acc[1] += r1; acc[2] += r2;
          acc[3] -= z * r1; acc[4] += z * r0; acc[5] += z * r2;
        }
      }
      for (int a = 0; a < 6; a++) bc[6 * A + a] = acc[a];
    });
    shell_cycle<T>(M, 0);
    const T *ec = LevelVec<T>::x(C);
    // prolongation into the iterate y = w x: the scaled correction
    // sinv (P e) times w is the unscaled (P e) on free DOFs
    const uint8_t *mask = M->dofmask.p;
    be::pfor(M->g.G, s, [=] __host__ __device__(int64_t gi) {
      int32_t A = nodeAgg[gi];
      if (A < 0) return;
      const T *e = ec + 6 * (int64_t)A;
      T z = zp[gi];
      if (mask[3 * gi]) x[3 * gi] += e[0] + z * e[4];
      if (mask[3 * gi + 1]) x[3 * gi + 1] += e[1] - z * e[3];
      if (mask[3 * gi + 2]) x[3 * gi + 2] += e[2] + z * e[5];
    });
  }
  for (int sw = 0; sw < M->nsweeps; sw++)
    for (int c = 3; c >= 0; c--) fine_color_sweep<T>(M, o, M->fineColors[c].p, M->nFineColor[c], b, x, r, dx, s);
}

// ---------------------------------------------------------------------------
// IC(0) with the Manteuffel shift (ICPC + scalarUpperCSR, src/core/10-solver.js:
// 718-800).  Node order is grid order (z fastest) and the reference's scalar
// row 3n+p, so the upper pattern of node n is itself and those of its 26 grid
// neighbours with a larger index that share an element with it: at most 14
// "slots" of 3x3, slot t at the grid offset ic_off(t), in increasing index
// order.  The factor U is kept in that layout (126 doubles per grid node; the
// self slot uses its upper triangle).  The factor is left-looking: entry (a, b)
// of row a receives -u_ia u_ib from every earlier row i holding both columns,
// in ascending i, then the pivot's square root and the division, which is the
// sequence of updates the reference's right-looking loop applies.  Nodes are
// processed level by level over the dependency graph (a node depends on its
// lower neighbours), one wavefront per node; both triangular solves run the
// same levels, forward ascending and backward descending.
// ---------------------------------------------------------------------------
constexpr int IC_SLOTS = 14, IC_STRIDE = 9 * IC_SLOTS, IC_WAVES = 4;

FE_HD void ic_off(int t, int &di, int &dj, int &dk) {
  if (t == 0) { di = 0; dj = 0; dk = 0; }
  else if (t == 1) { di = 0; dj = 0; dk = 1; }
  else if (t <= 4) { di = 0; dj = 1; dk = t - 3; }
  else { di = 1; dj = (t - 5) / 3 - 1; dk = (t - 5) % 3 - 1; }
}

// slot of an offset in [-1, 1]^3 that points to a larger grid index (or 0 for
// the node itself); -1 otherwise
FE_HD int ic_slot(int di, int dj, int dk) {
  if (di == 0 && dj == 0) return dk == 0 ? 0 : (dk == 1 ? 1 : -1);
  if (di == 0) return dj == 1 ? 3 + dk : -1;
  if (di == 1) return 5 + (dj + 1) * 3 + (dk + 1);
  return -1;
}

// Lower-neighbour updates of the entries (n p, nb q), p = 0..2, of node
// (i, j, k), nb in slot t: acc[p] -= u_{(m,r),(n,p)} u_{(m,r),(nb,q)} over the
// lower neighbours m in ascending order (descending slot of n in m) and their
// rows r = 0..2.
FE_HD void ic_lower_updates(const fe::Grid &g, int i, int j, int k, int t, int q, const uint16_t *mask,
                                   const double *U, double acc[3]) {
  int di, dj, dk;
  ic_off(t, di, dj, dk);
  for (int tl = IC_SLOTS - 1; tl >= 1; tl--) {
    int ei, ej, ek;
    ic_off(tl, ei, ej, ek);
    const int mi = i - ei, mj = j - ej, mk = k - ek;
    if (mi < 0 || mj < 0 || mj >= g.ny || mk < 0 || mk >= g.nz) continue;
    const int64_t gm = ((int64_t)mi * g.ny + mj) * g.nz + mk;
    const uint32_t mm = mask[gm];
    if (!((mm >> tl) & 1u)) continue;
    const int si = di + ei, sj = dj + ej, sk = dk + ek;  // offset of nb from m
    if (si > 1 || sj < -1 || sj > 1 || sk < -1 || sk > 1) continue;
    const int sb = ic_slot(si, sj, sk);
    if (sb <= 0 || !((mm >> sb) & 1u)) continue;
    const double *Um = U + gm * IC_STRIDE;
    for (int r = 0; r < 3; r++) {
      const double ub = Um[sb * 9 + 3 * r + q];
      for (int p = 0; p < 3; p++) acc[p] -= Um[tl * 9 + 3 * r + p] * ub;
    }
  }
}

// forward solve U^T y = r, lane (slot tl of node n in its lower neighbour m, row r of m):
// s[p] = u_{(m,r),(n,p)} y_{(m,r)}
FE_HD void ic_fwd_lane(const fe::Grid &g, int i, int j, int k, int tl, int r, const uint16_t *mask,
                              const double *U, const double *y, double s[3]) {
  int ei, ej, ek;
  ic_off(tl, ei, ej, ek);
  const int mi = i - ei, mj = j - ej, mk = k - ek;
  if (mi < 0 || mj < 0 || mj >= g.ny || mk < 0 || mk >= g.nz) return;
  const int64_t gm = ((int64_t)mi * g.ny + mj) * g.nz + mk;
  if (!((mask[gm] >> tl) & 1u)) return;
  const double ym = y[3 * gm + r];
  const double *u = U + gm * IC_STRIDE + tl * 9 + 3 * r;
  for (int p = 0; p < 3; p++) s[p] = u[p] * ym;
}

// the node's own rows: y_p = (r_p - s_p - sum_{r<p} u_rp y_r) / u_pp
FE_HD void ic_fwd_finish(const double *Un, const double *rn, const double s[3], double *yn) {
  const double y0 = (rn[0] - s[0]) / Un[0];
  const double y1 = (rn[1] - s[1] - Un[1] * y0) / Un[4];
  const double y2 = (rn[2] - s[2] - Un[2] * y0 - Un[5] * y1) / Un[8];
  yn[0] = y0; yn[1] = y1; yn[2] = y2;
}

// backward solve U z = y, lane (upper slot t >= 1, column q): s[p] = u_{(n,p),(nb,q)} z_{(nb,q)}
FE_HD void ic_bwd_lane(const fe::Grid &g, int64_t gi, int i, int j, int k, int t, int q, uint32_t mk,
                              const double *U, const double *z, double s[3]) {
  if (!((mk >> t) & 1u)) return;
  int di, dj, dk;
  ic_off(t, di, dj, dk);
  const int64_t gb = ((int64_t)(i + di) * g.ny + (j + dj)) * g.nz + (k + dk);
  const double zb = z[3 * gb + q];
  const double *u = U + gi * IC_STRIDE + t * 9 + q;
  for (int p = 0; p < 3; p++) s[p] = u[3 * p] * zb;
}

// zn holds y on entry
FE_HD void ic_bwd_finish(const double *Un, const double s[3], double *zn) {
  const double z2 = (zn[2] - s[2]) / Un[8];
  const double z1 = (zn[1] - s[1] - Un[5] * z2) / Un[4];
  const double z0 = (zn[0] - s[0] - Un[1] * z1 - Un[2] * z2) / Un[0];
  zn[0] = z0; zn[1] = z1; zn[2] = z2;
}

#if FCBGA_GPU
__device__ inline double wave_sum(double v) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// factor rows of node gi (one wavefront; lane 3t+q < 42 owns the column (slot
// t, q) of the node's three rows); a non-positive pivot sets fail[0]
__device__ inline void ic_factor_node(const fe::Grid &g, int64_t gi, const uint16_t *__restrict__ mask, double *U,
                                      int32_t *fail, int lane) {
  const int k = (int)(gi % g.nz);
  const int64_t cj = gi / g.nz;
  const int j = (int)(cj % g.ny), i = (int)(cj / g.ny);
  const uint32_t mk = mask[gi];
  const int t = lane / 3, q = lane - 3 * (lane / 3);
  const bool act = lane < 3 * IC_SLOTS && ((mk >> t) & 1u);
  double *Un = U + gi * IC_STRIDE;
  double acc[3] = {0.0, 0.0, 0.0};
  if (act) {
    for (int p = 0; p < 3; p++) acc[p] = Un[t * 9 + 3 * p + q];
    ic_lower_updates(g, i, j, k, t, q, mask, U, acc);
  }
  // the node's own rows, in order; lane p holds the diagonal of row p
  const double d0 = __shfl(acc[0], 0);
  const double u00 = sqrt(d0);
  const double u0 = lane == 0 ? u00 : acc[0] / u00;
  const double u01 = __shfl(u0, 1), u02 = __shfl(u0, 2);
  acc[1] -= u01 * u0;
  const double d1 = __shfl(acc[1], 1);
  const double u11 = sqrt(d1);
  const double u1 = lane == 1 ? u11 : acc[1] / u11;
  const double u12 = __shfl(u1, 2);
  acc[2] -= u02 * u0;
  acc[2] -= u12 * u1;
  const double d2 = __shfl(acc[2], 2);
  const double u22 = sqrt(d2);
  const double u2 = lane == 2 ? u22 : acc[2] / u22;
  if (act) {
    Un[t * 9 + q] = u0;
    if (t > 0 || q >= 1) Un[t * 9 + 3 + q] = u1;
    if (t > 0 || q >= 2) Un[t * 9 + 6 + q] = u2;
  }
  if (lane == 0 && !(d0 > 0 && d1 > 0 && d2 > 0)) fail[0] = 1;
}

// forward solve of node gi (one wavefront): z_n from r_n and the lower neighbours' z
__device__ inline void ic_forward_node(const fe::Grid &g, int64_t gi, const uint16_t *__restrict__ mask,
                                       const double *__restrict__ U, const double *__restrict__ r, double *z,
                                       int lane) {
  const int k = (int)(gi % g.nz);
  const int64_t cj = gi / g.nz;
  const int j = (int)(cj % g.ny), i = (int)(cj / g.ny);
  double s[3] = {0.0, 0.0, 0.0};
  if (lane >= 3 && lane < 3 * IC_SLOTS) ic_fwd_lane(g, i, j, k, lane / 3, lane % 3, mask, U, z, s);
  for (int p = 0; p < 3; p++) s[p] = wave_sum(s[p]);
  if (lane == 0) ic_fwd_finish(U + gi * IC_STRIDE, r + 3 * gi, s, z + 3 * gi);
}

// backward solve of node gi (one wavefront): z_n from y_n (in z) and the upper neighbours' z
__device__ inline void ic_backward_node(const fe::Grid &g, int64_t gi, const uint16_t *__restrict__ mask,
                                        const double *__restrict__ U, double *z, int lane) {
  const int k = (int)(gi % g.nz);
  const int64_t cj = gi / g.nz;
  const int j = (int)(cj % g.ny), i = (int)(cj / g.ny);
  double s[3] = {0.0, 0.0, 0.0};
  if (lane >= 3 && lane < 3 * IC_SLOTS) ic_bwd_lane(g, gi, i, j, k, lane / 3, lane % 3, mask[gi], U, z, s);
  for (int p = 0; p < 3; p++) s[p] = wave_sum(s[p]);
  if (lane == 0) ic_bwd_finish(U + gi * IC_STRIDE, s, z + 3 * gi);
}

// one dependency level per launch, one wavefront per node
__global__ void __launch_bounds__(64 * IC_WAVES) k_ic_factor(fe::Grid g, const int32_t *__restrict__ nodes,
                                                             int32_t cnt, const uint16_t *__restrict__ mask,
                                                             double *U, int32_t *fail) {
  const int64_t w = (int64_t)blockIdx.x * IC_WAVES + (threadIdx.x >> 6);
  if (w >= cnt) return;
  ic_factor_node(g, nodes[w], mask, U, fail, threadIdx.x & 63);
}

__global__ void __launch_bounds__(64 * IC_WAVES) k_ic_forward(fe::Grid g, const int32_t *__restrict__ nodes,
                                                              int32_t cnt, const uint16_t *__restrict__ mask,
                                                              const double *__restrict__ U,
                                                              const double *__restrict__ r, double *z) {
  const int64_t w = (int64_t)blockIdx.x * IC_WAVES + (threadIdx.x >> 6);
  if (w >= cnt) return;
  ic_forward_node(g, nodes[w], mask, U, r, z, threadIdx.x & 63);
}

__global__ void __launch_bounds__(64 * IC_WAVES) k_ic_backward(fe::Grid g, const int32_t *__restrict__ nodes,
                                                               int32_t cnt, const uint16_t *__restrict__ mask,
                                                               const double *__restrict__ U, double *z) {
  const int64_t w = (int64_t)blockIdx.x * IC_WAVES + (threadIdx.x >> 6);
  if (w >= cnt) return;
  ic_backward_node(g, nodes[w], mask, U, z, threadIdx.x & 63);
}
#endif

// dependency levels of the nodes present in mask (bit 0): a node's level is one
// more than the deepest of its lower neighbours; nodes grouped by level, in
// grid order within a level (ptr: level starts, nl + 1 entries)
void ic_levels(const fe::Grid &g, const std::vector<uint16_t> &mask, std::vector<int32_t> &nodes,
               std::vector<int64_t> &ptr) {
  std::vector<int32_t> level(g.G, -1);
  int32_t nl = 0;
  for (int i = 0; i < g.nx; i++)
    for (int j = 0; j < g.ny; j++)
      for (int k = 0; k < g.nz; k++) {
        const int64_t gi = ((int64_t)i * g.ny + j) * g.nz + k;
        if (!(mask[gi] & 1)) continue;
        int32_t L = 0;
        for (int tl = 1; tl < IC_SLOTS; tl++) {
          int ei, ej, ek;
          ic_off(tl, ei, ej, ek);
          const int mi = i - ei, mj = j - ej, mk = k - ek;
          if (mi < 0 || mj < 0 || mj >= g.ny || mk < 0 || mk >= g.nz) continue;
          const int64_t gm = ((int64_t)mi * g.ny + mj) * g.nz + mk;
          if ((mask[gm] >> tl) & 1u) L = std::max(L, level[gm] + 1);
        }
        level[gi] = L;
        nl = std::max(nl, L + 1);
      }
  ptr.assign(nl + 1, 0);
  for (int64_t q = 0; q < g.G; q++)
    if (level[q] >= 0) ptr[level[q] + 1]++;
  for (int32_t l = 0; l < nl; l++) ptr[l + 1] += ptr[l];
  nodes.assign(ptr[nl], 0);
  std::vector<int64_t> fillp(ptr.begin(), ptr.end() - 1);
  for (int64_t q = 0; q < g.G; q++)
    if (level[q] >= 0) nodes[fillp[level[q]]++] = (int32_t)q;
}

// pattern slots and dependency levels (once per model)
void ic_pattern(fcbga_model *M) {
  if (M->icPattern) return;
  const fe::Grid &g = M->g;
  std::vector<uint16_t> mask(g.G, 0);
  for (int64_t q = 0; q < g.G; q++)
    if (M->node_id[q] >= 0) mask[q] = 1;
  for (int ci = 0; ci < g.ncx; ci++)
    for (int cj = 0; cj < g.ncy; cj++)
      for (int ck = 0; ck < g.ncz; ck++) {
        const int64_t c = ((int64_t)ci * g.ncy + cj) * g.ncz + ck;
        if (M->cell_elems[2 * c] < 0 && M->cell_elems[2 * c + 1] < 0) continue;
        for (int a = 0; a < 8; a++) {
          const int ai = ci + fe::corner_li(a), aj = cj + fe::corner_lj(a), ak = ck + fe::corner_lk(a);
          const int64_t ga = ((int64_t)ai * g.ny + aj) * g.nz + ak;
          for (int b = 0; b < 8; b++) {
            const int bi = ci + fe::corner_li(b), bj = cj + fe::corner_lj(b), bk = ck + fe::corner_lk(b);
            const int t = ic_slot(bi - ai, bj - aj, bk - ak);
            if (t > 0) mask[ga] |= (uint16_t)(1u << t);
          }
        }
      }
  std::vector<int32_t> nodes;
  std::vector<int64_t> ptr;
  ic_levels(g, mask, nodes, ptr);
  M->icMask.upload(mask, M->stream);
  M->icNodes.upload(nodes, M->stream);
  M->icLvPtr = ptr;
  M->icFail.alloc(1);
  M->icPattern = true;
}

#if !FCBGA_GPU
// host builds: the factor node by node in level order (the device's
// per-node computation with the lanes as a loop); false on a non-positive pivot
bool ic_factor_host(const fe::Grid &g, const int32_t *nodes, int64_t nn, const uint16_t *mask, double *U) {
  bool ok = true;
  for (int64_t w = 0; w < nn; w++) {
    const int64_t gi = nodes[w];
    const int k = (int)(gi % g.nz);
    const int64_t cj = gi / g.nz;
    const int j = (int)(cj % g.ny), i = (int)(cj / g.ny);
    const uint32_t mk = mask[gi];
    double *Un = U + gi * IC_STRIDE;
    double acc[3][3 * IC_SLOTS];
    for (int l = 0; l < 3 * IC_SLOTS; l++) {
      const int t = l / 3, q = l % 3;
      double a3[3] = {0.0, 0.0, 0.0};
      if ((mk >> t) & 1u) {
        for (int p = 0; p < 3; p++) a3[p] = Un[t * 9 + 3 * p + q];
        ic_lower_updates(g, i, j, k, t, q, mask, U, a3);
      }
      for (int p = 0; p < 3; p++) acc[p][l] = a3[p];
    }
    for (int p = 0; p < 3; p++) {
      for (int r = 0; r < p; r++)
        for (int l = p; l < 3 * IC_SLOTS; l++) acc[p][l] -= acc[r][p] * acc[r][l];
      const double d = acc[p][p];
      ok = ok && d > 0;
      const double u = std::sqrt(d);
      acc[p][p] = u;
      for (int l = p + 1; l < 3 * IC_SLOTS; l++) acc[p][l] /= u;
    }
    for (int l = 0; l < 3 * IC_SLOTS; l++) {
      const int t = l / 3, q = l % 3;
      if (!((mk >> t) & 1u)) continue;
      for (int p = 0; p < 3; p++)
        if (t > 0 || q >= p) Un[t * 9 + 3 * p + q] = acc[p][l];
    }
  }
  return ok;
}

void ic_apply_host(const fe::Grid &g, const int32_t *nodes, int64_t nn, const uint16_t *mask, const double *U,
                   const double *r, double *z) {
  for (int64_t w = 0; w < nn; w++) {
    const int64_t gi = nodes[w];
    const int k = (int)(gi % g.nz);
    const int64_t cj = gi / g.nz;
    const int j = (int)(cj % g.ny), i = (int)(cj / g.ny);
    double s[3] = {0.0, 0.0, 0.0};
    for (int l = 3; l < 3 * IC_SLOTS; l++) {
      double t3[3] = {0.0, 0.0, 0.0};
      ic_fwd_lane(g, i, j, k, l / 3, l % 3, mask, U, z, t3);
      for (int p = 0; p < 3; p++) s[p] += t3[p];
    }
    ic_fwd_finish(U + gi * IC_STRIDE, r + 3 * gi, s, z + 3 * gi);
  }
  for (int64_t w = nn - 1; w >= 0; w--) {
    const int64_t gi = nodes[w];
    const int k = (int)(gi % g.nz);
    const int64_t cj = gi / g.nz;
    const int j = (int)(cj % g.ny), i = (int)(cj / g.ny);
    double s[3] = {0.0, 0.0, 0.0};
    for (int l = 3; l < 3 * IC_SLOTS; l++) {
      double t3[3] = {0.0, 0.0, 0.0};
      ic_bwd_lane(g, gi, i, j, k, l / 3, l % 3, mask[gi], U, z, t3);
      for (int p = 0; p < 3; p++) s[p] += t3[p];
    }
    ic_bwd_finish(U + gi * IC_STRIDE, s, z + 3 * gi);
  }
}
#endif

// values of the scaled, constraint-reduced matrix in the upper pattern with
// the diagonal times (1 + shift), then the factor level by level; false on a
// non-positive pivot
bool ic_try(fcbga_model *M, double shift) {
  const fe::Grid g = M->g;
  be::stream_t s = M->stream;
  M->icU.alloc((size_t)IC_STRIDE * g.G);
  const uint16_t *mask = M->icMask.p;
  const uint8_t *dm = M->dofmask.p;
  const int32_t *ct = M->cellType.p;
  const double *cs = M->cellScale64.p, *Kc = M->Kc64.p, *sv = M->s64.p;
  double *U = M->icU.p;
  int32_t *fail = M->icFail.p;
  be::pfor(1, s, [=] __host__ __device__(int64_t) { fail[0] = 0; });
  be::pfor(g.G, s, [=] __host__ __device__(int64_t gi) {
    const uint32_t mk = mask[gi];
    if (!(mk & 1u)) return;
    const int k = (int)(gi % g.nz);
    const int64_t cj = gi / g.nz;
    const int j = (int)(cj % g.ny), i = (int)(cj / g.ny);
    double *Un = U + gi * IC_STRIDE;
    for (int t = 0; t < IC_SLOTS; t++) {
      if (!((mk >> t) & 1u)) continue;
      int di, dj, dk;
      ic_off(t, di, dj, dk);
      const int64_t gj = ((int64_t)(i + di) * g.ny + (j + dj)) * g.nz + (k + dk);
      double Kb[9];
      fe::pair_block(g, i, j, k, i + di, j + dj, k + dk, ct, cs, Kc, Kb);
      for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) {
          double v = 0.0;
          if (t == 0 && q < p) { Un[3 * p + q] = 0.0; continue; }
          // applyConstraints keeps a constrained DOF's diagonal (1 when it is 0), then _scale
          if (dm[3 * gi + p] && dm[3 * gj + q]) v = Kb[3 * p + q] * (sv[3 * gi + p] * sv[3 * gj + q]);
          else if (t == 0 && p == q) v = Kb[3 * p + p] != 0.0 ? Kb[3 * p + p] * (sv[3 * gi + p] * sv[3 * gi + p]) : 1.0;
          if (t == 0 && p == q) v *= 1.0 + shift;
          Un[t * 9 + 3 * p + q] = v;
        }
    }
  });
  const int64_t nl = (int64_t)M->icLvPtr.size() - 1;
  const int32_t *nodes = M->icNodes.p;
#if FCBGA_GPU
  {
    KScope ks(M, FCBGA_KCLASS_IC_FACTOR);
    for (int64_t l = 0; l < nl; l++) {
      const int64_t a = M->icLvPtr[l], cnt = M->icLvPtr[l + 1] - a;
      if (cnt <= 0) continue;
      hipLaunchKernelGGL(k_ic_factor, dim3((unsigned)((cnt + IC_WAVES - 1) / IC_WAVES)), dim3(64 * IC_WAVES), 0, s,
                         g, nodes + a, (int32_t)cnt, mask, U, fail);
      BE_CHECK(hipGetLastError());
    }
  }
  int32_t h = 0;
  be::d2h(&h, fail, sizeof(int32_t), s);
  return h == 0;
#else
  return ic_factor_host(g, nodes, M->icLvPtr[nl], mask, U);
#endif
}

// ICPC.build: shifts 0, 1e-3, 1e-2, ... capped at 1, eight attempts
// (IC_SHIFT0, IC_SHIFT_MAX: src/core/00-constants.js:83-84); returns ms
double ic_setup(fcbga_model *M) {
  const double t0 = now_ms();
  ic_pattern(M);
  double shift = 0;
  for (int attempt = 0; attempt < 8; attempt++) {
    if (ic_try(M, shift)) {
      M->icShift = shift;
      be::sync(M->stream);
      return now_ms() - t0;
    }
    shift = shift == 0 ? 1e-3 : shift * 10;
    if (shift > 1) shift = 1;
  }
  throw be::Error("IC(0) failed even with the maximum diagonal shift");
}

// z = (U^T U)^-1 r (ICPC.apply); z at absent nodes is left as it was
void ic_apply(fcbga_model *M, const double *r, double *z) {
  const fe::Grid g = M->g;
  const int64_t nl = (int64_t)M->icLvPtr.size() - 1;
  const int32_t *nodes = M->icNodes.p;
  const uint16_t *mask = M->icMask.p;
  const double *U = M->icU.p;
#if FCBGA_GPU
  be::stream_t s = M->stream;
  for (int64_t l = 0; l < nl; l++) {
    const int64_t a = M->icLvPtr[l], cnt = M->icLvPtr[l + 1] - a;
    if (cnt <= 0) continue;
    hipLaunchKernelGGL(k_ic_forward, dim3((unsigned)((cnt + IC_WAVES - 1) / IC_WAVES)), dim3(64 * IC_WAVES), 0, s,
                       g, nodes + a, (int32_t)cnt, mask, U, r, z);
    BE_CHECK(hipGetLastError());
  }
  for (int64_t l = nl - 1; l >= 0; l--) {
    const int64_t a = M->icLvPtr[l], cnt = M->icLvPtr[l + 1] - a;
    if (cnt <= 0) continue;
    hipLaunchKernelGGL(k_ic_backward, dim3((unsigned)((cnt + IC_WAVES - 1) / IC_WAVES)), dim3(64 * IC_WAVES), 0, s,
                       g, nodes + a, (int32_t)cnt, mask, U, z);
    BE_CHECK(hipGetLastError());
  }
#else
  ic_apply_host(g, nodes, M->icLvPtr[nl], mask, U, r, z);
#endif
}

// ---------------------------------------------------------------------------
// PCG (LinearSolver._pcg, src/core/10-solver.js:1001-1042), cold start
// ---------------------------------------------------------------------------

double dotd(fcbga_model *M, const double *a, const double *b, int slot) {
  const int64_t n3 = 3 * M->g.G;
  be::dot_async(n3, a, b, M->partials.p, M->scal.p + slot, M->stream);
  double v;
  be::d2h(&v, M->scal.p + slot, sizeof(double), M->stream);
  return v;
}

// r32_ready: the mixed-precision cycle's FP32 copy of r (M->b32v) is already
// written (by the PCG update that formed r)
void precondition(fcbga_model *M, int precision, int precond, bool r32_ready = false) {
  // z = M^-1 r, then z = 0 on constrained / absent DOFs
  const int64_t n3 = 3 * M->g.G;
  be::stream_t s = M->stream;
  const uint8_t *mask = M->dofmask.p;
  const double *r = M->r.p;
  double *z = M->z.p;
  if (precond == FCBGA_PC_JACOBI) {
    be::pfor(n3, s, [=] __host__ __device__(int64_t i) { z[i] = mask[i] ? r[i] : 0.0; });
    return;
  }
  if (precond == FCBGA_PC_IC) {
    {
      KScope ks(M, FCBGA_KCLASS_IC_APPLY);
      ic_apply(M, r, z);
    }
    be::pfor(n3, s, [=] __host__ __device__(int64_t i) { if (!mask[i]) z[i] = 0.0; });
    return;
  }
  const bool coarse = precond == FCBGA_PC_MG;
  if (precision == FCBGA_PREC_FP64) {
    fine_cycle<double>(M, r, z, M->tmp64.p, M->dx64.p, coarse);
  } else {
    float *b32 = M->b32v.p, *x32 = M->x32v.p;
    if (!r32_ready) be::pfor(n3, s, [=] __host__ __device__(int64_t i) { b32[i] = (float)r[i]; });
    fine_cycle<float>(M, b32, x32, M->r32v.p, M->dx32.p, coarse);
    // the cycle's iterate is y = w z: z = y / s on free DOFs
    const double *sinv = M->sinv64.p;
    be::pfor(n3, s, [=] __host__ __device__(int64_t i) { z[i] = mask[i] ? (double)x32[i] * sinv[i] : 0.0; });
    return;
  }
  const double *sinv = M->sinv64.p;
  be::pfor(n3, s, [=] __host__ __device__(int64_t i) { z[i] = mask[i] ? z[i] * sinv[i] : 0.0; });
}

// device scalar slots of the PCG
enum { SC_RR = 0, SC_RZ = 1, SC_PQ = 2, SC_RZO = 3, SC_RZN = 4, SC_ALPHA = 5, SC_BETA = 6, SC_FLAG = 7, SC_COUNT = 8 };

#if FCBGA_GPU
// x += alpha p, r -= alpha q, and per block the partial sums of r.r and
// r.z over the updated r (the grid-stride order and block reduction of
// be::dot_async, so the sums are the same as two separate dots)
__global__ void __launch_bounds__(be::RED_BLOCK) k_cg_update(int64_t n, const double *__restrict__ sc,
                                                            double *__restrict__ x, double *__restrict__ r,
                                                            const double *__restrict__ p,
                                                            const double *__restrict__ q,
                                                            const double *__restrict__ z, double *__restrict__ part,
                                                            float *__restrict__ r32) {
  __shared__ double s1[be::RED_BLOCK / 64], s2[be::RED_BLOCK / 64];
  const double a = sc[SC_ALPHA];
  double rr = 0, rz = 0;
  for (int64_t i = (int64_t)blockIdx.x * be::RED_BLOCK + threadIdx.x; i < n; i += (int64_t)be::RED_GRID * be::RED_BLOCK) {
    x[i] += a * p[i];
    r[i] -= a * q[i];
    const double ri = r[i];
    rr += ri * ri;
    if (z) rz += ri * z[i];
    if (r32) r32[i] = (float)ri;
  }
  rr = be::block_sum(rr, s1);
  rz = be::block_sum(rz, s2);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rr;
    part[be::RED_GRID + blockIdx.x] = rz;
  }
}
#endif

// r32 (optional): the updated r rounded to FP32, the mixed-precision cycle's input
void cg_update(int64_t n3, double *sc, double *x, double *r, const double *p, const double *q, const double *z,
               double *part, float *r32, be::stream_t s) {
#if FCBGA_GPU
  hipLaunchKernelGGL(k_cg_update, dim3(be::RED_GRID), dim3(be::RED_BLOCK), 0, s, n3, sc, x, r, p, q, z, part, r32);
  BE_CHECK(hipGetLastError());
  // r.r into SC_RR and, with z, r.z into SC_RZO: one launch, a block per sum
  hipLaunchKernelGGL(be::k_sum_partials, dim3(z ? 2 : 1), dim3(be::RED_BLOCK), 0, s, be::RED_GRID, part, sc + SC_RR,
                     SC_RZO - SC_RR);
  BE_CHECK(hipGetLastError());
#else
  (void)part; (void)s;
  const double a = sc[SC_ALPHA];
  double rr = 0, rz = 0;
  for (int64_t i = 0; i < n3; i++) {
    x[i] += a * p[i];
    r[i] -= a * q[i];
    rr += r[i] * r[i];
    if (z) rz += r[i] * z[i];
    if (r32) r32[i] = (float)r[i];
  }
  sc[SC_RR] = rr;
  if (z) sc[SC_RZO] = rz;
#endif
}

// One PCG iteration (LinearSolver._pcg, src/core/10-solver.js:1021-1040) with
// every scalar kept on the device: q = A p, alpha = rz / pq, x += alpha p,
// r -= alpha q, rr = r.r, z = M^-1 r, beta (Fletcher-Reeves, or Polak-Ribiere
// for the flexible mixed-precision PCG), p = z + beta p.  A non-positive p.Ap
// sets the flag and leaves x and r unchanged.
void iteration_body(fcbga_model *M, const fcbga_solve_opts &o, bool flex) {
  be::stream_t s = M->stream;
  const int64_t n3 = 3 * M->g.G;
  FineOps<double> f64 = fine_ops<double>(M);
  double *x = M->x.p, *r = M->r.p, *z = M->z.p, *p = M->p.p, *q = M->q.p, *sc = M->scal.p;
  double *part = M->partials.p;
  {
    KScope ks(M, FCBGA_KCLASS_APPLY);
    fine_apply<double>(f64, p, q, s);
  }
  be::dot_async(n3, p, q, part, sc + SC_PQ, s);
  be::pfor(1, s, [=] __host__ __device__(int64_t) {
    double pq = sc[SC_PQ];
    bool ok = pq > 0;
    sc[SC_FLAG] = ok ? 0.0 : 1.0;
    sc[SC_ALPHA] = ok ? sc[SC_RZ] / pq : 0.0;
  });
  // x += alpha p, r -= alpha q with r.r and (flexible) r.z_old of the new r
  // in the same pass, before the preconditioner overwrites z
  // the mixed multigrid cycle's FP32 input is written by the same pass
  const bool r32 = o.precision == FCBGA_PREC_MIXED && (o.precond == FCBGA_PC_MG || o.precond == FCBGA_PC_COLUMN);
  cg_update(n3, sc, x, r, p, q, flex ? z : nullptr, part, r32 ? M->b32v.p : nullptr, s);
  precondition(M, o.precision, o.precond, r32);
  be::dot_async(n3, r, z, part, sc + SC_RZN, s);
  be::pfor(1, s, [=] __host__ __device__(int64_t) {
    double rz = sc[SC_RZ], rzn = sc[SC_RZN];
    double beta = flex ? (rzn - sc[SC_RZO]) / rz : rzn / rz;
    sc[SC_BETA] = sc[SC_FLAG] != 0 ? 0.0 : beta;
    if (sc[SC_FLAG] == 0) sc[SC_RZ] = rzn;
  });
  be::pfor(n3, s, [=] __host__ __device__(int64_t i) {
    if (sc[SC_FLAG] == 0) p[i] = z[i] + sc[SC_BETA] * p[i];
  });
}

// One PCG iteration, replayed from a captured HIP graph per key.  IC(0)
// iterations run eagerly: their level-scheduled triangular solves are one
// launch per dependency level (about 2,000 on the lidded mesh), and a graph
// of that size gains nothing over eager launches of a latency-bound apply
// while it was the one graph form a profiled run crashed in.
void run_iteration(fcbga_model *M, const fcbga_solve_opts &o, bool flex, int gkey) {
#if FCBGA_GPU
  if (M->useGraphs && !M->timing && o.precond != FCBGA_PC_IC) {
    auto it = M->graphs.find(gkey);
    if (it == M->graphs.end()) {
      hipGraph_t graph;
      BE_CHECK(hipStreamBeginCapture(M->stream, hipStreamCaptureModeThreadLocal));
      iteration_body(M, o, flex);
      BE_CHECK(hipStreamEndCapture(M->stream, &graph));
      hipGraphExec_t exec;
      BE_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      (void)hipGraphDestroy(graph);
      it = M->graphs.emplace(gkey, exec).first;
    }
    BE_CHECK(hipGraphLaunch(it->second, M->stream));
    return;
  }
#endif
  (void)gkey;
  iteration_body(M, o, flex);
}

void read_scalars(fcbga_model *M, double h[2]) {
#if FCBGA_GPU
  if (!M->hostScal) BE_CHECK(hipHostMalloc((void **)&M->hostScal, sizeof(double) * SC_COUNT, hipHostMallocDefault));
  BE_CHECK(hipMemcpyAsync(M->hostScal, M->scal.p, sizeof(double) * SC_COUNT, hipMemcpyDeviceToHost, M->stream));
  BE_CHECK(hipStreamSynchronize(M->stream));
  h[0] = M->hostScal[SC_RR];
  h[1] = M->hostScal[SC_FLAG];
#else
  h[0] = M->scal.p[SC_RR];
  h[1] = M->scal.p[SC_FLAG];
#endif
}

void alloc_vectors(fcbga_model *M, int precision) {
  const size_t n3 = 3 * (size_t)M->g.G;
  const void *before[6] = {M->x.p, M->z.p, M->b32v.p, M->partials.p, M->scal.p, M->dx64.p};
  M->x.alloc(n3); M->r.alloc(n3); M->z.alloc(n3); M->p.alloc(n3); M->q.alloc(n3);
  M->tmp64.alloc(n3); M->dx64.alloc(n3);
  if (precision == FCBGA_PREC_MIXED) {
    M->b32v.alloc(n3); M->x32v.alloc(n3); M->r32v.alloc(n3); M->dx32.alloc(n3);
  }
  M->partials.alloc(2 * be::RED_GRID);  // two partial sums per pass (cg_update)
  const bool fresh = M->scal.n != SC_COUNT;
  M->scal.alloc(SC_COUNT);
  if (fresh) M->scal.zero(M->stream);
  const void *after[6] = {M->x.p, M->z.p, M->b32v.p, M->partials.p, M->scal.p, M->dx64.p};
#if FCBGA_GPU
  // captured iteration graphs hold these addresses: drop them when a buffer moved
  for (int q = 0; q < 6; q++)
    if (before[q] && before[q] != after[q]) {
      drop_graphs(M);
      break;
    }
#else
  (void)before; (void)after;
#endif
}

// Gauss-Jordan inverse with partial pivoting, then x = A^-1 b (solveDense /
// invDense, src/core/01-util.js:111-142); false when singular
bool solve_dense_gj(std::vector<double> A, int n, const std::vector<double> &b, std::vector<double> &x) {
  const int w = 2 * n;
  std::vector<double> Mx((size_t)n * w, 0.0);
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) Mx[(size_t)i * w + j] = A[(size_t)i * n + j];
    Mx[(size_t)i * w + n + i] = 1;
  }
  for (int c = 0; c < n; c++) {
    int p = c;
    double best = std::fabs(Mx[(size_t)c * w + c]);
    for (int r = c + 1; r < n; r++) {
      const double v = std::fabs(Mx[(size_t)r * w + c]);
      if (v > best) { best = v; p = r; }
    }
    if (best == 0) return false;
    if (p != c)
      for (int j = 0; j < w; j++) std::swap(Mx[(size_t)c * w + j], Mx[(size_t)p * w + j]);
    const double inv = 1 / Mx[(size_t)c * w + c];
    for (int j = 0; j < w; j++) Mx[(size_t)c * w + j] *= inv;
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      const double f = Mx[(size_t)r * w + c];
      if (f == 0) continue;
      for (int j = 0; j < w; j++) Mx[(size_t)r * w + j] -= f * Mx[(size_t)c * w + j];
    }
  }
  x.assign(n, 0.0);
  for (int i = 0; i < n; i++) {
    double acc = 0;
    for (int j = 0; j < n; j++) acc += Mx[(size_t)i * w + n + j] * b[j];
    x[i] = acc;
  }
  return true;
}

// Galerkin projection of the load onto the stored solutions (_warmStart,
// src/core/10-solver.js:885-915), in the unscaled space: A_ij = v_i.K v_j
// (symmetrised), b_i = v_i.f, x0 = sum c_i v_i; M->x receives x0 / s (the
// PCG's scaled space).  f is the unscaled load (grid layout).  False (x left
// 0) without a basis or when the projection is singular or not finite.
bool warm_start(fcbga_model *M, const double *f) {
  const int m = (int)M->basis.size();
  if (!m) return false;
  be::stream_t s = M->stream;
  const int64_t n3 = 3 * M->g.G;
  const double *si = M->sinv64.p;
  M->warmTmp.alloc(n3);
  M->warmKv.alloc((size_t)m * n3);
  double *tmp = M->warmTmp.p;
  const FineOps<double> ops = fine_ops<double>(M);
  std::vector<double> A((size_t)m * m), b(m);
  for (int i = 0; i < m; i++) {
    const double *v = M->basis[i]->p;
    double *kv = M->warmKv.p + (size_t)i * n3;
    // K v = S^-1 Ks S^-1 v
    be::pfor(n3, s, [=] __host__ __device__(int64_t q) { tmp[q] = v[q] * si[q]; });
    fine_apply<double>(ops, tmp, kv, s);
    be::pfor(n3, s, [=] __host__ __device__(int64_t q) { kv[q] *= si[q]; });
  }
  for (int i = 0; i < m; i++) {
    b[i] = dotd(M, M->basis[i]->p, f, SC_RR);
    for (int j = 0; j < m; j++) A[(size_t)i * m + j] = dotd(M, M->basis[i]->p, M->warmKv.p + (size_t)j * n3, SC_RR);
  }
  for (int i = 0; i < m; i++)
    for (int j = i + 1; j < m; j++) {
      const double v = 0.5 * (A[(size_t)i * m + j] + A[(size_t)j * m + i]);
      A[(size_t)i * m + j] = A[(size_t)j * m + i] = v;
    }
  std::vector<double> c;
  if (!solve_dense_gj(A, m, b, c)) return false;
  for (int i = 0; i < m; i++)
    if (!std::isfinite(c[i])) return false;
  double *x = M->x.p;
  be::pfor(n3, s, [=] __host__ __device__(int64_t q) { x[q] = 0.0; });
  for (int i = 0; i < m; i++) {
    const double *v = M->basis[i]->p;
    const double ci = c[i];
    be::pfor(n3, s, [=] __host__ __device__(int64_t q) { x[q] += ci * v[q]; });
  }
  be::pfor(n3, s, [=] __host__ __device__(int64_t q) { x[q] *= si[q]; });
  return true;
}

// a converged solution joins the basis (_storeSolution, src/core/10-solver.js:917-926):
// two passes of modified Gram-Schmidt, normalised, the oldest dropped beyond nmax
void store_solution(fcbga_model *M, int nmax) {
  be::stream_t s = M->stream;
  const int64_t n3 = 3 * M->g.G;
  const double *x = M->x.p, *sv = M->s64.p;
  auto v = std::make_unique<DBuf<double>>();
  v->alloc(n3);
  double *vp = v->p;
  be::pfor(n3, s, [=] __host__ __device__(int64_t q) { vp[q] = x[q] * sv[q]; });
  const double nu = std::sqrt(dotd(M, vp, vp, SC_RR));
  for (int pass = 0; pass < 2; pass++)
    for (auto &bv : M->basis) {
      const double d = dotd(M, bv->p, vp, SC_RR);
      const double *bp = bv->p;
      be::pfor(n3, s, [=] __host__ __device__(int64_t q) { vp[q] -= d * bp[q]; });
    }
  const double nv = std::sqrt(dotd(M, vp, vp, SC_RR));
  if (!(nv > 1e-14 * (nu + 1e-300))) return;
  const double inv = 1 / nv;
  be::pfor(n3, s, [=] __host__ __device__(int64_t q) { vp[q] *= inv; });
  M->basis.push_back(std::move(v));
  while ((int)M->basis.size() > nmax) M->basis.erase(M->basis.begin());
}

fcbga_solve_opts check_opts(const fcbga_solve_opts *op) {
  fcbga_solve_opts o{1e-8, 2000, FCBGA_PREC_FP64, FCBGA_PC_MG, 0, 0, FCBGA_SMOOTH_GS};
  if (op) o = *op;
  if (!(o.tol > 0) || o.maxit < 1) throw be::Error("tol must be > 0 and maxit >= 1");
  if (o.precision != FCBGA_PREC_FP64 && o.precision != FCBGA_PREC_MIXED) throw be::Error("unknown precision");
  if (o.precond < FCBGA_PC_MG || o.precond > FCBGA_PC_IC) throw be::Error("unknown preconditioner");
  if (o.sweeps < 0 || o.sweeps > 8) throw be::Error("sweeps must be 0 (default 2) .. 8");
  if (o.warm < 0 || o.warm > 64) throw be::Error("warm must be 0 .. 64");
  if (o.smoother != FCBGA_SMOOTH_GS && o.smoother != FCBGA_SMOOTH_COARSE_JACOBI) throw be::Error("unknown smoother");
  return o;
}

// set-up of a preconditioner on the prepared operator (_buildPC, src/core/10-solver.js:868-883); ms
double build_pc(fcbga_model *M, int kind, fcbga_solve_stats &S) {
  const double t0 = now_ms();
  if ((kind == FCBGA_PC_MG || kind == FCBGA_PC_COLUMN) && !M->colsFactored) {
    factor_fine_columns(M);
    M->colsFactored = true;
  }
  if (kind == FCBGA_PC_MG) {
    mg_setup(M);
    S.levels = 1 + (int32_t)M->lv.size();
    S.coarsest_dof = M->lv.back()->ndense;
  } else if (kind == FCBGA_PC_IC) {
    ic_setup(M);
    S.ic_shift = M->icShift;
  }
  be::sync(M->stream);
  return now_ms() - t0;
}

// assembly, column factors and preconditioner set-up of a step (the operator
// and preconditioner every later pcg_run of this step uses)
void prepare_step(fcbga_model *M, const fcbga_step *st, const fcbga_solve_opts &o, fcbga_solve_stats &S) {
  check_step(M, st);
  M->nsweeps = o.sweeps ? o.sweeps : 2;
  // column Jacobi from the third shell level (lv[2]) on
  M->jacobiFrom = o.smoother == FCBGA_SMOOTH_COARSE_JACOBI ? 2 : 1 << 20;
  S.ndof = 3 * (int64_t)M->nn;
  S.ntypes = M->ntypes;
  S.levels = 1;
  be::stream_t s = M->stream;
  const double t0 = now_ms();
  assemble_step(M, st);
  M->colsFactored = false;
  S.sym_tables = M->symTables ? 1 : 0;
  alloc_vectors(M, o.precision);
  be::sync(s);
  S.ms_assemble = now_ms() - t0;
  S.ms_setup = build_pc(M, o.precond, S);
  M->prepGen++;
  M->prepPrecond = o.precond;
  M->prepPrecision = o.precision;
}

// r = b - A x on the scaled system (b = the scaled load bl * s)
void residual_from_x(fcbga_model *M, const double *bl) {
  be::stream_t s = M->stream;
  const int64_t n3 = 3 * M->g.G;
  fine_apply<double>(fine_ops<double>(M), M->x.p, M->q.p, s);
  const double *q = M->q.p, *sv = M->s64.p;
  double *r = M->r.p;
  be::pfor(n3, s, [=] __host__ __device__(int64_t i) { r[i] = bl[i] * sv[i] - q[i]; });
}

// PCG iterations from the current x with r = b - A x (_pcg, src/core/10-solver.js:1001-1042);
// returns the iteration count, rnorm and conv updated
// progress event to the model's callback; false when it asks to stop
bool report(fcbga_model *M, int event, int it, const fcbga_solve_opts &o, double relres, double r0) {
  if (!M->progFn) return true;
  fcbga_progress p{event, M->itersBase + it, o.maxit, o.precond, relres, r0, o.tol, now_ms() - M->solveT0};
  if (M->progFn(&p, M->progUser) != 0) M->cancelled = true;
  return !M->cancelled;
}

int pcg_iterate(fcbga_model *M, const fcbga_solve_opts &o, double bnorm, double &rnorm, bool &conv) {
  be::stream_t s = M->stream;
  const int64_t n3 = 3 * M->g.G;
  double *r = M->r.p, *z = M->z.p, *p = M->p.p, *sc = M->scal.p;
  const double r0 = rnorm / bnorm;
  if (!report(M, 0, 0, o, r0, r0)) return 0;
  if (rnorm / bnorm < o.tol) { conv = true; report(M, 2, 0, o, r0, r0); return 0; }
  precondition(M, o.precision, o.precond);
  be::pfor(n3, s, [=] __host__ __device__(int64_t i) { p[i] = z[i]; });
  be::dot_async(n3, r, z, M->partials.p, sc + SC_RZ, s);
  // the FP32 cycles need the flexible form; IC(0) runs in FP64 whatever the precision
  const bool flex = o.precision == FCBGA_PREC_MIXED && o.precond != FCBGA_PC_IC;
  // every host-side launch choice baked into a captured iteration is part of
  // the key: smoother, precision, preconditioner, sweeps and the element-table
  // form (compact mirror-symmetric or full), which assemble_step decides per step
  const int gkey = (o.smoother ? 8192 : 0) + (M->symTables ? 4096 : 0) + o.precision * 64 + o.precond * 8 + M->nsweeps;
  int it;
  for (it = 1; it <= o.maxit; it++) {
    run_iteration(M, o, flex, gkey);
    double h[2];
    read_scalars(M, h);  // rr, breakdown flag
    if (h[1] != 0) break;  // p.Ap <= 0: stop before the update; the iteration counts, as in the reference
    rnorm = std::sqrt(h[0]);
    const bool done = rnorm / bnorm < o.tol;
    if (M->progFn && (done || it % M->progEvery == 0) && !report(M, 1, it, o, rnorm / bnorm, r0)) break;
    if (done) { conv = true; break; }
  }
  const int n = it > o.maxit ? o.maxit : it;
  if (!M->cancelled) report(M, 2, n, o, rnorm / bnorm, r0);
  return n;
}

// PCG on the prepared step: right-hand side the step's load, or rhs_grid
// (grid layout, unscaled, zero on constrained DOFs); du_dev receives the
// solution in the reference's node order.  A run with MG or IC(0) that does
// not converge continues once with the other from the same iterate
// (LinearSolver.solve, src/core/10-solver.js:965-973).
void pcg_run(fcbga_model *M, const fcbga_solve_opts &o, const double *rhs_grid, double *du_dev, double *du_host,
             fcbga_solve_stats &S) {
  be::stream_t s = M->stream;
  const double t2 = now_ms();
  M->cancelled = false;
  M->solveT0 = t2;
  const int64_t n3 = 3 * M->g.G;
  double *x = M->x.p, *r = M->r.p;
  const double *bl = rhs_grid ? rhs_grid : M->bload.p, *sv = M->s64.p;
  // scaled load; x = 0, r = b
  be::pfor(n3, s, [=] __host__ __device__(int64_t i) {
    double bs = bl[i] * sv[i];
    r[i] = bs;
    x[i] = 0.0;
  });
  const double bnorm = std::sqrt(dotd(M, r, r, SC_RR));
  const bool warm = o.warm > 0 && !rhs_grid;
  if (warm && bnorm != 0 && warm_start(M, bl)) residual_from_x(M, bl);
  int it = 0;
  bool conv = false;
  double rnorm = bnorm == 0 ? 0 : std::sqrt(dotd(M, r, r, SC_RR));
  S.precond = o.precond;
  S.fallback_first = -1;
  if (bnorm == 0) {
    conv = true;
    S.r0 = 0;
  } else {
    S.r0 = rnorm / bnorm;
    M->itersBase = 0;
    it = pcg_iterate(M, o, bnorm, rnorm, conv);
    if (!conv && !M->cancelled && (o.precond == FCBGA_PC_MG || o.precond == FCBGA_PC_IC)) {
      fcbga_solve_opts o2 = o;
      o2.precond = o.precond == FCBGA_PC_MG ? FCBGA_PC_IC : FCBGA_PC_MG;
      S.fallback_first = o.precond;
      S.fallback_iters = it;
      S.fallback_relres = rnorm / bnorm;
      S.ms_setup += build_pc(M, o2.precond, S);
      residual_from_x(M, bl);
      rnorm = std::sqrt(dotd(M, r, r, SC_RR));
      S.r0 = rnorm / bnorm;  // the record's iters0 is that of the last run
      M->itersBase = it;
      it += pcg_iterate(M, o2, bnorm, rnorm, conv);
      S.precond = o2.precond;
    }
  }
  // du = S x in the reference's node order
  {
    const int64_t *gof = M->gridOf.p;
    double *out = du_dev;
    be::pfor(M->nn, s, [=] __host__ __device__(int64_t nidx) {
      int64_t gi = gof[nidx];
      for (int c = 0; c < 3; c++) out[3 * nidx + c] = x[3 * gi + c] * sv[3 * gi + c];
    });
  }
  if (du_host) be::d2h(du_host, du_dev, sizeof(double) * 3 * (size_t)M->nn, s);
  if (warm && conv && bnorm != 0) store_solution(M, o.warm);
  be::sync(s);
  ktimer_collect(M);
  const double t3 = now_ms();
  S.ms_pcg = t3 - t2;
  S.iters = it;
  S.converged = conv ? 1 : 0;
  S.relres = bnorm > 0 ? rnorm / bnorm : 0;
  S.warm = (int32_t)M->basis.size();
}

void step_solve(fcbga_model *M, const fcbga_step *st, const fcbga_solve_opts *op, double *du_dev, double *du_host,
                fcbga_solve_stats *stats) {
  const fcbga_solve_opts o = check_opts(op);
  fcbga_solve_stats S{};
  const double t0 = now_ms();
  prepare_step(M, st, o, S);
  pcg_run(M, o, nullptr, du_dev, du_host, S);
  S.ms_total = now_ms() - t0;
  if (stats) *stats = S;
}

// ---------------------------------------------------------------------------
// stress recovery: element stress histories and point stresses
// ---------------------------------------------------------------------------
void post_tables(fcbga_model *M) {
  if (M->d_eCell.n) return;
  be::stream_t s = M->stream;
  M->d_xs.upload(M->xs, s);
  M->d_ys.upload(M->ys, s);
  M->d_zs.upload(M->zs, s);
  M->d_nodeId.upload(M->node_id, s);
  std::vector<int32_t> ecell(M->ne, -1);
  const int64_t ncells = (int64_t)M->g.ncx * M->g.ncy * M->g.ncz;
  for (int64_t c = 0; c < ncells; c++)
    for (int q = 0; q < 2; q++) {
      const int32_t e = M->cell_elems[2 * c + q];
      if (e >= 0 && ecell[e] < 0) ecell[e] = (int32_t)c;
    }
  M->d_eCell.upload(ecell, s);
}

void check_elems(const fcbga_model *M, int64_t ne, const int32_t *elems) {
  for (int64_t t = 0; t < ne; t++) {
    const int32_t e = elems[t];
    if (e < 0 || e >= M->ne) throw be::Error("element index out of range");
    if (M->e_section[e] >= 0) throw be::Error("stress histories of layered-section elements are not supported");
  }
}

struct ElemGeom {
  int i, j, k;
  double hx, hy, hz;
};

FE_HD ElemGeom elem_geom(const fe::Grid &g, int32_t c, const double *xs, const double *ys, const double *zs) {
  ElemGeom o;
  o.k = c % g.ncz;
  const int32_t cij = c / g.ncz;
  o.j = cij % g.ncy;
  o.i = cij / g.ncy;
  o.hx = xs[o.i + 1] - xs[o.i];
  o.hy = ys[o.j + 1] - ys[o.j];
  o.hz = zs[o.k + 1] - zs[o.k];
  return o;
}

// accumulateStep (src/core/09-assembly.js:494-545) for a list of homogeneous
// elements: states (n x STATE_SIZE) += one step with increment du (3 nn, the
// reference's node order)
// per-group shape matrices, stiffness, thermal stress increment and birth of a
// step on the device (the data of stepRecovery, src/core/09-assembly.js:480-492)
void upload_step_groups(fcbga_model *M, const fcbga_step *st) {
  be::stream_t s = M->stream;
  std::vector<double> Mg((size_t)M->ngroups * 6 * 1089);
  for (int32_t q = 0; q < M->ngroups; q++) shape_matrices(nullptr, st->D + 36 * q, Mg.data() + (size_t)6 * 1089 * q);
  M->postM.upload(Mg, s);
  M->postD.upload(st->D, 36 * (size_t)M->ngroups, s);
  M->postS0.upload(st->s0, 6 * (size_t)M->ngroups, s);
  M->postBorn.upload(st->born, (size_t)M->ngroups, s);
}

// one step added to the states of n elements (device arrays; du in the
// reference's node order) with the group data of upload_step_groups
void recover_states(fcbga_model *M, int64_t n, const int32_t *elems_dev, const double *du_dev, double *states_dev,
                    bool qOnly) {
  post_tables(M);
  const fe::Grid g = M->g;
  const double *xs = M->d_xs.p, *ys = M->d_ys.p, *zs = M->d_zs.p, *Mp = M->postM.p, *Dp = M->postD.p,
               *S0 = M->postS0.p, *dup = du_dev, *gH = M->tabs.p + 99;
  const int32_t *ep = elems_dev, *ecell = M->d_eCell.p, *eg = M->d_egroup.p, *nid = M->d_nodeId.p;
  const uint8_t *born = M->postBorn.p;
  double *stp = states_dev;
  const bool qo = qOnly;
  be::pfor(n, M->stream, [=] __host__ __device__(int64_t t) {
    const int32_t e = ep[t];
    const ElemGeom eg_ = elem_geom(g, ecell[e], xs, ys, zs);
    const int32_t grp = eg[e];
    double du8[24];
    for (int a = 0; a < 8; a++) {
      const int64_t gi = ((int64_t)(eg_.i + fe::corner_li(a)) * g.ny + eg_.j + fe::corner_lj(a)) * g.nz + eg_.k +
                         fe::corner_lk(a);
      const int32_t nd = nid[gi];
      for (int p = 0; p < 3; p++) du8[3 * a + p] = dup ? dup[3 * (int64_t)nd + p] : 0.0;
    }
    fe::element_recover(eg_.hx, eg_.hy, eg_.hz, Mp + (int64_t)grp * 6 * 1089, Dp + 36 * grp, S0 + 6 * grp,
                        born[grp] != 0, qo, gH, du8, stp + (int64_t)t * fe::STATE_SIZE);
  }, 64);
}

// accumulateStep (src/core/09-assembly.js:494-545) for a list of homogeneous
// elements: states (n x STATE_SIZE) += one step with increment du (3 nn, the
// reference's node order)
void state_accumulate(fcbga_model *M, const fcbga_step *st, const double *du, int64_t n, const int32_t *elems,
                      double *states, bool qOnly) {
  check_step(M, st);
  check_elems(M, n, elems);
  post_tables(M);
  be::stream_t s = M->stream;
  upload_step_groups(M, st);
  M->postDu.upload(du, 3 * (size_t)M->nn, s);
  DBuf<int32_t> de;
  DBuf<double> dst;
  de.upload(elems, (size_t)n, s);
  dst.upload(states, (size_t)n * fe::STATE_SIZE, s);
  recover_states(M, n, de.p, M->postDu.p, dst.p, qOnly);
  be::d2h(states, dst.p, sizeof(double) * (size_t)n * fe::STATE_SIZE, s);
}

// stateStress (src/core/09-assembly.js:556-571) at natural points: point p of
// element elems[row[p]] with state row row[p]
void state_points(fcbga_model *M, int64_t n, const int32_t *elems, const double *states, int64_t npts,
                  const int32_t *row, const double *xez, double *sig, double *eps) {
  check_elems(M, n, elems);
  for (int64_t p = 0; p < npts; p++)
    if (row[p] < 0 || row[p] >= n) throw be::Error("point row out of range");
  post_tables(M);
  be::stream_t s = M->stream;
  DBuf<int32_t> de, dr;
  DBuf<double> dst, dx, dsig, deps;
  de.upload(elems, (size_t)n, s);
  dst.upload(states, (size_t)n * fe::STATE_SIZE, s);
  dr.upload(row, (size_t)npts, s);
  dx.upload(xez, 3 * (size_t)npts, s);
  dsig.alloc(6 * (size_t)npts);
  deps.alloc(6 * (size_t)npts);
  const fe::Grid g = M->g;
  const double *xs = M->d_xs.p, *ys = M->d_ys.p, *zs = M->d_zs.p, *stp = dst.p, *xp = dx.p;
  const int32_t *ep = de.p, *rp = dr.p, *ecell = M->d_eCell.p;
  double *sp = dsig.p, *epp = deps.p;
  be::pfor(npts, s, [=] __host__ __device__(int64_t p) {
    const int32_t r = rp[p];
    const ElemGeom eg_ = elem_geom(g, ecell[ep[r]], xs, ys, zs);
    fe::state_stress(eg_.hx, eg_.hy, eg_.hz, stp + (int64_t)r * fe::STATE_SIZE, xp[3 * p], xp[3 * p + 1],
                     xp[3 * p + 2], sp + 6 * p, epp + 6 * p);
  });
  be::d2h(sig, dsig.p, sizeof(double) * 6 * (size_t)npts, s);
  if (eps) be::d2h(eps, deps.p, sizeof(double) * 6 * (size_t)npts, s);
  be::sync(s);
}

#if FCBGA_GPU
// per-group sums of two element arrays, one block per group (fixed order:
// deterministic)
__global__ void __launch_bounds__(256) k_group_sum(int64_t ne, const int32_t *__restrict__ eg,
                                                   const double *__restrict__ a, const double *__restrict__ b,
                                                   int ng, double *__restrict__ out) {
  __shared__ double sa[256], sb[256];
  const int grp = blockIdx.x, tid = threadIdx.x;
  double x = 0, y = 0;
  for (int64_t e = tid; e < ne; e += 256)
    if (eg[e] == grp) {
      x += a[e];
      y += b[e];
    }
  sa[tid] = x;
  sb[tid] = y;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (tid < off) {
      sa[tid] += sa[tid + off];
      sb[tid] += sb[tid + off];
    }
    __syncthreads();
  }
  if (tid == 0) {
    out[grp] = sa[0];
    out[ng + grp] = sb[0];
  }
}
#endif

void group_sums(fcbga_model *M, const double *a, const double *b, double *outA, double *outB) {
  const int ng = M->ngroups;
  std::vector<double> h(2 * (size_t)ng, 0.0);
#if FCBGA_GPU
  DBuf<double> d;
  d.alloc(2 * (size_t)ng);
  hipLaunchKernelGGL(k_group_sum, dim3(ng), dim3(256), 0, M->stream, (int64_t)M->ne, M->d_egroup.p, a, b, ng, d.p);
  BE_CHECK(hipGetLastError());
  be::d2h(h.data(), d.p, sizeof(double) * h.size(), M->stream);
  be::sync(M->stream);
#else
  for (int64_t e = 0; e < M->ne; e++) {
    h[M->e_group[e]] += a[e];
    h[ng + M->e_group[e]] += b[e];
  }
#endif
  for (int q = 0; q < ng; q++) {
    outA[q] = h[q];
    outB[q] = h[ng + q];
  }
}

// ---------------------------------------------------------------------------
// adjoint sensitivities: forward (solve + objective) and backward (one adjoint
// PCG solve + per-element contractions)
// ---------------------------------------------------------------------------
double dev_dot(fcbga_model *M, int64_t n, const double *a, const double *b) {
  AdjTape &T = M->tape;
  T.red.alloc(1);
  be::dot_async(n, a, b, M->partials.p, T.red.p, M->stream);
  double v;
  be::d2h(&v, T.red.p, sizeof(double), M->stream);
  be::sync(M->stream);
  return v;
}

void fill(double *v, int64_t n, double x, be::stream_t s) {
  be::pfor(n, s, [=] __host__ __device__(int64_t i) { v[i] = x; });
}

fcbga_step tape_step(const AdjTape &T) {
  fcbga_step st{};
  st.Ta = T.Ta; st.Tb = T.Tb; st.stage = T.stage;
  st.D = T.D.data(); st.s0 = T.s0.data(); st.born = T.born.data();
  st.secW = T.secW.empty() ? nullptr : T.secW.data();
  st.secV = T.secV.empty() ? nullptr : T.secV.data();
  st.sec_born = T.secBorn.empty() ? nullptr : T.secBorn.data();
  return st;
}

constexpr double GPN = 0.57735026918962576451;  // 1 / sqrt(3)

// Gauss points of the box element, ordered like the nodes (src/core/12-post.js:201-205)
FE_HD void gauss_point(int g, double *x) {
  x[0] = fe::HEX_XI_[g] * GPN; x[1] = fe::HEX_ETA_[g] * GPN; x[2] = fe::HEX_ZETA_[g] * GPN;
}

void adjoint_forward(fcbga_model *M, const fcbga_step *st, const fcbga_solve_opts *op, const fcbga_objective *obj,
                     double *J, double *du_host, fcbga_solve_stats *stats) {
  if (!obj) throw be::Error("objective is NULL");
  if (obj->k_top < 0 || obj->k_top >= M->g.nz) throw be::Error("k_top outside the grid");
  if (obj->n_vm < 0 || (obj->n_vm && !obj->vm_elems)) throw be::Error("vm_elems missing");
  if (obj->n_vm && !(obj->p >= 1)) throw be::Error("p-norm exponent must be >= 1");
  check_elems(M, obj->n_vm, obj->vm_elems);
  const fcbga_solve_opts o = check_opts(op);
  AdjTape &T = M->tape;
  T.valid = false;
  be::stream_t s = M->stream;
  fcbga_solve_stats S{};
  const double t0 = now_ms();
  prepare_step(M, st, o, S);
  T.u.alloc(3 * (size_t)M->nn);
  pcg_run(M, o, nullptr, T.u.p, du_host, S);
  S.ms_total = now_ms() - t0;
  // the step, to re-prepare the same operator in backward if another solve ran in between
  T.Ta = st->Ta; T.Tb = st->Tb; T.stage = st->stage;
  T.D.assign(st->D, st->D + 36 * (size_t)M->ngroups);
  T.s0.assign(st->s0, st->s0 + 6 * (size_t)M->ngroups);
  T.born.assign(st->born, st->born + M->ngroups);
  T.secW.clear(); T.secV.clear(); T.secBorn.clear();
  if (M->nsec) {
    T.secW.assign(st->secW, st->secW + 108 * (size_t)M->nsec);
    T.secV.assign(st->secV, st->secV + 18 * (size_t)M->nsec);
    T.secBorn.assign(st->sec_born, st->sec_born + M->nsec);
  }
  T.opts = o;
  T.gen = M->prepGen;
  T.p = obj->p;
  // warpage proxy: nodes of plane k_top in (i, j) order
  std::vector<int32_t> top;
  for (int i = 0; i < M->g.nx; i++)
    for (int j = 0; j < M->g.ny; j++) {
      const int32_t nd = M->node_id[((int64_t)i * M->g.ny + j) * M->g.nz + obj->k_top];
      if (nd >= 0) top.push_back(nd);
    }
  if (top.empty()) throw be::Error("no nodes on plane k_top");
  T.ntop = (int32_t)top.size();
  T.topNodes.upload(top, s);
  T.work.alloc(std::max<size_t>(top.size(), 8 * (size_t)std::max(obj->n_vm, 1)));
  T.ones.alloc(T.work.n);
  fill(T.ones.p, (int64_t)T.ones.n, 1.0, s);
  {
    const int32_t *tn = T.topNodes.p;
    const double *u = T.u.p;
    double *w = T.work.p;
    be::pfor(T.ntop, s, [=] __host__ __device__(int64_t t) { w[t] = u[3 * (int64_t)tn[t] + 2]; });
    T.wmean = dev_dot(M, T.ntop, w, T.ones.p) / T.ntop;
    const double wm = T.wmean;
    be::pfor(T.ntop, s, [=] __host__ __device__(int64_t t) { w[t] -= wm; });
    T.J[0] = dev_dot(M, T.ntop, w, w);
  }
  // die von Mises p-norm over the Gauss points of vm_elems
  T.nvm = obj->n_vm;
  T.J[1] = 0;
  T.S = 0;
  if (T.nvm) {
    const int64_t n = T.nvm;
    T.vmElems.upload(obj->vm_elems, (size_t)n, s);
    std::vector<int32_t> slot(M->ne, -1);
    for (int64_t q = 0; q < n; q++) {
      if (slot[obj->vm_elems[q]] >= 0) throw be::Error("vm_elems repeats an element");
      slot[obj->vm_elems[q]] = (int32_t)q;
    }
    T.vmSlot.upload(slot, s);
    upload_step_groups(M, st);
    T.states.alloc(n * fe::STATE_SIZE);
    T.statesTh.alloc(n * fe::STATE_SIZE);
    fill(T.states.p, n * fe::STATE_SIZE, 0.0, s);
    fill(T.statesTh.p, n * fe::STATE_SIZE, 0.0, s);
    recover_states(M, n, T.vmElems.p, T.u.p, T.states.p, false);
    recover_states(M, n, T.vmElems.p, nullptr, T.statesTh.p, false);
    T.sig.alloc(n * 48);
    T.sigTh.alloc(n * 48);
    const fe::Grid g = M->g;
    const double *xs = M->d_xs.p, *ys = M->d_ys.p, *zs = M->d_zs.p, *stp = T.states.p, *stth = T.statesTh.p;
    const int32_t *ep = T.vmElems.p, *ecell = M->d_eCell.p;
    double *sg = T.sig.p, *sgt = T.sigTh.p, *w = T.work.p;
    const double p = T.p;
    be::pfor(n * 8, s, [=] __host__ __device__(int64_t q) {
      const int64_t t = q / 8;
      const int gpi = (int)(q % 8);
      const ElemGeom eg_ = elem_geom(g, ecell[ep[t]], xs, ys, zs);
      double x[3], eps[6];
      gauss_point(gpi, x);
      fe::state_stress(eg_.hx, eg_.hy, eg_.hz, stp + t * fe::STATE_SIZE, x[0], x[1], x[2], sg + 6 * q, eps);
      fe::state_stress(eg_.hx, eg_.hy, eg_.hz, stth + t * fe::STATE_SIZE, x[0], x[1], x[2], sgt + 6 * q, eps);
      w[q] = pow(fe::von_mises(sg + 6 * q), p);
    });
    T.S = dev_dot(M, n * 8, w, T.ones.p);
    T.J[1] = T.S > 0 ? pow(T.S, 1.0 / p) : 0.0;
  }
  T.valid = true;
  J[0] = T.J[0];
  J[1] = T.J[1];
  if (stats) *stats = S;
}

// dJdu (optional, host, 3 nn in node order): a gradient of the caller's own
// objective with respect to the step solution, added to the adjoint load
void adjoint_backward(fcbga_model *M, const double *gJ, double *dlnE, double *dlnA, double *dlnAe,
                      fcbga_solve_stats *stats, const double *dJdu = nullptr) {
  AdjTape &T = M->tape;
  if (!T.valid) throw be::Error("adjoint backward without a forward on this model");
  be::stream_t s = M->stream;
  fcbga_solve_stats S{};
  const double t0 = now_ms();
  const fcbga_step st = tape_step(T);
  if (M->prepGen != T.gen) {  // another solve replaced the operator: rebuild the forward's
    prepare_step(M, &st, T.opts, S);
    T.gen = M->prepGen;
  }
  const int64_t nn = M->nn, n3 = 3 * nn;
  T.rhs.alloc(n3);
  fill(T.rhs.p, n3, 0.0, s);
  // dJ_w/du: 2 (w - mean w) on the z components of the plane's nodes
  {
    const int32_t *tn = T.topNodes.p;
    const double *u = T.u.p;
    double *r = T.rhs.p;
    const double wm = T.wmean, g0 = gJ[0];
    be::pfor(T.ntop, s, [=] __host__ __device__(int64_t t) {
      const int64_t nd = tn[t];
      r[3 * nd + 2] = g0 * 2 * (u[3 * nd + 2] - wm);
    });
  }
  const int64_t nvm = T.nvm;
  const bool vm = nvm > 0 && gJ[1] != 0 && T.S > 0;
  if (vm) {
    upload_step_groups(M, &st);
    post_tables(M);
    // dJ_vm/dsigma at each Gauss point: S^(1/p - 1) sigma_vm^(p-1) dsigma_vm/dsigma
    T.dJds.alloc(nvm * 48);
    {
      const double *sg = T.sig.p;
      double *d = T.dJds.p;
      const double p = T.p, fac = gJ[1] * pow(T.S, 1.0 / p - 1.0);
      be::pfor(nvm * 8, s, [=] __host__ __device__(int64_t q) {
        const double *x = sg + 6 * q;
        const double v = fe::von_mises(x);
        double *o = d + 6 * q;
        if (!(v > 0)) {
          for (int c = 0; c < 6; c++) o[c] = 0;
          return;
        }
        const double f = fac * pow(v, p - 1) / v;
        o[0] = f * (2 * x[0] - x[1] - x[2]) * 0.5;
        o[1] = f * (2 * x[1] - x[2] - x[0]) * 0.5;
        o[2] = f * (2 * x[2] - x[0] - x[1]) * 0.5;
        o[3] = f * 3 * x[3];
        o[4] = f * 3 * x[4];
        o[5] = f * 3 * x[5];
      });
    }
    // per element: dJ/du_e (24)
    T.gvm.alloc(nvm * 24);
    {
      const fe::Grid g = M->g;
      const double *xs = M->d_xs.p, *ys = M->d_ys.p, *zs = M->d_zs.p, *Mp = M->postM.p, *Dp = M->postD.p,
                   *d = T.dJds.p;
      const int32_t *ep = T.vmElems.p, *ecell = M->d_eCell.p, *eg = M->d_egroup.p;
      double *gv = T.gvm.p;
      be::pfor(nvm, s, [=] __host__ __device__(int64_t t) {
        const int32_t e = ep[t];
        const ElemGeom eg_ = elem_geom(g, ecell[e], xs, ys, zs);
        const int32_t grp = eg[e];
        double xez[24];
        for (int q = 0; q < 8; q++) gauss_point(q, xez + 3 * q);
        fe::element_stress_pullback(eg_.hx, eg_.hy, eg_.hz, Mp + (int64_t)grp * 6 * 1089, Dp + 36 * grp, 8, xez,
                                    d + 48 * t, gv + 24 * t);
      }, 64);
    }
    // gathered by node (fixed order over the node's cells: deterministic)
    {
      const fe::Grid g = M->g;
      const int64_t *gof = M->gridOf.p;
      const int32_t *ce = M->cellElem.p, *slot = T.vmSlot.p;
      const double *gv = T.gvm.p;
      double *r = T.rhs.p;
      be::pfor(nn, s, [=] __host__ __device__(int64_t nd) {
        const int64_t gi = gof[nd];
        const int k = (int)(gi % g.nz);
        const int64_t cij = gi / g.nz;
        const int j = (int)(cij % g.ny), i = (int)(cij / g.ny);
        double acc[3] = {0, 0, 0};
        for (int ci = i - 1; ci <= i; ci++)
          for (int cj = j - 1; cj <= j; cj++)
            for (int ck = k - 1; ck <= k; ck++) {
              if (ci < 0 || cj < 0 || ck < 0 || ci >= g.ncx || cj >= g.ncy || ck >= g.ncz) continue;
              const int64_t c = ((int64_t)ci * g.ncy + cj) * g.ncz + ck;
              for (int q = 0; q < 2; q++) {
                const int32_t e = ce[2 * c + q];
                if (e < 0 || slot[e] < 0) continue;
                const int a = fe::corner_of(i - ci, j - cj, k - ck);
                for (int p = 0; p < 3; p++) acc[p] += gv[24 * (int64_t)slot[e] + 3 * a + p];
              }
            }
        for (int p = 0; p < 3; p++) r[3 * nd + p] += acc[p];
      });
    }
  }
  if (dJdu) {
    T.dJdu.upload(dJdu, (size_t)n3, s);
    const double *gu = T.dJdu.p;
    double *r = T.rhs.p;
    be::pfor(n3, s, [=] __host__ __device__(int64_t i) { r[i] += gu[i]; });
  }
  // adjoint solve K lambda = dJ/du with the forward's operator and preconditioner
  {
    const int64_t G3 = 3 * M->g.G;
    T.rhsGrid.alloc(G3);
    fill(T.rhsGrid.p, G3, 0.0, s);
    const int64_t *gof = M->gridOf.p;
    const uint8_t *mask = M->dofmask.p;
    const double *r = T.rhs.p;
    double *rg = T.rhsGrid.p;
    be::pfor(nn, s, [=] __host__ __device__(int64_t nd) {
      const int64_t gi = gof[nd];
      for (int p = 0; p < 3; p++) rg[3 * gi + p] = mask[3 * gi + p] ? r[3 * nd + p] : 0.0;
    });
    T.lam.alloc(n3);
    pcg_run(M, T.opts, T.rhsGrid.p, T.lam.p, nullptr, S);
  }
  // per element: d/dlnE = lambda.(f_e - K_e u_e), d/dlnCTE = lambda.f_e (K_e and f_e scale with E,
  // f_e with CTE; the element tables are the step's condensed ones)
  post_tables(M);
  const int64_t ne = M->ne;
  T.dE.alloc(ne);
  T.dA.alloc(ne);
  {
    const fe::Grid g = M->g;
    const int32_t *ecell = M->d_eCell.p, *ce = M->cellElem.p, *ct = M->cellType.p, *nid = M->d_nodeId.p;
    const double *cs = M->cellScale64.p, *Kc = M->Kc64.p, *fc = M->fc64.p, *u = T.u.p, *lam = T.lam.p;
    double *dE = T.dE.p, *dA = T.dA.p;
    be::pfor(ne, s, [=] __host__ __device__(int64_t e) {
      const int32_t c = ecell[e];
      const int q = ce[2 * (int64_t)c] == (int32_t)e ? 0 : 1;
      const int32_t t = ct[2 * (int64_t)c + q];
      const double sc = cs[2 * (int64_t)c + q];
      const int k = c % g.ncz;
      const int32_t cij = c / g.ncz;
      const int j = cij % g.ncy, i = cij / g.ncy;
      double ue[24], le[24];
      for (int a = 0; a < 8; a++) {
        const int64_t gi = ((int64_t)(i + fe::corner_li(a)) * g.ny + j + fe::corner_lj(a)) * g.nz + k + fe::corner_lk(a);
        const int64_t nd = nid[gi];
        for (int p = 0; p < 3; p++) { ue[3 * a + p] = u[3 * nd + p]; le[3 * a + p] = lam[3 * nd + p]; }
      }
      const double *K = Kc + 576 * (int64_t)t, *f = fc + 24 * (int64_t)t;
      double ku = 0, lf = 0;
      for (int r = 0; r < 24; r++) {
        double row = 0;
        for (int q2 = 0; q2 < 24; q2++) row += K[24 * r + q2] * ue[q2];
        ku += le[r] * row;
        lf += le[r] * f[r];
      }
      dE[e] = sc * (lf - ku);
      dA[e] = sc * lf;
    });
  }
  // explicit terms of the p-norm at fixed u: dsigma/dlnE = sigma, dsigma/dlnCTE = sigma at u = 0
  if (vm) {
    const int32_t *ep = T.vmElems.p;
    const double *d = T.dJds.p, *sg = T.sig.p, *sgt = T.sigTh.p;
    double *dE = T.dE.p, *dA = T.dA.p;
    be::pfor(nvm, s, [=] __host__ __device__(int64_t t) {
      double aE = 0, aA = 0;
      for (int q = 0; q < 48; q++) {
        aE += d[48 * t + q] * sg[48 * t + q];
        aA += d[48 * t + q] * sgt[48 * t + q];
      }
      dE[ep[t]] += aE;
      dA[ep[t]] += aA;
    });
  }
  group_sums(M, T.dE.p, T.dA.p, dlnE, dlnA);
  if (dlnAe) be::d2h(dlnAe, T.dA.p, sizeof(double) * (size_t)ne, s);
  be::sync(s);
  S.ms_total = now_ms() - t0;
  if (stats) *stats = S;
}

int fail(const std::exception &e, int code) {
  g_err = e.what();
  return code;
}

}  // namespace

#include "submodel.h"
#include "viscoplastic.h"

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

namespace {

// Validation shared by the two screening entry points.
void check_screen_inputs(const fcbga_anand_params *solder, const fcbga_screen_steps *st, int32_t ntemps,
                         int32_t i25) {
  const int32_t n = st->nsteps;
  if (n < 2 || st->cycles < 1 || ntemps < 2 || i25 < -1 || i25 >= ntemps)
    throw be::Error("screening needs nsteps >= 2, cycles >= 1, ntemps >= 2 and i25 in -1..ntemps-1");
  if (!st->T || !st->dt || !st->E || !st->Fth || !st->wt || !st->wi || !st->cycle || !st->same)
    throw be::Error("screening step arrays missing");
  if ((st->Cp == nullptr) != (st->Fp == nullptr)) throw be::Error("pillar needs both Cp and Fp");
  if (!(st->fs > 0 && st->fs <= 1)) throw be::Error("solder height fraction fs must be in (0, 1]");
  if (!(solder->nu < 0.5)) throw be::Error("nu < 0.5 required");
  for (int32_t k = 0; k < n; k++) {
    if (st->wi[k] < 0 || st->wi[k] > ntemps - 2) throw be::Error("table interval wi out of range");
    if (st->cycle[k] < 0 || st->cycle[k] > st->cycles) throw be::Error("cycle index out of range");
    if (k > 0 && !(st->E[k] > 0 && st->dt[k] >= 0)) throw be::Error("E > 0 and dt >= 0 required");
  }
}

fe::AnandParams screen_params(const fcbga_anand_params *p) {
  return fe::AnandParams{p->A, p->QR, p->xi, p->m, p->shat, p->n, p->h0, p->a, p->s0, p->nu};
}

// One device thread per joint through the whole cycle history.
void launch_screen(const fe::AnandParams P, const fe::ScreenSteps h, int32_t i25, const double *tab, int64_t nsites,
                   int32_t ntemps, double *dW, double *dE, double *stab, be::stream_t s) {
  const int64_t stride = (int64_t)ntemps * 6;
  be::pfor(nsites, s, [=] __host__ __device__(int64_t q) {
    double r[3];
    fe::screen_site(P, h, i25, tab + q * stride, r);
    dW[q] = r[0];
    dE[q] = r[1];
    stab[q] = r[2];
  }, 64);
}

// Scratch of the allocation-free screening entry point: the per-step history
// arrays staged for the kernel, FCBGA_SCREEN_NOALLOC_MAX_STEPS steps at most
// (9 doubles, 2 int32 and 1 byte per step: 332 KB).  One copy per device,
// held until the process ends; every call on a device reuses it, so calls
// that overlap on one device race.
constexpr size_t SCR_STEPS = FCBGA_SCREEN_NOALLOC_MAX_STEPS;
constexpr size_t SCR_BYTES = SCR_STEPS * (9 * sizeof(double) + 2 * sizeof(int32_t) + 1);
static __device__ __align__(256) unsigned char scratch[SCR_BYTES];

struct ScreenScratch {
  double *T, *dt, *E, *Fth, *wt, *Cp, *Fp;
  int32_t *wi, *cycle;
  uint8_t *same;
};

// The device address of this device's scratch array, looked up once per
// device (with that device current) and kept: the array's name is a host-side
// stand-in, never a device address.
unsigned char *scratch_base() {
#if FCBGA_GPU
  constexpr int MAXDEV = 64;
  static unsigned char *addr[MAXDEV] = {};
  static std::mutex mu;
  int dev = 0;
  BE_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAXDEV) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || !p)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(rc));
    addr[dev] = (unsigned char *)p;
  }
  return addr[dev];
#else
  return scratch;
#endif
}

// The scratch array cut into the history arrays of n steps (each 8-byte aligned).
ScreenScratch screen_scratch(int32_t n) {
  unsigned char *b = scratch_base();
  ScreenScratch w;
  double *d = (double *)b;
  w.T = d; w.dt = d + SCR_STEPS; w.E = d + 2 * SCR_STEPS; w.Fth = d + 3 * SCR_STEPS; w.wt = d + 4 * SCR_STEPS;
  w.Cp = d + 5 * SCR_STEPS; w.Fp = d + 8 * SCR_STEPS;
  int32_t *ii = (int32_t *)(d + 9 * SCR_STEPS);
  w.wi = ii; w.cycle = ii + SCR_STEPS;
  w.same = (uint8_t *)(ii + 2 * SCR_STEPS);
  (void)n;
  return w;
}

}  // namespace

extern "C" {

int fcbga_model_create(const fcbga_mesh *mesh, fcbga_model **out) {
  if (!out) { g_err = "out is NULL"; return FCBGA_ERR_ARG; }
  *out = nullptr;
  fcbga_model *M = new fcbga_model();
  {
    const char *ev = getenv("FCBGA_FULL_TABLES");
    M->sym_allowed = !(ev && ev[0] == '1');
  }
  try {
    model_create(mesh, M);
  } catch (const be::Error &e) {
    delete M;
    return fail(e, std::string(e.what()).rfind("HIP", 0) == 0 ? FCBGA_ERR_DEVICE : FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    delete M;
    return fail(e, FCBGA_ERR_INTERNAL);
  }
  *out = M;
  return FCBGA_OK;
}

void fcbga_model_destroy(fcbga_model *m) {
  if (!m) return;
#if FCBGA_GPU
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  for (auto &g : m->graphs) (void)hipGraphExecDestroy(g.second);
  if (m->hostScal) (void)hipHostFree(m->hostScal);
  for (hipEvent_t e : m->evPool) (void)hipEventDestroy(e);
  for (auto &u : m->evUsed) { (void)hipEventDestroy(u.second.first); (void)hipEventDestroy(u.second.second); }
  hipStream_t s = m->stream;
  delete m;
  if (s) (void)hipStreamDestroy(s);
#else
  delete m;
#endif
}

int64_t fcbga_model_ndof(const fcbga_model *m) { return m ? 3 * (int64_t)m->nn : 0; }

static int solve_common(fcbga_model *m, const fcbga_step *step, const fcbga_solve_opts *opts, double *du_dev,
                        double *du_host, fcbga_solve_stats *stats) {
  if (!m) { g_err = "model is NULL"; return FCBGA_ERR_ARG; }
  try {
    DBuf<double> tmp;
    double *dst = du_dev;
    if (!dst) {
      tmp.alloc(3 * (size_t)m->nn);
      dst = tmp.p;
    }
    fcbga_solve_stats S{};
    step_solve(m, step, opts, dst, du_host, &S);
    if (stats) *stats = S;
    if (m->cancelled) {
      g_err = "cancelled";
      return FCBGA_ERR_CANCELLED;
    }
    if (!S.converged) {
      char buf[160];
      snprintf(buf, sizeof buf, "Linear solver did not converge: relative residual %.2e after %d iterations",
               S.relres, S.iters);
      g_err = buf;
      return FCBGA_ERR_NOT_CONVERGED;
    }
  } catch (const be::Error &e) {
    return fail(e, std::string(e.what()).rfind("HIP", 0) == 0 ? FCBGA_ERR_DEVICE : FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_INTERNAL);
  }
  return FCBGA_OK;
}

int fcbga_set_progress_cb(fcbga_model *m, fcbga_progress_fn fn, void *user, int32_t every) {
  if (!m) { g_err = "model is NULL"; return FCBGA_ERR_ARG; }
  if (every < 1) { g_err = "every must be >= 1"; return FCBGA_ERR_ARG; }
  m->progFn = fn;
  m->progUser = user;
  m->progEvery = every;
  return FCBGA_OK;
}

int fcbga_warm_reset(fcbga_model *m) {
  if (!m) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  m->basis.clear();
  return FCBGA_OK;
}

int fcbga_step_solve(fcbga_model *m, const fcbga_step *step, const fcbga_solve_opts *opts, double *du_out,
                     fcbga_solve_stats *stats) {
  if (!du_out) { g_err = "du_out is NULL"; return FCBGA_ERR_ARG; }
  return solve_common(m, step, opts, nullptr, du_out, stats);
}

int fcbga_step_solve_device(fcbga_model *m, const fcbga_step *step, const fcbga_solve_opts *opts, double *du_dev,
                            fcbga_solve_stats *stats) {
  if (!du_dev) { g_err = "du_dev is NULL"; return FCBGA_ERR_ARG; }
  return solve_common(m, step, opts, du_dev, nullptr, stats);
}

int fcbga_path_reset(fcbga_model *m) {
  if (!m) { g_err = "model is NULL"; return FCBGA_ERR_ARG; }
  m->pathDu.clear();
  return FCBGA_OK;
}

int fcbga_path_solve(fcbga_model *m, int32_t k, const fcbga_step *step, const fcbga_solve_opts *opts,
                     fcbga_solve_stats *stats) {
  if (!m || k < 0) { g_err = "model is NULL or k < 0"; return FCBGA_ERR_ARG; }
  try {
    if ((size_t)k >= m->pathDu.size()) m->pathDu.resize(k + 1);
    if (!m->pathDu[k]) m->pathDu[k] = std::make_unique<DBuf<double>>();
    m->pathDu[k]->alloc(3 * (size_t)m->nn);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return solve_common(m, step, opts, m->pathDu[k]->p, nullptr, stats);
}

int fcbga_path_sum(fcbga_model *m, int32_t n, const int32_t *ids, double *out) {
  if (!m || n < 0 || (n && !ids) || !out) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    for (int32_t q = 0; q < n; q++)
      if (ids[q] < 0 || (size_t)ids[q] >= m->pathDu.size() || !m->pathDu[ids[q]])
        throw be::Error("path step not solved");
    const int64_t n3 = 3 * (int64_t)m->nn;
    be::stream_t s = m->stream;
    m->pathSum.alloc(n3);
    double *acc = m->pathSum.p;
    be::pfor(n3, s, [=] __host__ __device__(int64_t i) { acc[i] = 0.0; });
    // in list order, as the reference adds the increments
    for (int32_t q = 0; q < n; q++) {
      const double *d = m->pathDu[ids[q]]->p;
      be::pfor(n3, s, [=] __host__ __device__(int64_t i) { acc[i] += d[i]; });
    }
    be::d2h(out, acc, sizeof(double) * n3, s);
  } catch (const be::Error &e) {
    return fail(e, FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_path_du(fcbga_model *m, int32_t k, float *out) {
  if (!m || !out || k < 0 || (size_t)k >= m->pathDu.size() || !m->pathDu[k]) {
    g_err = "path step not solved";
    return FCBGA_ERR_ARG;
  }
  try {
    const int64_t n3 = 3 * (int64_t)m->nn;
    DBuf<float> f;
    f.alloc(n3);
    const double *d = m->pathDu[k]->p;
    float *fp = f.p;
    be::pfor(n3, m->stream, [=] __host__ __device__(int64_t i) { fp[i] = (float)d[i]; });
    be::d2h(out, fp, sizeof(float) * n3, m->stream);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_path_load(fcbga_model *m, int32_t k, const float *du) {
  if (!m || !du || k < 0) { g_err = "model or du is NULL, or k < 0"; return FCBGA_ERR_ARG; }
  try {
    if ((size_t)k >= m->pathDu.size()) m->pathDu.resize(k + 1);
    if (!m->pathDu[k]) m->pathDu[k] = std::make_unique<DBuf<double>>();
    const int64_t n3 = 3 * (int64_t)m->nn;
    m->pathDu[k]->alloc(n3);
    DBuf<float> f;
    f.alloc(n3);
    be::h2d(f.p, du, sizeof(float) * n3, m->stream);
    const float *fp = f.p;
    double *d = m->pathDu[k]->p;
    // exact: every FP32 value is an FP64 value
    be::pfor(n3, m->stream, [=] __host__ __device__(int64_t i) { d[i] = (double)fp[i]; });
    be::sync(m->stream);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_path_load64(fcbga_model *m, int32_t k, const double *du) {
  if (!m || !du || k < 0) { g_err = "model or du is NULL, or k < 0"; return FCBGA_ERR_ARG; }
  try {
    if ((size_t)k >= m->pathDu.size()) m->pathDu.resize(k + 1);
    if (!m->pathDu[k]) m->pathDu[k] = std::make_unique<DBuf<double>>();
    const int64_t n3 = 3 * (int64_t)m->nn;
    m->pathDu[k]->alloc(n3);
    be::h2d(m->pathDu[k]->p, du, sizeof(double) * n3, m->stream);
    be::sync(m->stream);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_path_copy(fcbga_model *dst, const fcbga_model *src, int32_t k) {
  if (!dst || !src || k < 0 || (size_t)k >= src->pathDu.size() || !src->pathDu[k]) {
    g_err = "model is NULL or the source has not kept step k";
    return FCBGA_ERR_ARG;
  }
  if (dst->nn != src->nn) { g_err = "models of different meshes"; return FCBGA_ERR_ARG; }
  if (dst == src) return FCBGA_OK;
  try {
    if ((size_t)k >= dst->pathDu.size()) dst->pathDu.resize(k + 1);
    if (!dst->pathDu[k]) dst->pathDu[k] = std::make_unique<DBuf<double>>();
    const int64_t n3 = 3 * (int64_t)dst->nn;
    dst->pathDu[k]->alloc(n3);
    // the source's kernels for step k are finished before its bytes are read
    be::sync(src->stream);
    be::d2d(dst->pathDu[k]->p, src->pathDu[k]->p, sizeof(double) * n3, dst->stream);
    be::sync(dst->stream);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_prepare_step(int32_t ngroups, const fcbga_group_grids *groups, int32_t nsections,
                       const fcbga_section_grids *sections, double Ta, double Tb, int32_t stage, double *D,
                       double *s0, uint8_t *born, double *secW, double *secV, uint8_t *sec_born) {
  static_assert(FCBGA_TGRID_MIN == sd::TGRID_MIN && FCBGA_TGRID_N == sd::TGRID_N, "property grid");
  if (ngroups < 0 || nsections < 0 || (ngroups && (!groups || !D || !s0 || !born)) ||
      (nsections && (!sections || !secW || !secV || !sec_born))) {
    g_err = "NULL argument or negative size";
    return FCBGA_ERR_ARG;
  }
  if (!std::isfinite(Ta) || !std::isfinite(Tb)) {
    g_err = "step temperatures must be finite";
    return FCBGA_ERR_ARG;
  }
  for (int32_t q = 0; q < ngroups; q++) {
    const fcbga_group_grids &g = groups[q];
    if (!g.layered && (!g.F || (!g.Dgrid && !g.Dconst))) {
      g_err = "group " + std::to_string(q) + ": thermal strain grid and a stiffness grid or constant required";
      return FCBGA_ERR_ARG;
    }
  }
  for (int32_t q = 0; q < nsections; q++) {
    const fcbga_section_grids &c = sections[q];
    if (c.ns < 0 || c.group < -1 || c.group >= ngroups || !c.W || (c.ns && (!c.D || !c.F || !c.Ls))) {
      g_err = "section " + std::to_string(q) + ": grids missing, negative sub-layer count or group out of range";
      return FCBGA_ERR_ARG;
    }
  }
  for (int32_t q = 0; q < ngroups; q++) {
    const fcbga_group_grids &g = groups[q];
    double *Dq = D + 36 * (int64_t)q, *sq = s0 + 6 * (int64_t)q;
    born[q] = g.stage <= stage ? 1 : 0;
    for (int r = 0; r < 6; r++) sq[r] = 0;
    if (g.layered) {
      for (int r = 0; r < 36; r++) Dq[r] = 0;
      continue;
    }
    if (g.Dgrid) sd::step_mean(g.Dgrid, 36, 36, Ta, Tb, Dq);
    else for (int r = 0; r < 36; r++) Dq[r] = g.Dconst[r];
    if (born[q]) sd::step_stress(g.Dgrid, g.Dconst, g.F, Ta, Tb, sq);
  }
  for (int32_t q = 0; q < nsections; q++) {
    const fcbga_section_grids &c = sections[q];
    sd::step_mean(c.W, 108, 108, Ta, Tb, secW + 108 * (int64_t)q);
    sec_born[q] = c.group < 0 || groups[c.group].stage <= stage ? 1 : 0;
    double *V = secV + 18 * (int64_t)q;
    if (sec_born[q]) sd::section_stress_weights(c.ns, c.D, c.F, c.Ls, Ta, Tb, V);
    else for (int r = 0; r < 18; r++) V[r] = 0;
  }
  return FCBGA_OK;
}

int fcbga_mesh_fingerprint(int32_t nn, int32_t ne, const double *xs, int32_t nx, const double *ys, int32_t ny,
                           const double *zs, int32_t nz, const int32_t *e_group, const int8_t *e_part,
                           const double *e_scale, char out[15]) {
  if (!out || nx < 0 || ny < 0 || nz < 0 || ne < 0 || (nx && !xs) || (ny && !ys) || (nz && !zs) ||
      (ne && (!e_group || !e_part || !e_scale))) {
    g_err = "NULL argument or negative size";
    return FCBGA_ERR_ARG;
  }
  // two 32-bit multiplicative hashes over the bytes (Math.imul is the low 32 bits of the product)
  uint32_t h1 = 0xdeadbeefu ^ (uint32_t)nn, h2 = 0x41c6ce57u ^ (uint32_t)ne;
  const auto feed = [&](const void *p, size_t nbytes) {
    const uint8_t *b = static_cast<const uint8_t *>(p);
    for (size_t i = 0; i < nbytes; i++) {
      h1 = (h1 ^ b[i]) * 2654435761u;
      h2 = (h2 ^ b[i]) * 1597334677u;
    }
  };
  feed(xs, sizeof(double) * (size_t)nx);
  feed(ys, sizeof(double) * (size_t)ny);
  feed(zs, sizeof(double) * (size_t)nz);
  feed(e_group, sizeof(int32_t) * (size_t)ne);
  feed(e_part, sizeof(int8_t) * (size_t)ne);
  feed(e_scale, sizeof(double) * (size_t)ne);
  h1 = ((h1 ^ (h1 >> 16)) * 2246822507u) ^ ((h2 ^ (h2 >> 13)) * 3266489909u);
  h2 = ((h2 ^ (h2 >> 16)) * 2246822507u) ^ ((h1 ^ (h1 >> 13)) * 3266489909u);
  const unsigned long long v = ((unsigned long long)(h2 & 0x1fffffu) << 32) | h1;
  snprintf(out, 15, "%014llx", v);
  return FCBGA_OK;
}

int fcbga_adjoint_forward(fcbga_model *m, const fcbga_step *step, const fcbga_solve_opts *opts,
                          const fcbga_objective *obj, double *J, double *du_out, fcbga_solve_stats *stats) {
  if (!m || !step || !obj || !J) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    fcbga_solve_stats S{};
    adjoint_forward(m, step, opts, obj, J, du_out, &S);
    if (stats) *stats = S;
    if (m->cancelled) {
      g_err = "cancelled";
      return FCBGA_ERR_CANCELLED;
    }
    if (!S.converged) {
      g_err = "not converged";
      return FCBGA_ERR_NOT_CONVERGED;
    }
  } catch (const be::Error &e) {
    return fail(e, std::string(e.what()).rfind("HIP", 0) == 0 ? FCBGA_ERR_DEVICE : FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_INTERNAL);
  }
  return FCBGA_OK;
}

static int adjoint_backward_entry(fcbga_model *m, const double *gJ, const double *dJdu, double *dlnE,
                                  double *dlnA, double *dlnA_elem, fcbga_solve_stats *stats);

int fcbga_adjoint_backward_u(fcbga_model *m, const double *dJdu, double *dlnE, double *dlnA, double *dlnA_elem,
                             fcbga_solve_stats *stats) {
  if (!m || !dJdu) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  const double gJ[2] = {0.0, 0.0};
  return adjoint_backward_entry(m, gJ, dJdu, dlnE, dlnA, dlnA_elem, stats);
}

int fcbga_adjoint_backward(fcbga_model *m, const double *gJ, double *dlnE, double *dlnA, double *dlnA_elem,
                           fcbga_solve_stats *stats) {
  if (!m || !gJ) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  return adjoint_backward_entry(m, gJ, nullptr, dlnE, dlnA, dlnA_elem, stats);
}

static int adjoint_backward_entry(fcbga_model *m, const double *gJ, const double *dJdu, double *dlnE,
                                  double *dlnA, double *dlnA_elem, fcbga_solve_stats *stats) {
  if (!dlnE || !dlnA) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    fcbga_solve_stats S{};
    adjoint_backward(m, gJ, dlnE, dlnA, dlnA_elem, &S, dJdu);
    if (stats) *stats = S;
    if (m->cancelled) {
      g_err = "cancelled";
      return FCBGA_ERR_CANCELLED;
    }
    if (!S.converged) {
      g_err = "adjoint solve not converged";
      return FCBGA_ERR_NOT_CONVERGED;
    }
  } catch (const be::Error &e) {
    return fail(e, std::string(e.what()).rfind("HIP", 0) == 0 ? FCBGA_ERR_DEVICE : FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_INTERNAL);
  }
  return FCBGA_OK;
}

int fcbga_state_accumulate(fcbga_model *m, const fcbga_step *step, const double *du, int64_t nelem,
                           const int32_t *elems, double *states, int32_t q_only) {
  if (!m || !step || !du || nelem < 0 || (nelem && (!elems || !states))) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  if (!nelem) return FCBGA_OK;
  try {
    state_accumulate(m, step, du, nelem, elems, states, q_only != 0);
  } catch (const be::Error &e) {
    return fail(e, std::string(e.what()).rfind("HIP", 0) == 0 ? FCBGA_ERR_DEVICE : FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_INTERNAL);
  }
  return FCBGA_OK;
}

int fcbga_state_stress(fcbga_model *m, int64_t nelem, const int32_t *elems, const double *states, int64_t npts,
                       const int32_t *row, const double *xez, double *sig, double *eps) {
  if (!m || nelem < 0 || npts < 0 || (npts && (!elems || !states || !row || !xez || !sig))) {
    g_err = "NULL argument";
    return FCBGA_ERR_ARG;
  }
  if (!npts) return FCBGA_OK;
  try {
    state_points(m, nelem, elems, states, npts, row, xez, sig, eps);
  } catch (const be::Error &e) {
    return fail(e, std::string(e.what()).rfind("HIP", 0) == 0 ? FCBGA_ERR_DEVICE : FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_INTERNAL);
  }
  return FCBGA_OK;
}

int fcbga_step_load(fcbga_model *m, const fcbga_step *step, double *f_out) {
  if (!m || !f_out) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    check_step(m, step);
    assemble_step(m, step);
    DBuf<double> out;
    out.alloc(3 * (size_t)m->nn);
    const int64_t *gof = m->gridOf.p;
    const double *bl = m->bload.p;
    double *o = out.p;
    be::pfor(m->nn, m->stream, [=] __host__ __device__(int64_t n) {
      for (int c = 0; c < 3; c++) o[3 * n + c] = bl[3 * gof[n] + c];
    });
    be::d2h(f_out, out.p, sizeof(double) * 3 * (size_t)m->nn, m->stream);
  } catch (const be::Error &e) {
    return fail(e, FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_INTERNAL);
  }
  return FCBGA_OK;
}

int fcbga_step_apply(fcbga_model *m, const fcbga_step *step, const double *x, double *y_out) {
  if (!m || !x || !y_out) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    check_step(m, step);
    assemble_step(m, step);
    const fe::Grid g = m->g;
    DBuf<double> xin, xg, yg, out;
    xin.upload(x, 3 * (size_t)m->nn, m->stream);
    xg.alloc(3 * g.G);
    yg.alloc(3 * g.G);
    out.alloc(3 * (size_t)m->nn);
    xg.zero(m->stream);
    const int64_t *gof = m->gridOf.p;
    const double *xi = xin.p;
    double *xgp = xg.p, *ygp = yg.p, *o = out.p;
    be::pfor(m->nn, m->stream, [=] __host__ __device__(int64_t n) {
      for (int c = 0; c < 3; c++) xgp[3 * gof[n] + c] = xi[3 * n + c];
    });
    // the solver's own operator kernel (full element tables) with unit
    // scaling on every DOF of an existing node: y = K x, no constraints
    DBuf<double> wone;
    wone.alloc(3 * g.G);
    wone.zero(m->stream);
    double *wo = wone.p;
    be::pfor(m->nn, m->stream, [=] __host__ __device__(int64_t n) {
      for (int c = 0; c < 3; c++) wo[3 * gof[n] + c] = 1.0;
    });
    FineOps<double> fo = fine_ops<double>(m);
    fo.w = wone.p;
    fo.sym = false;
    fine_apply<double>(fo, xgp, ygp, m->stream);
    be::pfor(m->nn, m->stream, [=] __host__ __device__(int64_t n) {
      for (int c = 0; c < 3; c++) o[3 * n + c] = ygp[3 * gof[n] + c];
    });
    be::d2h(y_out, out.p, sizeof(double) * 3 * (size_t)m->nn, m->stream);
  } catch (const be::Error &e) {
    return fail(e, FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_INTERNAL);
  }
  return FCBGA_OK;
}

int fcbga_anand_update(const fcbga_anand_params *params, int64_t n, double T, double dt, double E, const double *eps,
                       double *state, double *trial, double *sig, double *D, double *dep) {
  if (!params || n < 0 || !eps || !state || !trial || !sig) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  if (n == 0) return FCBGA_OK;
  try {
    be::stream_t s = nullptr;
#if FCBGA_GPU
    BE_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
#endif
    DBuf<double> de, dst, dtr, dsig, dD, ddep;
    de.upload(eps, 6 * (size_t)n, s);
    dst.upload(state, FCBGA_ANAND_STATE * (size_t)n, s);
    dtr.alloc(FCBGA_ANAND_STATE * (size_t)n);
    dsig.alloc(6 * (size_t)n);
    if (D) dD.alloc(36 * (size_t)n);
    ddep.alloc((size_t)n);
    int rc = fcbga_anand_update_device(params, n, T, dt, E, de.p, dst.p, dtr.p, dsig.p, D ? dD.p : nullptr, ddep.p,
                                       (void *)s);
    if (rc != FCBGA_OK) return rc;
    be::d2h(state, dst.p, sizeof(double) * FCBGA_ANAND_STATE * n, s);
    be::d2h(trial, dtr.p, sizeof(double) * FCBGA_ANAND_STATE * n, s);
    be::d2h(sig, dsig.p, sizeof(double) * 6 * n, s);
    if (D) be::d2h(D, dD.p, sizeof(double) * 36 * n, s);
    if (dep) be::d2h(dep, ddep.p, sizeof(double) * n, s);
#if FCBGA_GPU
    (void)hipStreamDestroy(s);
#endif
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_anand_update_device(const fcbga_anand_params *params, int64_t n, double T, double dt, double E,
                              const double *eps, double *state, double *trial, double *sig, double *D, double *dep,
                              void *stream) {
  if (!params || n < 0 || !eps || !state || !trial || !sig) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  if (!(E > 0) || !(dt >= 0) || !(params->nu < 0.5)) { g_err = "E > 0, dt >= 0 and nu < 0.5 required"; return FCBGA_ERR_ARG; }
  try {
    const fe::AnandParams P{params->A, params->QR, params->xi, params->m, params->shat, params->n, params->h0,
                            params->a, params->s0, params->nu};
    be::pfor(n, (be::stream_t)stream, [=] __host__ __device__(int64_t q) {
      double sg[6];
      double dd = fe::anand_step(P, E, T, dt, eps + 6 * q, state + FCBGA_ANAND_STATE * q,
                                 trial + FCBGA_ANAND_STATE * q, sg, D ? D + 36 * q : nullptr);
      for (int i = 0; i < 6; i++) sig[6 * q + i] = sg[i];
      if (dep) dep[q] = dd;
    }, 64);
    be::sync((be::stream_t)stream);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_screen_joints(const fcbga_anand_params *solder, const fcbga_screen_steps *st, int64_t nsites,
                        int32_t ntemps, int32_t i25, const double *table, double *dW, double *dE, double *stab) {
  if (!solder || !st || nsites < 0 || !table || !dW || !dE || !stab) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  if (nsites == 0) return FCBGA_OK;
  try {
    check_screen_inputs(solder, st, ntemps, i25);
    const int32_t n = st->nsteps;
    be::stream_t s = nullptr;
#if FCBGA_GPU
    BE_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
#endif
    DBuf<double> dT, ddt, dEm, dF, dwt, dCp, dFp, dtab, dout;
    DBuf<int32_t> dwi, dcy;
    DBuf<uint8_t> dsame;
    dT.upload(st->T, n, s); ddt.upload(st->dt, n, s); dEm.upload(st->E, n, s); dF.upload(st->Fth, n, s);
    dwt.upload(st->wt, n, s); dwi.upload(st->wi, n, s); dcy.upload(st->cycle, n, s); dsame.upload(st->same, n, s);
    if (st->Cp) { dCp.upload(st->Cp, 3 * (size_t)n, s); dFp.upload(st->Fp, n, s); }
    dtab.upload(table, (size_t)nsites * ntemps * 6, s);
    dout.alloc(3 * (size_t)nsites);
    fe::ScreenSteps h{n, st->cycles, dT.p, ddt.p, dEm.p, dF.p, dwt.p, st->Cp ? dCp.p : nullptr,
                      st->Cp ? dFp.p : nullptr, dwi.p, dcy.p, dsame.p, st->fs};
    launch_screen(screen_params(solder), h, i25, dtab.p, nsites, ntemps, dout.p, dout.p + nsites,
                  dout.p + 2 * nsites, s);
    be::d2h(dW, dout.p, sizeof(double) * nsites, s);
    be::d2h(dE, dout.p + nsites, sizeof(double) * nsites, s);
    be::d2h(stab, dout.p + 2 * nsites, sizeof(double) * nsites, s);
#if FCBGA_GPU
    (void)hipStreamDestroy(s);
#endif
  } catch (const be::Error &e) {
    return fail(e, FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_screen_joints_noalloc(const fcbga_anand_params *solder, const fcbga_screen_steps *st, int64_t nsites,
                                int32_t ntemps, int32_t i25, const double *table_dev, double *dW_dev,
                                double *dE_dev, double *stab_dev, void *stream) {
  if (!solder || !st || nsites < 0 || !table_dev || !dW_dev || !dE_dev || !stab_dev) {
    g_err = "NULL argument";
    return FCBGA_ERR_ARG;
  }
  try {
    check_screen_inputs(solder, st, ntemps, i25);
    const int32_t n = st->nsteps;
    if (n > FCBGA_SCREEN_NOALLOC_MAX_STEPS)
      throw be::Error("fcbga_screen_joints_noalloc: nsteps exceeds FCBGA_SCREEN_NOALLOC_MAX_STEPS (" +
                      std::to_string(FCBGA_SCREEN_NOALLOC_MAX_STEPS) + ")");
    if (nsites == 0) return FCBGA_OK;
    be::stream_t s = (be::stream_t)stream;
    const ScreenScratch w = screen_scratch(n);
    // stage the history arrays in the scratch array (copies only, no allocation)
    be::h2d(w.T, st->T, sizeof(double) * n, s);
    be::h2d(w.dt, st->dt, sizeof(double) * n, s);
    be::h2d(w.E, st->E, sizeof(double) * n, s);
    be::h2d(w.Fth, st->Fth, sizeof(double) * n, s);
    be::h2d(w.wt, st->wt, sizeof(double) * n, s);
    if (st->Cp) {
      be::h2d(w.Cp, st->Cp, sizeof(double) * 3 * n, s);
      be::h2d(w.Fp, st->Fp, sizeof(double) * n, s);
    }
    be::h2d(w.wi, st->wi, sizeof(int32_t) * n, s);
    be::h2d(w.cycle, st->cycle, sizeof(int32_t) * n, s);
    be::h2d(w.same, st->same, n, s);
    fe::ScreenSteps h{n, st->cycles, w.T, w.dt, w.E, w.Fth, w.wt, st->Cp ? w.Cp : nullptr,
                      st->Cp ? w.Fp : nullptr, w.wi, w.cycle, w.same, st->fs};
    launch_screen(screen_params(solder), h, i25, table_dev, nsites, ntemps, dW_dev, dE_dev, stab_dev, s);
    be::sync(s);
  } catch (const be::Error &e) {
    return fail(e, FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_set_kernel_timing(fcbga_model *m, int enable) {
  if (!m) { g_err = "model is NULL"; return FCBGA_ERR_ARG; }
  m->timing = enable != 0;
  for (int c = 0; c < FCBGA_KCLASS_COUNT; c++) { m->kt_ms[c] = 0; m->kt_n[c] = 0; }
  return FCBGA_OK;
}

int fcbga_kernel_times(const fcbga_model *m, double *ms_total, int32_t *launches) {
  if (!m || !ms_total || !launches) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  for (int c = 0; c < FCBGA_KCLASS_COUNT; c++) { ms_total[c] = m->kt_ms[c]; launches[c] = m->kt_n[c]; }
  return FCBGA_OK;
}

int fcbga_last_error(char *buf, size_t len) {
  if (!buf || !len) return (int)g_err.size();
  size_t n = g_err.size() < len - 1 ? g_err.size() : len - 1;
  std::memcpy(buf, g_err.data(), n);
  buf[n] = 0;
  return (int)n;
}

int fcbga_submodel_create(const fcbga_submodel_mesh *mesh, const fcbga_anand_params *solder, fcbga_submodel **out) {
  if (!out) { g_err = "out is NULL"; return FCBGA_ERR_ARG; }
  *out = nullptr;
  fcbga_submodel *S = new fcbga_submodel();
  try {
    sub_create(mesh, solder, S);
  } catch (const be::Error &e) {
    fcbga_submodel_destroy(S);
    return fail(e, std::string(e.what()).rfind("HIP", 0) == 0 ? FCBGA_ERR_DEVICE : FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    fcbga_submodel_destroy(S);
    return fail(e, FCBGA_ERR_INTERNAL);
  }
  *out = S;
  return FCBGA_OK;
}

void fcbga_submodel_destroy(fcbga_submodel *s) {
  if (!s) return;
#if FCBGA_GPU
  if (s->stream) (void)hipStreamSynchronize(s->stream);
#endif
  be::stream_t st = s->stream;
  delete s;
#if FCBGA_GPU
  if (st) (void)hipStreamDestroy(st);
#else
  (void)st;
#endif
}

int32_t fcbga_submodel_nboundary(const fcbga_submodel *s) { return s ? s->nb : -1; }

int fcbga_submodel_step(fcbga_submodel *s, const fcbga_submodel_step_in *in, fcbga_submodel_step_out *out) {
  if (!s || !in || !out) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    sub_step(s, *in, *out);
  } catch (const be::Error &e) {
    return fail(e, std::string(e.what()).rfind("HIP", 0) == 0 ? FCBGA_ERR_DEVICE : FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_INTERNAL);
  }
  return FCBGA_OK;
}

int fcbga_submodel_save(fcbga_submodel *s) {
  if (!s) { g_err = "submodel is NULL"; return FCBGA_ERR_ARG; }
  try {
    sub_save(s);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_submodel_restore(fcbga_submodel *s) {
  if (!s) { g_err = "submodel is NULL"; return FCBGA_ERR_ARG; }
  try {
    sub_restore(s);
  } catch (const be::Error &e) {
    return fail(e, FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_submodel_layer_means(fcbga_submodel *s, double *out4) {
  if (!s || !out4) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    sub_layer_means(s, out4);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_submodel_loop_sample(fcbga_submodel *s, double T, double E, double fsolder, double cphi, double sphi,
                               double *out4) {
  if (!s || !out4) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  if (!(E > 0)) { g_err = "E > 0 required"; return FCBGA_ERR_ARG; }
  try {
    sub_loop_sample(s, T, E, fsolder, cphi, sphi, out4);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_submodel_section(fcbga_submodel *s, double E, double fsolder, float *out) {
  if (!s || !out) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  if (!(E > 0)) { g_err = "E > 0 required"; return FCBGA_ERR_ARG; }
  try {
    sub_section(s, E, fsolder, out);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_vp_create(const fcbga_mesh *mesh, const fcbga_vp_solder *solder, fcbga_vp **out) {
  if (!out || !mesh || !solder) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  *out = nullptr;
  fcbga_vp *V = new fcbga_vp();
  try {
    vp_create(mesh, solder, V);
  } catch (const be::Error &e) {
    fcbga_vp_destroy(V);
    return fail(e, std::string(e.what()).rfind("HIP", 0) == 0 ? FCBGA_ERR_DEVICE : FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    fcbga_vp_destroy(V);
    return fail(e, FCBGA_ERR_INTERNAL);
  }
  *out = V;
  return FCBGA_OK;
}

void fcbga_vp_destroy(fcbga_vp *v) {
  if (!v) return;
  fcbga_model *M = v->M;
  v->M = nullptr;
#if FCBGA_GPU
  if (M && M->stream) (void)hipStreamSynchronize(M->stream);
#endif
  delete v;  // device buffers of the engine first, then the model and its stream
  if (M) fcbga_model_destroy(M);
}

int32_t fcbga_vp_nsolder(const fcbga_vp *v) { return v ? v->ns : 0; }

int fcbga_vp_step(fcbga_vp *v, const fcbga_step *elastic, double T, double dt, const fcbga_vp_opts *opts,
                  fcbga_vp_stats *stats) {
  if (!v || !elastic || !opts) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    if (!(dt > 0) || !std::isfinite(T)) throw be::Error("dt > 0 and a finite T required");
    if (opts->maxit < 1 || !(opts->res_tol > 0) || !(opts->du_tol > 0) || !(opts->inner_floor > 0))
      throw be::Error("Newton options out of range");
    fcbga_vp_opts o = *opts;
    o.inner = check_opts(&opts->inner);
    if (o.inner.precond == FCBGA_PC_IC) throw be::Error("the viscoplastic Newton takes the multigrid inner solver");
    check_step(v->M, elastic);
    const fcbga_vp_stats R = vp_newton(v, elastic, T, dt, o);
    if (stats) *stats = R;
    if (!R.converged) {
      g_err = R.nonfinite ? "viscoplastic step: non-finite material-point stress" : "viscoplastic step did not converge";
      return FCBGA_ERR_NOT_CONVERGED;
    }
  } catch (const be::Error &e) {
    return fail(e, FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_vp_save(fcbga_vp *v) {
  if (!v) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try { vp_save(v); } catch (const std::exception &e) { return fail(e, FCBGA_ERR_DEVICE); }
  return FCBGA_OK;
}

int fcbga_vp_restore(fcbga_vp *v) {
  if (!v) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try { vp_restore(v); } catch (const be::Error &e) { return fail(e, FCBGA_ERR_ARG); }
  catch (const std::exception &e) { return fail(e, FCBGA_ERR_DEVICE); }
  return FCBGA_OK;
}

int fcbga_vp_points(fcbga_vp *v, double *st, int32_t *elems) {
  if (!v || !st) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    be::d2h(st, v->st.p, sizeof(double) * 88 * (size_t)v->ns, v->M->stream);
    if (elems) std::copy(v->elemOf.begin(), v->elemOf.end(), elems);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_vp_displacement(fcbga_vp *v, double *u) {
  if (!v || !u) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    DBuf<double> tmp;
    tmp.alloc(3 * (size_t)v->M->nn);
    vp_to_nodes(v, v->u.p, tmp.p);
    be::d2h(u, tmp.p, sizeof(double) * tmp.n, v->M->stream);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_vp_set_prescribed(fcbga_vp *v, const double *values) {
  if (!v || !values) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    if (v->bcGrid.n) {
      be::h2d(v->bcVal.p, values, sizeof(double) * v->bcVal.n, v->M->stream);
      be::sync(v->M->stream);
    }
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_vp_set_displacement(fcbga_vp *v, const double *u) {
  if (!v || !u) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    DBuf<double> tmp;
    tmp.upload(u, 3 * (size_t)v->M->nn, v->M->stream);
    vp_from_nodes(v, tmp.p, v->u.p);
    be::sync(v->M->stream);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_vp_residual(fcbga_vp *v, const fcbga_step *elastic, double T, double dt, double *r, const double *x,
                      double *y) {
  if (!v || !elastic || !r || (x && !y)) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    if (!(dt > 0)) throw be::Error("dt > 0 required");
    check_step(v->M, elastic);
    fcbga_model *M = v->M;
    be::stream_t s = M->stream;
    vp_elements(v, T, dt, 1);
    fcbga_solve_opts o{1e-8, 1, FCBGA_PREC_FP64, FCBGA_PC_JACOBI, 0, 0, FCBGA_SMOOTH_GS};
    fcbga_solve_stats S{};
    vp_prepare(v, elastic, o, S);
    vp_residual(v);
    const size_t nn3 = 3 * (size_t)M->nn;
    DBuf<double> tmp;
    tmp.alloc(nn3);
    vp_to_nodes(v, v->rg.p, tmp.p);
    be::d2h(r, tmp.p, sizeof(double) * nn3, s);
    if (x) {
      DBuf<double> xg, xn;
      xn.upload(x, nn3, s);
      xg.alloc(3 * M->g.G);
      vp_from_nodes(v, xn.p, xg.p);
      const int64_t n3 = 3 * M->g.G;
      const double *si = M->sinv64.p, *wv = M->w64.p;
      double *xp = xg.p, *q = M->q.p;
      be::pfor(n3, s, [=] __host__ __device__(int64_t i) { xp[i] = wv[i] != 0 ? xp[i] * si[i] : 0.0; });
      fine_apply<double>(fine_ops<double>(M), xp, q, s);
      be::pfor(n3, s, [=] __host__ __device__(int64_t i) { xp[i] = wv[i] != 0 ? q[i] * si[i] : 0.0; });
      vp_to_nodes(v, xp, tmp.p);
      be::d2h(y, tmp.p, sizeof(double) * nn3, s);
    }
  } catch (const be::Error &e) {
    return fail(e, FCBGA_ERR_ARG);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_submodel_points(fcbga_submodel *s, double *st) {
  if (!s || !st) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    be::d2h(st, s->st.p, sizeof(double) * 8 * FCBGA_ANAND_STATE * (size_t)s->ne, s->stream);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

int fcbga_submodel_displacement(fcbga_submodel *s, double *u) {
  if (!s || !u) { g_err = "NULL argument"; return FCBGA_ERR_ARG; }
  try {
    be::d2h(u, s->u.p, sizeof(double) * s->ndof, s->stream);
  } catch (const std::exception &e) {
    return fail(e, FCBGA_ERR_DEVICE);
  }
  return FCBGA_OK;
}

}  // extern "C"
