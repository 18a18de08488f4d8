// submodel.h -- bump submodel engine (BumpSubmodel, src/core/13-fatigue.js:514-1105),
// included by fcbga_hip.hip.
//
// One pitch cell around a bump on a full structured grid of B-bar hexahedra
// (one element per cell, nodes and cells z fastest): substrate slice,
// solder, underfill and silicon; the nodes on the six faces carry prescribed
// displacements.  Per time step: Newton-Raphson with the algorithmic tangent
// (rebuilt on the first two iterations, then kept while the residual keeps
// shrinking), elastic parts incremental with element matrices cached per step,
// solder points through the Anand return map, and an IC(0)-preconditioned CG
// on the Jacobi-scaled tangent with the inexact-Newton tolerance.
//
// Device layout: the tangent in the upper block pattern of each node (the
// node and its 13 grid neighbours of larger index, 14 slots of 3x3, the
// diagonal block full), the same layout as the IC(0) factor; element
// matrices 24x24 per element; Anand states 11 per Gauss point (8 per element).
// Element kernels: one wavefront per element (Gauss points on lanes 0..7,
// stiffness entries spread over the lanes).  The inner CG runs in one
// workgroup (level-scheduled IC(0) sweeps separated by workgroup barriers),
// one launch per solve.

constexpr int SUB_NMAT = 4, SUB_SOLDER = 1, SUB_WG = 1024, SUB_EW = 4;

struct fcbga_submodel {
  fe::Grid g{};
  int64_t ne = 0, ndof = 0;
  std::vector<double> xs, ys, zs;
  fe::AnandParams P{};
  be::stream_t stream = nullptr;
  DBuf<double> dxs, dys, dzs;
  DBuf<int8_t> cellMat;
  DBuf<uint8_t> cellLayer, isB;
  int32_t nb = 0;
  DBuf<int32_t> bnodes;
  DBuf<double> ub;
  // displacement, committed displacement, internal force, Newton increment
  DBuf<double> u, uN, rint, du;
  // element data: tangent (576), cached elastic thermal load (24), committed internal force (24),
  // committed elastic Gauss-point stresses (48), current internal force (24), largest plastic increment
  DBuf<double> eK, eF, eFn, sigE, fe, dmaxE;
  // Anand states: committed and trial (11 per Gauss point)
  DBuf<double> st, tr;
  DBuf<double> elasD, elasSth;
  // assembled tangent (constrained), scaled copy, IC(0) factor (126 per node)
  DBuf<double> K, Ks, U, sv;
  DBuf<uint16_t> mask;
  DBuf<int32_t> nodes, fail, itOut;
  DBuf<int64_t> lvPtr;
  int64_t nl = 0;
  // inner CG vectors
  DBuf<double> b, x, r, z, p, q;
  DBuf<double> red;
  bool pcBuilt = false, pcStale = false;
  int pcIters = 0;
  double shift = 0;
  double Tn = 25, keyTn = NAN, keyT = NAN;
  // snapshot for cutbacks
  DBuf<double> snapU, snapSigE, snapFn, snapSt;
  double snapTn = 25;
  bool haveSnap = false;
  double *hostRed = nullptr;
};

namespace {

struct SubView {
  fe::Grid g;
  const double *xs, *ys, *zs;
  const int8_t *cellMat;
};

FE_HD void sub_cell(const fe::Grid &g, int64_t e, int &i, int &j, int &k) {
  k = (int)(e % g.ncz);
  const int64_t c = e / g.ncz;
  j = (int)(c % g.ncy);
  i = (int)(c / g.ncy);
}

FE_HD int64_t sub_node(const fe::Grid &g, int i, int j, int k) { return ((int64_t)i * g.ny + j) * g.nz + k; }

#if FCBGA_GPU
// Element pass: internal force (and, when asked, the tangent) of one element per
// wavefront.  Solder: Anand trial states from the committed ones at the
// mechanical strain B u - F_solder; elastic: incremental from the committed
// state, element matrix and thermal load rebuilt when the step is fresh.
__global__ void __launch_bounds__(64 * SUB_EW) k_sub_elements(SubView v, fe::AnandParams P, int64_t ne, double T,
                                                             double dt, double E, double fsol, int wantK, int fresh,
                                                             const double *__restrict__ u, const double *__restrict__ uN,
                                                             const double *__restrict__ elasD,
                                                             const double *__restrict__ elasSth, double *st,
                                                             double *tr, double *eK, double *eF,
                                                             const double *__restrict__ eFn, double *fe,
                                                             double *dmaxE) {
  __shared__ double Bs[SUB_EW][8 * 144];
  __shared__ double Ds[SUB_EW][8 * 36];
  __shared__ double Ss[SUB_EW][8 * 6];
  __shared__ double Us[SUB_EW][24];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t e = (int64_t)blockIdx.x * SUB_EW + wv;
  const bool valid = e < ne;
  const fe::Grid g = v.g;
  int i = 0, j = 0, k = 0;
  if (valid) sub_cell(g, e, i, j, k);
  const int mat = valid ? v.cellMat[e] : 0;
  const bool anand = mat == SUB_SOLDER;
  const double hx = valid ? v.xs[i + 1] - v.xs[i] : 1, hy = valid ? v.ys[j + 1] - v.ys[j] : 1,
               hz = valid ? v.zs[k + 1] - v.zs[k] : 1;
  const double w = hx * hy * hz / 8;
  double *B = Bs[wv], *Dl = Ds[wv], *Sg = Ss[wv], *ue = Us[wv];
  double d = 0;
  if (valid) {
    if (lane < 24) {
      const int a = lane / 3, c = lane % 3;
      const int64_t n = sub_node(g, i + fe::corner_li(a), j + fe::corner_lj(a), k + fe::corner_lk(a));
      ue[lane] = anand ? u[3 * n + c] : u[3 * n + c] - uN[3 * n + c];
    }
    if (lane < 8) {
      const double gp = 0x1.279a74590331dp-1;  // 1 / sqrt(3)
      fe::bbar_matrix(hx, hy, hz, (2 * fe::corner_li(lane) - 1) * gp, (2 * fe::corner_lj(lane) - 1) * gp,
                      (2 * fe::corner_lk(lane) - 1) * gp, B + 144 * lane);
    }
  }
  __syncthreads();
  if (valid && lane < 8) {
    if (anand) {
      const double *Bg = B + 144 * lane;
      double eps[6], sig[6];
      for (int r = 0; r < 6; r++) {
        double s = 0;
        for (int c = 0; c < 24; c++) s += Bg[r * 24 + c] * ue[c];
        eps[r] = s;
      }
      eps[0] -= fsol; eps[1] -= fsol; eps[2] -= fsol;
      const int64_t gq = 8 * e + lane;
      d = fe::anand_step(P, E, T, dt, eps, st + 11 * gq, tr + 11 * gq, sig,
                         wantK ? Dl + 36 * lane : nullptr);
      for (int r = 0; r < 6; r++) Sg[6 * lane + r] = sig[r];
    } else if (fresh) {
      for (int q = 0; q < 36; q++) Dl[36 * lane + q] = elasD[36 * mat + q];
    }
  }
  __syncthreads();
  double *Ke = eK + 576 * e;
  const bool buildK = valid && ((anand && wantK) || (!anand && fresh));
  if (buildK) {
    // K_e[a][b] (b >= a) = sum over points of w B^T D B, then the mirror
    for (int idx = lane; idx < 300; idx += 64) {
      int a = 0, rem = idx;
      while (rem >= 24 - a) { rem -= 24 - a; a++; }
      const int bcol = a + rem;
      double acc = 0;
      for (int gq = 0; gq < 8; gq++) {
        const double *Bg = B + 144 * gq, *Dg = Dl + 36 * gq;
        double s = 0;
        for (int r = 0; r < 6; r++) {
          double db = 0;
          for (int t = 0; t < 6; t++) db += Dg[r * 6 + t] * Bg[t * 24 + bcol];
          s += Bg[r * 24 + a] * db;
        }
        acc += w * s;
      }
      Ke[a * 24 + bcol] = acc;
      Ke[bcol * 24 + a] = acc;
    }
    if (!anand && lane < 24) {
      double f = 0;
      for (int gq = 0; gq < 8; gq++) {
        const double *Bg = B + 144 * gq;
        double s = 0;
        for (int r = 0; r < 6; r++) s += Bg[r * 24 + lane] * elasSth[6 * mat + r];
        f += w * s;
      }
      eF[24 * e + lane] = f;
    }
  }
  __syncthreads();
  if (valid && lane < 24) {
    double f;
    if (anand) {
      f = 0;
      for (int gq = 0; gq < 8; gq++) {
        const double *Bg = B + 144 * gq;
        double s = 0;
        for (int r = 0; r < 6; r++) s += Bg[r * 24 + lane] * Sg[6 * gq + r];
        f += w * s;
      }
    } else {
      f = eFn[24 * e + lane] - eF[24 * e + lane];
      for (int c = 0; c < 24; c++) f += Ke[lane * 24 + c] * ue[c];
    }
    fe[24 * e + lane] = f;
  }
  // largest plastic increment of the element's points
  double m = d;
  for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if (valid && lane == 0) dmaxE[e] = m;
}
#endif

#if !FCBGA_GPU
// host builds: one element at a time, the reference's loop order
void sub_element_host(const SubView &v, const fe::AnandParams &P, int64_t e, double T, double dt, double E,
                      double fsol, int wantK, int fresh, const double *u, const double *uN, const double *elasD,
                      const double *elasSth, double *st, double *tr, double *eK, double *eF, const double *eFn,
                      double *fe, double *dmaxE) {
  const fe::Grid &g = v.g;
  int i, j, k;
  sub_cell(g, e, i, j, k);
  const int mat = v.cellMat[e];
  const bool anand = mat == SUB_SOLDER;
  const double hx = v.xs[i + 1] - v.xs[i], hy = v.ys[j + 1] - v.ys[j], hz = v.zs[k + 1] - v.zs[k];
  const double w = hx * hy * hz / 8;
  double B[8 * 144], ue[24], Dg[8 * 36], Sg[8 * 6];
  for (int a = 0; a < 8; a++) {
    const int64_t n = sub_node(g, i + fe::corner_li(a), j + fe::corner_lj(a), k + fe::corner_lk(a));
    for (int c = 0; c < 3; c++) ue[3 * a + c] = anand ? u[3 * n + c] : u[3 * n + c] - uN[3 * n + c];
  }
  const double gp = 0x1.279a74590331dp-1;
  for (int gq = 0; gq < 8; gq++)
    fe::bbar_matrix(hx, hy, hz, (2 * fe::corner_li(gq) - 1) * gp, (2 * fe::corner_lj(gq) - 1) * gp,
                    (2 * fe::corner_lk(gq) - 1) * gp, B + 144 * gq);
  double dm = 0;
  for (int gq = 0; gq < 8; gq++) {
    if (anand) {
      double eps[6], sig[6];
      for (int r = 0; r < 6; r++) {
        double s2 = 0;
        for (int c = 0; c < 24; c++) s2 += B[144 * gq + r * 24 + c] * ue[c];
        eps[r] = s2;
      }
      eps[0] -= fsol; eps[1] -= fsol; eps[2] -= fsol;
      const int64_t q = 8 * e + gq;
      const double d = fe::anand_step(P, E, T, dt, eps, st + 11 * q, tr + 11 * q, sig, wantK ? Dg + 36 * gq : nullptr);
      if (d > dm) dm = d;
      for (int r = 0; r < 6; r++) Sg[6 * gq + r] = sig[r];
    } else {
      for (int q2 = 0; q2 < 36; q2++) Dg[36 * gq + q2] = elasD[36 * mat + q2];
    }
  }
  double *Ke = eK + 576 * e;
  if ((anand && wantK) || (!anand && fresh)) {
    for (int a = 0; a < 24; a++)
      for (int bc = a; bc < 24; bc++) {
        double acc = 0;
        for (int gq = 0; gq < 8; gq++) {
          const double *Bg = B + 144 * gq, *D = Dg + 36 * gq;
          double s2 = 0;
          for (int r = 0; r < 6; r++) {
            double db = 0;
            for (int t = 0; t < 6; t++) db += D[r * 6 + t] * Bg[t * 24 + bc];
            s2 += Bg[r * 24 + a] * db;
          }
          acc += w * s2;
        }
        Ke[a * 24 + bc] = acc;
        Ke[bc * 24 + a] = acc;
      }
    if (!anand)
      for (int c = 0; c < 24; c++) {
        double f = 0;
        for (int gq = 0; gq < 8; gq++) {
          double s2 = 0;
          for (int r = 0; r < 6; r++) s2 += B[144 * gq + r * 24 + c] * elasSth[6 * mat + r];
          f += w * s2;
        }
        eF[24 * e + c] = f;
      }
  }
  for (int c = 0; c < 24; c++) {
    double f;
    if (anand) {
      f = 0;
      for (int gq = 0; gq < 8; gq++) {
        double s2 = 0;
        for (int r = 0; r < 6; r++) s2 += B[144 * gq + r * 24 + c] * Sg[6 * gq + r];
        f += w * s2;
      }
    } else {
      f = eFn[24 * e + c] - eF[24 * e + c];
      for (int b = 0; b < 24; b++) f += Ke[c * 24 + b] * ue[b];
    }
    fe[24 * e + c] = f;
  }
  dmaxE[e] = dm;
}
#endif

// node pass: internal force gathered from the elements around the node, and
// (wantK) the tangent's upper blocks from the elements holding both nodes, in
// ascending element order; rows and columns of the boundary DOFs zeroed but
// the diagonal (_constrain, src/core/13-fatigue.js:868-877)
FE_HD void sub_node_gather(const fe::Grid &g, int64_t gi, const uint8_t *isB, const double *fe,
                                  const double *eK, int wantK, double *rint, double *K) {
  const int k = (int)(gi % g.nz);
  const int64_t cj = gi / g.nz;
  const int j = (int)(cj % g.ny), i = (int)(cj / g.ny);
  double f[3] = {0, 0, 0};
  for (int ci = i - 1; ci <= i; ci++) {
    if (ci < 0 || ci >= g.ncx) continue;
    for (int cjj = j - 1; cjj <= j; cjj++) {
      if (cjj < 0 || cjj >= g.ncy) continue;
      for (int ck = k - 1; ck <= k; ck++) {
        if (ck < 0 || ck >= g.ncz) continue;
        const int64_t e = ((int64_t)ci * g.ncy + cjj) * g.ncz + ck;
        const int a = fe::corner_of(i - ci, j - cjj, k - ck);
        for (int p = 0; p < 3; p++) f[p] += fe[24 * e + 3 * a + p];
      }
    }
  }
  for (int p = 0; p < 3; p++) rint[3 * gi + p] = f[p];
  if (!wantK) return;
  const bool bi = isB[gi];
  double *Kn = K + gi * IC_STRIDE;
  for (int t = 0; t < IC_SLOTS; t++) {
    int di, dj, dk;
    ic_off(t, di, dj, dk);
    const int ni = i + di, nj = j + dj, nk = k + dk;
    double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (ni < g.nx && nj >= 0 && nj < g.ny && nk >= 0 && nk < g.nz) {
      const int64_t gn = sub_node(g, ni, nj, nk);
      const int cx0 = (i > ni ? i : ni) - 1, cx1 = i < ni ? i : ni;
      const int cy0 = (j > nj ? j : nj) - 1, cy1 = j < nj ? j : nj;
      const int cz0 = (k > nk ? k : nk) - 1, cz1 = k < nk ? k : nk;
      for (int ci = cx0; ci <= cx1; ci++) {
        if (ci < 0 || ci >= g.ncx) continue;
        for (int cjj = cy0; cjj <= cy1; cjj++) {
          if (cjj < 0 || cjj >= g.ncy) continue;
          for (int ck = cz0; ck <= cz1; ck++) {
            if (ck < 0 || ck >= g.ncz) continue;
            const int64_t e = ((int64_t)ci * g.ncy + cjj) * g.ncz + ck;
            const int a = fe::corner_of(i - ci, j - cjj, k - ck), b = fe::corner_of(ni - ci, nj - cjj, nk - ck);
            const double *Ke = eK + 576 * e;
            for (int p = 0; p < 3; p++)
              for (int q = 0; q < 3; q++) blk[3 * p + q] += Ke[(3 * a + p) * 24 + 3 * b + q];
          }
        }
      }
      const bool bn = isB[gn];
      for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++)
          if ((bi || bn) && !(t == 0 && p == q)) blk[3 * p + q] = 0;
    }
    for (int q = 0; q < 9; q++) Kn[t * 9 + q] = blk[q];
  }
}

// y = K x with K in upper blocks: own slots, then the transposed blocks of the lower neighbours
FE_HD void sub_matvec_node(const fe::Grid &g, int64_t gi, const uint16_t *mask, const double *K,
                                  const double *x, double *y) {
  const int k = (int)(gi % g.nz);
  const int64_t cj = gi / g.nz;
  const int j = (int)(cj % g.ny), i = (int)(cj / g.ny);
  const uint32_t mk = mask[gi];
  double acc[3] = {0, 0, 0};
  for (int t = 0; t < IC_SLOTS; t++) {
    if (!((mk >> t) & 1u)) continue;
    int di, dj, dk;
    ic_off(t, di, dj, dk);
    const int64_t gn = sub_node(g, i + di, j + dj, k + dk);
    const double *Bk = K + gi * IC_STRIDE + t * 9;
    for (int p = 0; p < 3; p++) acc[p] += Bk[3 * p] * x[3 * gn] + Bk[3 * p + 1] * x[3 * gn + 1] + Bk[3 * p + 2] * x[3 * gn + 2];
  }
  for (int tl = 1; tl < IC_SLOTS; tl++) {
    int ei, ej, ek;
    ic_off(tl, ei, ej, ek);
    const int mi = i - ei, mj = j - ej, mk2 = k - ek;
    if (mi < 0 || mj < 0 || mj >= g.ny || mk2 < 0 || mk2 >= g.nz) continue;
    const int64_t gm = sub_node(g, mi, mj, mk2);
    if (!((mask[gm] >> tl) & 1u)) continue;
    const double *Bk = K + gm * IC_STRIDE + tl * 9;
    for (int p = 0; p < 3; p++) acc[p] += Bk[p] * x[3 * gm] + Bk[3 + p] * x[3 * gm + 1] + Bk[6 + p] * x[3 * gm + 2];
  }
  for (int p = 0; p < 3; p++) y[3 * gi + p] = acc[p];
}

#if FCBGA_GPU
__device__ inline double wg_sum(double v, double *red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int w = SUB_WG / 2; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  const double s = red[0];
  __syncthreads();
  return s;
}

__device__ inline double wg_dot(const double *a, const double *b, int64_t n, double *red) {
  double s = 0;
  for (int64_t q = threadIdx.x; q < n; q += SUB_WG) s += a[q] * b[q];
  return wg_sum(s, red);
}

struct SubLevels {
  fe::Grid g;
  const int32_t *nodes;
  const int64_t *lvPtr;
  int64_t nl;
  const uint16_t *mask;
};

// z = (U^T U)^-1 r inside one workgroup: one wavefront per node, a barrier per level
__device__ inline void wg_ic_apply(const SubLevels &L, const double *U, const double *r, double *z) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = SUB_WG / 64;
  for (int64_t l = 0; l < L.nl; l++) {
    for (int64_t w = L.lvPtr[l] + wv; w < L.lvPtr[l + 1]; w += nw) ic_forward_node(L.g, L.nodes[w], L.mask, U, r, z, lane);
    __syncthreads();
  }
  for (int64_t l = L.nl - 1; l >= 0; l--) {
    for (int64_t w = L.lvPtr[l] + wv; w < L.lvPtr[l + 1]; w += nw) ic_backward_node(L.g, L.nodes[w], L.mask, U, z, lane);
    __syncthreads();
  }
}

__global__ void __launch_bounds__(SUB_WG) k_sub_ic_factor(SubLevels L, double *U, int32_t *fail) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = SUB_WG / 64;
  for (int64_t l = 0; l < L.nl; l++) {
    for (int64_t w = L.lvPtr[l] + wv; w < L.lvPtr[l + 1]; w += nw) ic_factor_node(L.g, L.nodes[w], L.mask, U, fail, lane);
    __syncthreads();
  }
}

// IC(0)-preconditioned CG from x = 0 on the scaled tangent (BumpSubmodel._solve,
// src/core/13-fatigue.js:880-917): stops when |r| / |b| < tol or after maxit
// iterations; itOut = the reference's iteration count (maxit + 1 when it ran out)
__global__ void __launch_bounds__(SUB_WG) k_sub_pcg(SubLevels L, const double *__restrict__ Ks,
                                                    const double *__restrict__ U, const double *__restrict__ bv,
                                                    double *x, double *r, double *z, double *p, double *q,
                                                    double tol, int maxit, int32_t *itOut) {
  __shared__ double red[SUB_WG];
  const int tid = threadIdx.x;
  const int64_t n3 = 3 * L.g.G;
  for (int64_t i = tid; i < n3; i += SUB_WG) { r[i] = bv[i]; x[i] = 0; }
  __syncthreads();
  const double bn = sqrt(wg_dot(bv, bv, n3, red));
  if (bn == 0) {
    if (tid == 0) itOut[0] = 0;
    return;
  }
  wg_ic_apply(L, U, r, z);
  for (int64_t i = tid; i < n3; i += SUB_WG) p[i] = z[i];
  __syncthreads();
  double rz = wg_dot(r, z, n3, red);
  int it;
  for (it = 1; it <= maxit; it++) {
    for (int64_t gi = tid; gi < L.g.G; gi += SUB_WG) sub_matvec_node(L.g, gi, L.mask, Ks, p, q);
    __syncthreads();
    const double alpha = rz / wg_dot(p, q, n3, red);
    for (int64_t i = tid; i < n3; i += SUB_WG) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
    __syncthreads();
    if (sqrt(wg_dot(r, r, n3, red)) / bn < tol) break;
    wg_ic_apply(L, U, r, z);
    const double rz2 = wg_dot(r, z, n3, red);
    const double beta = rz2 / rz;
    rz = rz2;
    for (int64_t i = tid; i < n3; i += SUB_WG) p[i] = z[i] + beta * p[i];
    __syncthreads();
  }
  if (tid == 0) itOut[0] = it;
}

// out[0] = sum over free DOFs of a^2, out[1] = sum of a^2 (deterministic order)
__global__ void __launch_bounds__(SUB_WG) k_sub_norms(const double *__restrict__ a, const uint8_t *__restrict__ isB,
                                                      int64_t n3, double *out) {
  __shared__ double red[SUB_WG];
  double s0 = 0, s1 = 0;
  for (int64_t i = threadIdx.x; i < n3; i += SUB_WG) {
    const double v = a[i] * a[i];
    s1 += v;
    if (!isB[i / 3]) s0 += v;
  }
  const double r0 = wg_sum(s0, red), r1 = wg_sum(s1, red);
  if (threadIdx.x == 0) { out[0] = r0; out[1] = r1; }
}

__global__ void __launch_bounds__(SUB_WG) k_sub_max(const double *__restrict__ a, int64_t n, double *out) {
  __shared__ double red[SUB_WG];
  double m = 0;
  for (int64_t i = threadIdx.x; i < n; i += SUB_WG) m = fmax(m, a[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = SUB_WG / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}
#endif

SubView sub_view(const fcbga_submodel *S) { return SubView{S->g, S->dxs.p, S->dys.p, S->dzs.p, S->cellMat.p}; }

void sub_create(const fcbga_submodel_mesh *m, const fcbga_anand_params *sp, fcbga_submodel *S) {
  if (!m || !sp) throw be::Error("submodel mesh or solder parameters missing");
  if (m->nx < 2 || m->ny < 2 || m->nz < 2) throw be::Error("submodel needs at least 2 grid lines per axis");
  if (!m->xs || !m->ys || !m->zs || !m->cell_mat || !m->cell_layer) throw be::Error("submodel arrays missing");
  if (!(sp->nu < 0.5)) throw be::Error("nu < 0.5 required");
  fe::Grid &g = S->g;
  g.nx = m->nx; g.ny = m->ny; g.nz = m->nz;
  g.ncx = g.nx - 1; g.ncy = g.ny - 1; g.ncz = g.nz - 1;
  g.G = (int64_t)g.nx * g.ny * g.nz;
  S->ne = (int64_t)g.ncx * g.ncy * g.ncz;
  S->ndof = 3 * g.G;
  S->xs.assign(m->xs, m->xs + g.nx);
  S->ys.assign(m->ys, m->ys + g.ny);
  S->zs.assign(m->zs, m->zs + g.nz);
  for (int64_t e = 0; e < S->ne; e++)
    if (m->cell_mat[e] < 0 || m->cell_mat[e] >= SUB_NMAT) throw be::Error("submodel cell material out of range");
  S->P = fe::AnandParams{sp->A, sp->QR, sp->xi, sp->m, sp->shat, sp->n, sp->h0, sp->a, sp->s0, sp->nu};
#if FCBGA_GPU
  BE_CHECK(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
#endif
  be::stream_t s = S->stream;
  S->dxs.upload(S->xs, s); S->dys.upload(S->ys, s); S->dzs.upload(S->zs, s);
  S->cellMat.upload(m->cell_mat, S->ne, s);
  S->cellLayer.upload(m->cell_layer, S->ne, s);
  // boundary nodes (the six faces) in grid order; upper pattern: every in-range neighbour
  std::vector<uint8_t> isB(g.G, 0);
  std::vector<int32_t> bn;
  std::vector<uint16_t> mask(g.G, 0);
  for (int i = 0; i < g.nx; i++)
    for (int j = 0; j < g.ny; j++)
      for (int k = 0; k < g.nz; k++) {
        const int64_t gi = sub_node(g, i, j, k);
        if (i == 0 || j == 0 || k == 0 || i == g.nx - 1 || j == g.ny - 1 || k == g.nz - 1) {
          isB[gi] = 1;
          bn.push_back((int32_t)gi);
        }
        uint16_t mk = 1;
        for (int t = 1; t < IC_SLOTS; t++) {
          int di, dj, dk;
          ic_off(t, di, dj, dk);
          if (i + di < g.nx && j + dj >= 0 && j + dj < g.ny && k + dk >= 0 && k + dk < g.nz) mk |= (uint16_t)(1u << t);
        }
        mask[gi] = mk;
      }
  S->isB.upload(isB, s);
  S->nb = (int32_t)bn.size();
  S->bnodes.upload(bn, s);
  S->ub.alloc(3 * (size_t)S->nb);
  std::vector<int32_t> nodes;
  std::vector<int64_t> ptr;
  ic_levels(g, mask, nodes, ptr);
  S->mask.upload(mask, s);
  S->nodes.upload(nodes, s);
  S->lvPtr.upload(ptr, s);
  S->nl = (int64_t)ptr.size() - 1;
  S->fail.alloc(1);
  S->itOut.alloc(1);
  const size_t n3 = (size_t)S->ndof, ne = (size_t)S->ne;
  for (DBuf<double> *v : {&S->u, &S->uN, &S->rint, &S->du, &S->sv, &S->b, &S->x, &S->r, &S->z, &S->p, &S->q}) {
    v->alloc(n3);
    v->zero(s);
  }
  S->eK.alloc(576 * ne); S->eK.zero(s);
  S->eF.alloc(24 * ne); S->eF.zero(s);
  S->eFn.alloc(24 * ne); S->eFn.zero(s);
  S->sigE.alloc(48 * ne); S->sigE.zero(s);
  S->fe.alloc(24 * ne);
  S->dmaxE.alloc(ne);
  S->st.alloc(88 * ne);
  S->tr.alloc(88 * ne);
  {
    double *st = S->st.p;
    const double s0 = S->P.s0;
    be::pfor(8 * (int64_t)ne, s, [=] __host__ __device__(int64_t q) {
      for (int c = 0; c < 11; c++) st[11 * q + c] = c == 6 ? s0 : 0.0;  // AnandPoint.reset
    });
  }
  S->tr.zero(s);
  S->elasD.alloc(SUB_NMAT * 36);
  S->elasSth.alloc(SUB_NMAT * 6);
  S->K.alloc(IC_STRIDE * (size_t)g.G); S->K.zero(s);
  S->Ks.alloc(IC_STRIDE * (size_t)g.G);
  S->U.alloc(IC_STRIDE * (size_t)g.G);
  S->red.alloc(8);
  be::sync(s);
}

// Internal force at u, and the tangent when wantK (BumpSubmodel.assemble, src/core/13-fatigue.js:
// 775-851); returns the largest equivalent plastic increment over the solder points
double sub_assemble(fcbga_submodel *S, double T, double dt, double E, double fsol, bool wantK) {
  be::stream_t s = S->stream;
  const bool fresh = !(S->keyTn == S->Tn && S->keyT == T);
  const SubView v = sub_view(S);
  const fe::AnandParams P = S->P;
  const int64_t ne = S->ne;
#if FCBGA_GPU
  hipLaunchKernelGGL(k_sub_elements, dim3((unsigned)((ne + SUB_EW - 1) / SUB_EW)), dim3(64 * SUB_EW), 0, s, v, P, ne,
                     T, dt, E, fsol, wantK ? 1 : 0, fresh ? 1 : 0, S->u.p, S->uN.p, S->elasD.p, S->elasSth.p, S->st.p,
                     S->tr.p, S->eK.p, S->eF.p, S->eFn.p, S->fe.p, S->dmaxE.p);
  BE_CHECK(hipGetLastError());
#else
  for (int64_t e = 0; e < ne; e++)
    sub_element_host(v, P, e, T, dt, E, fsol, wantK, fresh, S->u.p, S->uN.p, S->elasD.p, S->elasSth.p, S->st.p,
                     S->tr.p, S->eK.p, S->eF.p, S->eFn.p, S->fe.p, S->dmaxE.p);
#endif
  S->keyTn = S->Tn;
  S->keyT = T;
  const fe::Grid g = S->g;
  const uint8_t *isB = S->isB.p;
  const double *fe = S->fe.p, *eK = S->eK.p;
  double *rint = S->rint.p, *K = S->K.p;
  const int wk = wantK ? 1 : 0;
  be::pfor(g.G, s, [=] __host__ __device__(int64_t gi) { sub_node_gather(g, gi, isB, fe, eK, wk, rint, K); });
  double dmax = 0;
#if FCBGA_GPU
  hipLaunchKernelGGL(k_sub_max, dim3(1), dim3(SUB_WG), 0, s, S->dmaxE.p, ne, S->red.p);
  BE_CHECK(hipGetLastError());
  be::d2h(&dmax, S->red.p, sizeof(double), s);
#else
  for (int64_t e = 0; e < ne; e++) dmax = std::max(dmax, S->dmaxE.p[e]);
#endif
  return dmax;
}

// {sum over free DOFs of a^2, sum of a^2}
void sub_norms(fcbga_submodel *S, const double *a, double out[2]) {
#if FCBGA_GPU
  hipLaunchKernelGGL(k_sub_norms, dim3(1), dim3(SUB_WG), 0, S->stream, a, S->isB.p, S->ndof, S->red.p);
  BE_CHECK(hipGetLastError());
  be::d2h(out, S->red.p, 2 * sizeof(double), S->stream);
#else
  out[0] = out[1] = 0;
  for (int64_t i = 0; i < S->ndof; i++) {
    out[1] += a[i] * a[i];
    if (!S->isB.p[i / 3]) out[0] += a[i] * a[i];
  }
#endif
}

// du = K^-1 rhs by IC(0)-PCG on the Jacobi-scaled tangent (BumpSubmodel._solve,
// src/core/13-fatigue.js:880-917); rhs in S->b (unscaled) on entry; returns the
// iteration count.  The factor is kept until a solve takes more than twice the
// iterations of the first solve after its build.
int sub_solve(fcbga_submodel *S, double tol, int32_t &icBuilds) {
  be::stream_t s = S->stream;
  const fe::Grid g = S->g;
  const double *K = S->K.p;
  double *sv = S->sv.p, *Ks = S->Ks.p, *bv = S->b.p;
  const uint16_t *mask = S->mask.p;
  be::pfor(g.G, s, [=] __host__ __device__(int64_t gi) {
    const double *D = K + gi * IC_STRIDE;
    for (int p = 0; p < 3; p++) sv[3 * gi + p] = 1 / sqrt(fmax(D[4 * p], 1e-300));
  });
  be::pfor(g.G, s, [=] __host__ __device__(int64_t gi) {
    const int k = (int)(gi % g.nz);
    const int64_t cj = gi / g.nz;
    const int j = (int)(cj % g.ny), i = (int)(cj / g.ny);
    const uint32_t mk = mask[gi];
    for (int t = 0; t < IC_SLOTS; t++) {
      if (!((mk >> t) & 1u)) continue;
      int di, dj, dk;
      ic_off(t, di, dj, dk);
      const int64_t gn = sub_node(g, i + di, j + dj, k + dk);
      for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++)
          Ks[gi * IC_STRIDE + t * 9 + 3 * p + q] = K[gi * IC_STRIDE + t * 9 + 3 * p + q] * (sv[3 * gi + p] * sv[3 * gn + q]);
    }
    for (int p = 0; p < 3; p++) bv[3 * gi + p] *= sv[3 * gi + p];
  });
  if (!S->pcBuilt || S->pcStale) {
    // ICPC.build: shifts 0, 1e-3, x10 up to 1 (src/core/10-solver.js:748-760)
    double shift = 0;
    bool ok = false;
    for (int attempt = 0; attempt < 8 && !ok; attempt++) {
      double *U = S->U.p;
      const double sh = shift;
      be::pfor(g.G, s, [=] __host__ __device__(int64_t gi) {
        const uint32_t mk = mask[gi];
        for (int t = 0; t < IC_SLOTS; t++) {
          if (!((mk >> t) & 1u)) continue;
          for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++) {
              double v = (t == 0 && q < p) ? 0.0 : Ks[gi * IC_STRIDE + t * 9 + 3 * p + q];
              if (t == 0 && p == q && sh > 0) v *= 1 + sh;
              U[gi * IC_STRIDE + t * 9 + 3 * p + q] = v;
            }
        }
      });
#if FCBGA_GPU
      int32_t *fl = S->fail.p;
      be::pfor(1, s, [=] __host__ __device__(int64_t) { fl[0] = 0; });
      SubLevels L{g, S->nodes.p, S->lvPtr.p, S->nl, mask};
      hipLaunchKernelGGL(k_sub_ic_factor, dim3(1), dim3(SUB_WG), 0, s, L, U, fl);
      BE_CHECK(hipGetLastError());
      int32_t h = 0;
      be::d2h(&h, fl, sizeof(int32_t), s);
      ok = h == 0;
#else
      ok = ic_factor_host(g, S->nodes.p, S->lvPtr.p[S->nl], mask, U);
#endif
      if (!ok) {
        shift = shift == 0 ? 1e-3 : shift * 10;
        if (shift > 1) shift = 1;
      }
    }
    if (!ok) throw be::Error("IC(0) failed even with the maximum diagonal shift");
    S->shift = shift;
    S->pcBuilt = true;
    S->pcStale = false;
    S->pcIters = 0;
    icBuilds++;
  }
  int it = 0;
  const int64_t n3 = S->ndof;
#if FCBGA_GPU
  SubLevels L{g, S->nodes.p, S->lvPtr.p, S->nl, mask};
  hipLaunchKernelGGL(k_sub_pcg, dim3(1), dim3(SUB_WG), 0, s, L, (const double *)Ks, (const double *)S->U.p,
                     (const double *)bv, S->x.p, S->r.p, S->z.p, S->p.p, S->q.p, tol, 2000, S->itOut.p);
  BE_CHECK(hipGetLastError());
  int32_t hit = 0;
  be::d2h(&hit, S->itOut.p, sizeof(int32_t), s);
  it = hit;
#else
  {
    double *x = S->x.p, *r = S->r.p, *z = S->z.p, *p = S->p.p, *q = S->q.p;
    const double *U = S->U.p;
    auto dot = [&](const double *a, const double *c) {
      double v = 0;
      for (int64_t i = 0; i < n3; i++) v += a[i] * c[i];
      return v;
    };
    for (int64_t i = 0; i < n3; i++) { r[i] = bv[i]; x[i] = 0; }
    const double bn = std::sqrt(dot(bv, bv));
    if (bn != 0) {
      ic_apply_host(g, S->nodes.p, S->lvPtr.p[S->nl], mask, U, r, z);
      for (int64_t i = 0; i < n3; i++) p[i] = z[i];
      double rz = dot(r, z);
      for (it = 1; it <= 2000; it++) {
        for (int64_t gi = 0; gi < g.G; gi++) sub_matvec_node(g, gi, mask, Ks, p, q);
        const double alpha = rz / dot(p, q);
        for (int64_t i = 0; i < n3; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        if (std::sqrt(dot(r, r)) / bn < tol) break;
        ic_apply_host(g, S->nodes.p, S->lvPtr.p[S->nl], mask, U, r, z);
        const double rz2 = dot(r, z);
        const double beta = rz2 / rz;
        rz = rz2;
        for (int64_t i = 0; i < n3; i++) p[i] = z[i] + beta * p[i];
      }
    }
  }
#endif
  // du = s x
  {
    const double *x = S->x.p;
    double *du = S->du.p;
    be::pfor(n3, s, [=] __host__ __device__(int64_t i) { du[i] = x[i] * sv[i]; });
  }
  if (!S->pcIters) S->pcIters = it;
  else if (it > 2 * S->pcIters) S->pcStale = true;
  return it;
}

constexpr int SUB_NEWTON_MAXIT = 25;     // src/core/00-constants.js:116
constexpr double SUB_RES_TOL = 1e-6;     // :117
constexpr double SUB_DU_TOL = 1e-8;      // :118
constexpr double SUB_INNER_TOL = 1e-8;   // :119

// Commit at the converged displacement (stepTo's end and _commitElastic,
// src/core/13-fatigue.js:853-866, 955-959): solder trial states become the
// committed ones; elastic Gauss-point stresses and internal forces advance by
// the step's increment; u_n = u, T_n = T
void sub_commit(fcbga_submodel *S, double T) {
  be::stream_t s = S->stream;
  const SubView v = sub_view(S);
  const fe::Grid g = S->g;
  double *st = S->st.p, *sigE = S->sigE.p, *fn = S->eFn.p;
  const double *tr = S->tr.p, *u = S->u.p, *uN = S->uN.p, *eD = S->elasD.p, *eS = S->elasSth.p;
  be::pfor(S->ne, s, [=] __host__ __device__(int64_t e) {
    const int mat = v.cellMat[e];
    if (mat == SUB_SOLDER) {
      for (int gq = 0; gq < 8; gq++)
        for (int c = 0; c < 10; c++) st[11 * (8 * e + gq) + c] = tr[11 * (8 * e + gq) + c];
      return;
    }
    int i, j, k;
    sub_cell(g, e, i, j, k);
    const double hx = v.xs[i + 1] - v.xs[i], hy = v.ys[j + 1] - v.ys[j], hz = v.zs[k + 1] - v.zs[k];
    const double w = hx * hy * hz / 8;
    double due[24], B[144];
    for (int a = 0; a < 8; a++) {
      const int64_t n = sub_node(g, i + fe::corner_li(a), j + fe::corner_lj(a), k + fe::corner_lk(a));
      for (int c = 0; c < 3; c++) due[3 * a + c] = u[3 * n + c] - uN[3 * n + c];
    }
    double f[24];
    for (int c = 0; c < 24; c++) f[c] = 0;
    const double gp = 0x1.279a74590331dp-1;
    for (int gq = 0; gq < 8; gq++) {
      fe::bbar_matrix(hx, hy, hz, (2 * fe::corner_li(gq) - 1) * gp, (2 * fe::corner_lj(gq) - 1) * gp,
                      (2 * fe::corner_lk(gq) - 1) * gp, B);
      double de[6], sg[6];
      for (int r = 0; r < 6; r++) {
        double s2 = 0;
        for (int c = 0; c < 24; c++) s2 += B[r * 24 + c] * due[c];
        de[r] = s2;
      }
      double *so = sigE + 48 * e + 6 * gq;
      for (int r = 0; r < 6; r++) {
        double s2 = -eS[6 * mat + r];
        for (int c = 0; c < 6; c++) s2 += eD[36 * mat + 6 * r + c] * de[c];
        so[r] += s2;
        sg[r] = so[r];
      }
      for (int c = 0; c < 24; c++) {
        double s2 = 0;
        for (int r = 0; r < 6; r++) s2 += B[r * 24 + c] * sg[r];
        f[c] += w * s2;
      }
    }
    for (int c = 0; c < 24; c++) fn[24 * e + c] = f[c];
  });
  be::d2d(S->uN.p, S->u.p, sizeof(double) * S->ndof, s);
  S->Tn = T;
}

// One time step to T (BumpSubmodel.stepTo, src/core/13-fatigue.js:924-961)
void sub_step(fcbga_submodel *S, const fcbga_submodel_step_in &in, fcbga_submodel_step_out &out) {
  be::stream_t s = S->stream;
  const int64_t n3 = S->ndof;
  if (!(in.E > 0) || !(in.dt >= 0)) throw be::Error("E > 0 and dt >= 0 required");
  if (!in.elas_D || !in.elas_sth || !in.ub) throw be::Error("step data missing");
  // elastic data of the step from T_n to T (used when the element matrices are rebuilt)
  be::h2d(S->elasD.p, in.elas_D, sizeof(double) * SUB_NMAT * 36, s);
  be::h2d(S->elasSth.p, in.elas_sth, sizeof(double) * SUB_NMAT * 6, s);
  // prescribe the boundary
  be::h2d(S->ub.p, in.ub, sizeof(double) * 3 * S->nb, s);
  {
    const int32_t *bn = S->bnodes.p;
    const double *ub = S->ub.p;
    double *u = S->u.p;
    be::pfor(S->nb, s, [=] __host__ __device__(int64_t q) {
      for (int c = 0; c < 3; c++) u[3 * (int64_t)bn[q] + c] = ub[3 * q + c];
    });
  }
  const double T = in.T, dt = in.dt, E = in.E, fs = in.fsolder;
  double dMax = 0, rPrev = INFINITY;
  bool converged = false, haveK = false;
  int it;
  int32_t pcg = 0, builds = 0;
  const uint8_t *isB = S->isB.p;
  for (it = 1; it <= SUB_NEWTON_MAXIT; it++) {
    const bool wantK = it <= 2 || !haveK;
    dMax = sub_assemble(S, T, dt, E, fs, wantK);
    if (wantK) haveK = true;
    double nr[2];
    sub_norms(S, S->rint.p, nr);
    const double rn = std::sqrt(nr[0]), fn = std::sqrt(nr[1]);
    const double ref = std::max(fn, 1e-30);
    if (rn / ref < SUB_RES_TOL) { converged = true; break; }
    if (!wantK && rn > 0.5 * rPrev) sub_assemble(S, T, dt, E, fs, true);  // stalled with the frozen tangent
    rPrev = rn;
    {
      const double *ri = S->rint.p;
      double *bv = S->b.p;
      be::pfor(n3, s, [=] __host__ __device__(int64_t i) { bv[i] = isB[i / 3] ? 0.0 : -ri[i]; });
    }
    // inexact Newton: the inner solve beats the current residual ratio by a margin
    const double tol = std::max(SUB_INNER_TOL, std::min(1e-3, 0.01 * SUB_RES_TOL * ref / std::max(rn, 1e-300)));
    pcg += sub_solve(S, tol, builds);
    {
      const double *du = S->du.p;
      double *u = S->u.p;
      be::pfor(n3, s, [=] __host__ __device__(int64_t i) { if (!isB[i / 3]) u[i] += du[i]; });
    }
    double nd[2], nu[2];
    sub_norms(S, S->du.p, nd);
    sub_norms(S, S->u.p, nu);
    if (std::sqrt(nd[0]) < SUB_DU_TOL * std::max(std::sqrt(nu[1]), 1e-30)) { converged = true; it++; break; }
  }
  if (converged) {
    dMax = sub_assemble(S, T, dt, E, fs, false);  // trial states at the converged displacement
    sub_commit(S, T);
  }
  be::sync(s);
  out.converged = converged ? 1 : 0;
  out.newton = it;
  out.dmax = dMax;
  out.pcg_iters = pcg;
  out.ic_builds = builds;
}

void sub_save(fcbga_submodel *S) {
  be::stream_t s = S->stream;
  S->snapU.alloc(S->ndof); S->snapSigE.alloc(48 * S->ne); S->snapFn.alloc(24 * S->ne); S->snapSt.alloc(88 * S->ne);
  be::d2d(S->snapU.p, S->u.p, sizeof(double) * S->ndof, s);
  be::d2d(S->snapSigE.p, S->sigE.p, sizeof(double) * 48 * S->ne, s);
  be::d2d(S->snapFn.p, S->eFn.p, sizeof(double) * 24 * S->ne, s);
  be::d2d(S->snapSt.p, S->st.p, sizeof(double) * 88 * S->ne, s);
  S->snapTn = S->Tn;
  S->haveSnap = true;
  be::sync(s);
}

// restore (the solder points' Newton warm starts are kept, as AnandPoint.restore does)
void sub_restore(fcbga_submodel *S) {
  if (!S->haveSnap) throw be::Error("no saved submodel state");
  be::stream_t s = S->stream;
  be::d2d(S->u.p, S->snapU.p, sizeof(double) * S->ndof, s);
  be::d2d(S->uN.p, S->snapU.p, sizeof(double) * S->ndof, s);
  be::d2d(S->sigE.p, S->snapSigE.p, sizeof(double) * 48 * S->ne, s);
  be::d2d(S->eFn.p, S->snapFn.p, sizeof(double) * 24 * S->ne, s);
  double *st = S->st.p;
  const double *ss = S->snapSt.p;
  be::pfor(8 * S->ne, s, [=] __host__ __device__(int64_t q) {
    for (int c = 0; c < 10; c++) st[11 * q + c] = ss[11 * q + c];
  });
  S->Tn = S->snapTn;
  be::sync(s);
}

// per solder point of the averaging layers: weight, W, ebar (host reduction in
// element order); cellLayer bit 0 die side, bit 1 substrate side
void sub_layer_means(fcbga_submodel *S, double out[4]) {
  const int64_t ne = S->ne;
  std::vector<double> st(88 * ne);
  std::vector<int8_t> mat(ne);
  std::vector<uint8_t> lay(ne);
  be::d2h(st.data(), S->st.p, sizeof(double) * st.size(), S->stream);
  be::d2h(mat.data(), S->cellMat.p, ne, S->stream);
  be::d2h(lay.data(), S->cellLayer.p, ne, S->stream);
  double sW[2] = {0, 0}, sE[2] = {0, 0}, sw[2] = {0, 0};
  for (int64_t e = 0; e < ne; e++) {
    if (mat[e] != SUB_SOLDER) continue;
    int i, j, k;
    sub_cell(S->g, e, i, j, k);
    const double w = (S->xs[i + 1] - S->xs[i]) * (S->ys[j + 1] - S->ys[j]) * (S->zs[k + 1] - S->zs[k]) / 8;
    for (int L = 0; L < 2; L++) {
      if (!((lay[e] >> L) & 1)) continue;
      for (int gq = 0; gq < 8; gq++) {
        sW[L] += w * st[11 * (8 * e + gq) + 8];
        sE[L] += w * st[11 * (8 * e + gq) + 7];
        sw[L] += w;
      }
    }
  }
  for (int L = 0; L < 2; L++) {
    out[2 * L] = sw[L] ? sW[L] / sw[L] : NAN;
    out[2 * L + 1] = sw[L] ? sE[L] / sw[L] : NAN;
  }
}

// hysteresis sample of one element (sampleLoop, src/core/13-fatigue.js:1001-1019):
// for the solder points of the die-side (L = 0) and substrate-side (L = 1)
// layers, the stress at the committed plastic state (an Anand step of 1e-9 s,
// which moves the point's Newton warm start as the reference's does), the shear
// stress and strain along (cphi, sphi); o[3L..3L+2] = sums of w tau, w gamma, w
FE_HD void sub_loop_element(const SubView &v, const fe::AnandParams &P, int64_t e, double T, double E, double fsol,
                            double cphi, double sphi, const uint8_t *lay, const double *u, double *st, double o[6]) {
  for (int c = 0; c < 6; c++) o[c] = 0;
  if (v.cellMat[e] != SUB_SOLDER || !lay[e]) return;
  const fe::Grid &g = v.g;
  int i, j, k;
  sub_cell(g, e, i, j, k);
  const double hx = v.xs[i + 1] - v.xs[i], hy = v.ys[j + 1] - v.ys[j], hz = v.zs[k + 1] - v.zs[k];
  const double w = hx * hy * hz / 8;
  double ue[24], B[144];
  for (int a = 0; a < 8; a++) {
    const int64_t n = sub_node(g, i + fe::corner_li(a), j + fe::corner_lj(a), k + fe::corner_lk(a));
    for (int c = 0; c < 3; c++) ue[3 * a + c] = u[3 * n + c];
  }
  const double gp = 0x1.279a74590331dp-1;
  for (int L = 0; L < 2; L++) {
    if (!((lay[e] >> L) & 1)) continue;
    for (int gq = 0; gq < 8; gq++) {
      fe::bbar_matrix(hx, hy, hz, (2 * fe::corner_li(gq) - 1) * gp, (2 * fe::corner_lj(gq) - 1) * gp,
                      (2 * fe::corner_lk(gq) - 1) * gp, B);
      double eps[6], sig[6], trial[11];
      for (int r = 0; r < 6; r++) {
        double s2 = 0;
        for (int c = 0; c < 24; c++) s2 += B[r * 24 + c] * ue[c];
        eps[r] = s2;
      }
      eps[0] -= fsol; eps[1] -= fsol; eps[2] -= fsol;
      fe::anand_step(P, E, T, 1e-9, eps, st + 11 * (8 * e + gq), trial, sig, nullptr);
      o[3 * L] += w * (sig[5] * cphi + sig[4] * sphi);
      o[3 * L + 1] += w * (eps[5] * cphi + eps[4] * sphi);
      o[3 * L + 2] += w;
    }
  }
}

// stresses of one element at the committed state extrapolated to its corners
// (_stressSnapshot, src/core/13-fatigue.js:1036-1062): elastic parts the
// committed Gauss-point stresses, solder the Anand stress at the committed
// plastic strain (stressAt, :221-227) for the current strain; corner a gets
// sum over points of GP_TO_NODE[a][gq] sigma_gq (the trilinear functions of the
// points at sqrt(3) x the corner's natural coordinates)
FE_HD void sub_section_element(const SubView &v, const fe::AnandParams &P, int64_t e, double E, double fsol,
                               const double *u, const double *st, const double *sigE, float *out) {
  const fe::Grid &g = v.g;
  int i, j, k;
  sub_cell(g, e, i, j, k);
  double gpS[48];
  if (v.cellMat[e] != SUB_SOLDER) {
    for (int q = 0; q < 48; q++) gpS[q] = sigE[48 * e + q];
  } else {
    const double hx = v.xs[i + 1] - v.xs[i], hy = v.ys[j + 1] - v.ys[j], hz = v.zs[k + 1] - v.zs[k];
    double ue[24], B[144];
    for (int a = 0; a < 8; a++) {
      const int64_t n = sub_node(g, i + fe::corner_li(a), j + fe::corner_lj(a), k + fe::corner_lk(a));
      for (int c = 0; c < 3; c++) ue[3 * a + c] = u[3 * n + c];
    }
    const double G = E / (2 * (1 + P.nu)), K = E / (3 * (1 - 2 * P.nu));
    const double gp = 0x1.279a74590331dp-1;
    for (int gq = 0; gq < 8; gq++) {
      fe::bbar_matrix(hx, hy, hz, (2 * fe::corner_li(gq) - 1) * gp, (2 * fe::corner_lj(gq) - 1) * gp,
                      (2 * fe::corner_lk(gq) - 1) * gp, B);
      double eps[6];
      for (int r = 0; r < 6; r++) {
        double s2 = 0;
        for (int c = 0; c < 24; c++) s2 += B[r * 24 + c] * ue[c];
        eps[r] = s2;
      }
      eps[0] -= fsol; eps[1] -= fsol; eps[2] -= fsol;
      const double *ep = st + 11 * (8 * e + gq);
      const double vol = (eps[0] + eps[1] + eps[2]) / 3, sm = 3 * K * vol;
      for (int c = 0; c < 3; c++) gpS[6 * gq + c] = 2 * G * (eps[c] - vol - ep[c]) + sm;
      for (int c = 3; c < 6; c++) gpS[6 * gq + c] = 2 * G * (0.5 * eps[c] - ep[c]);
    }
  }
  const double s3 = 1.7320508075688772;  // sqrt(3)
  for (int a = 0; a < 8; a++) {
    const double x = (2 * fe::corner_li(a) - 1) * s3, y = (2 * fe::corner_lj(a) - 1) * s3,
                 z = (2 * fe::corner_lk(a) - 1) * s3;
    double N[8];
    for (int gq = 0; gq < 8; gq++)
      N[gq] = 0.125 * (1 + (2 * fe::corner_li(gq) - 1) * x) * (1 + (2 * fe::corner_lj(gq) - 1) * y) *
              (1 + (2 * fe::corner_lk(gq) - 1) * z);
    for (int c = 0; c < 6; c++) {
      double acc = 0;
      for (int gq = 0; gq < 8; gq++) acc += N[gq] * gpS[6 * gq + c];
      out[48 * e + 6 * a + c] = (float)acc;
    }
  }
}

void sub_section(fcbga_submodel *S, double E, double fsol, float *host_out) {
  be::stream_t s = S->stream;
  const SubView v = sub_view(S);
  const fe::AnandParams P = S->P;
  const double *u = S->u.p, *st = S->st.p, *sigE = S->sigE.p;
  DBuf<float> o;
  o.alloc(48 * S->ne);
  float *op = o.p;
  be::pfor(S->ne, s, [=] __host__ __device__(int64_t e) { sub_section_element(v, P, e, E, fsol, u, st, sigE, op); });
  be::d2h(host_out, op, sizeof(float) * 48 * S->ne, s);
}

void sub_loop_sample(fcbga_submodel *S, double T, double E, double fsol, double cphi, double sphi, double out[4]) {
  be::stream_t s = S->stream;
  const SubView v = sub_view(S);
  const fe::AnandParams P = S->P;
  const uint8_t *lay = S->cellLayer.p;
  const double *u = S->u.p;
  double *st = S->st.p;
  DBuf<double> part;
  part.alloc(6 * S->ne);
  double *pp = part.p;
  be::pfor(S->ne, s, [=] __host__ __device__(int64_t e) {
    sub_loop_element(v, P, e, T, E, fsol, cphi, sphi, lay, u, st, pp + 6 * e);
  });
  std::vector<double> h(6 * S->ne);
  be::d2h(h.data(), pp, sizeof(double) * h.size(), s);
  double a[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t e = 0; e < S->ne; e++)
    for (int c = 0; c < 6; c++) a[c] += h[6 * e + c];
  for (int L = 0; L < 2; L++) {
    out[2 * L] = a[3 * L + 2] ? a[3 * L] / a[3 * L + 2] : NAN;
    out[2 * L + 1] = a[3 * L + 2] ? a[3 * L + 1] / a[3 * L + 2] : NAN;
  }
}

}  // namespace
