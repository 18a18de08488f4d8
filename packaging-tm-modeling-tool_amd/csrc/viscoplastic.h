// viscoplastic.h -- global viscoplastic Newton of a package mesh with discrete
// solder bumps (the bump-array cycling workload): Anand material points at the
// 2 x 2 x 2 Gauss points of every solder element, B-bar (mean dilatation)
// solder elements, the incompatible-mode elastic elements of the rest of the
// mesh, and an inexact Newton whose inner solve is the step solve's flexible
// PCG with the shell-aggregate multigrid built on the current tangent.
//
// The reference has no global viscoplastic model: its Anand points live in
// the bump submodel.  The pieces it combines are the reference's own:
//   material point        AnandPoint.step          src/core/13-fatigue.js:67-143
//   B-bar element         bbarMatrix               src/core/13-fatigue.js:474-490
//   Newton time step      BumpSubmodel.stepTo      src/core/13-fatigue.js:924-961
//                         (residual and increment tolerances, the inexact inner
//                         tolerance, a final residual-only pass, then commit)
//   inner PCG             BumpSubmodel._solve      src/core/13-fatigue.js:886-917
// and the linear machinery of the step solve (element tables, Jacobi scaling,
// multigrid set-up, flexible PCG) for the elastic elements.
//
// The material points use the return map with the bracket-end guard
// (fe::anand_step robust = true): a non-positive effective stress is no flow.
//
// Layout: each solder element is an element type of its own in the model
// (fcbga_model::ownGroup), so the fine operator, the column smoother and the
// Galerkin products read its current 24 x 24 tangent from the type tables like
// any other element; the solder state lives here, per Gauss point (11 FP64:
// ep[6] tensor shear, s, ebar, W, sigEq, Newton warm start).
#pragma once

struct fcbga_vp {
  fcbga_model *M = nullptr;  // the global model (solder group with own types)
  fe::AnandParams P{};
  std::vector<double> ET, Ev;  // solder E(T) table (C, MPa), linear inside and beyond
  double cte = 0, Tref = 25;
  int32_t ns = 0;              // solder elements
  DBuf<int64_t> eNodes;        // 8 grid nodes per solder element (corner order)
  DBuf<double> eH;             // hx hy hz per solder element
  DBuf<int32_t> cellSolder;    // per cell: solder element index or -1
  DBuf<int32_t> sNodes;        // grid nodes with at least one solder element
  DBuf<double> st, tr, snapSt; // per Gauss point, 11 each
  DBuf<double> Ke, fe, ce, dmaxE;
  DBuf<int32_t> badE;
  DBuf<double> gpSig, gpD, gpDep;  // per Gauss point: stress (6), tangent (36), plastic increment
  DBuf<double> u, snapU;       // total displacement, grid layout (3 per grid node)
  DBuf<double> rg, fsv, xs_, duNode;
  std::vector<int32_t> elemOf; // mesh element of every solder element
  DBuf<int64_t> bcGrid;        // grid DOF of every constrained DOF (the mesh's constraint order)
  DBuf<double> bcVal;          // its prescribed displacement (zero unless set)
  double Tn = 25, snapTn = 25;
  bool haveSnap = false;
};

namespace {

constexpr int VP_EW = 4;  // solder elements (wavefronts) per workgroup

struct VpView {
  int32_t ns;
  const int64_t *nodes;
  const double *h;
};

double vp_E(const fcbga_vp *V, double T) {
  const std::vector<double> &X = V->ET, &Y = V->Ev;
  const size_t n = X.size();
  if (n == 1) return Y[0];
  size_t i = 0;
  while (i + 2 < n && T > X[i + 1]) i++;
  return Y[i] + (Y[i + 1] - Y[i]) * (T - X[i]) / (X[i + 1] - X[i]);
}

// A point whose consistent tangent is not finite (driven to the end of the
// return-map bracket, where the deviatoric stress vanishes and its direction
// is undefined) takes the secant tangent K 1x1 + 2 G theta I_dev with
// theta = sigEq / sigma_trial, at least 1e-3, so the element stays positive
// definite.
FE_HD void vp_secant_tangent(const fe::AnandParams &P, double E, double d, double sigEq, double *D) {
  const double G = E / (2 * (1 + P.nu)), K = E / (3 * (1 - 2 * P.nu));
  const double tri = sigEq + 3 * G * d;
  double th = tri > 0 ? sigEq / tri : 1.0;
  if (!(th >= 1e-3)) th = 1e-3;
  const double c1 = 2 * G * th;
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) {
      double v = 0;
      if (i < 3 && j < 3) v = K + c1 * ((i == j ? 1.0 : 0.0) - 1.0 / 3.0);
      else if (i == j) v = 0.5 * c1;
      D[6 * i + j] = v;
    }
}

FE_HD bool all_finite(const double *v, int n) {
  for (int q = 0; q < n; q++)
    if (!isfinite(v[q])) return false;
  return true;
}

// Gauss point gq of a box: natural coordinates (2 l - 1) / sqrt(3) of its corner
FE_HD void vp_gauss(int gq, double &xi, double &eta, double &zeta) {
  const double gp = 0x1.279a74590331dp-1;  // 1 / sqrt(3)
  xi = (2 * fe::corner_li(gq) - 1) * gp;
  eta = (2 * fe::corner_lj(gq) - 1) * gp;
  zeta = (2 * fe::corner_lk(gq) - 1) * gp;
}

// upper-triangle entry idx (0..299) of a 24 x 24 matrix -> (a, b), b >= a
FE_HD void upper_entry(int idx, int &a, int &b) {
  a = 0;
  int rem = idx;
  while (rem >= 24 - a) { rem -= 24 - a; a++; }
  b = a + rem;
}

// B-bar strain (engineering shear) at a point of a box from its corner
// displacements, without forming the 6 x 24 matrix: the normal rows share the
// volumetric part of the centre's derivatives, sum_a ((c_a - d_a) / 3) . u_a,
// and add their own derivative; the shear rows are the standard ones (the
// rows of fe::bbar_matrix applied to u).
FE_HD void bbar_strain(double hx, double hy, double hz, double xi, double eta, double zeta, const double *ue,
                       double eps[6]) {
  double dx[8], dy[8], dz[8], cx[8], cy[8], cz[8];
  fe::box_corner_derivs(hx, hy, hz, xi, eta, zeta, dx, dy, dz);
  fe::box_corner_derivs(hx, hy, hz, 0, 0, 0, cx, cy, cz);
  double vol = 0, e0 = 0, e1 = 0, e2 = 0, g01 = 0, g12 = 0, g20 = 0;
  for (int a = 0; a < 8; a++) {
    const double ux = ue[3 * a], uy = ue[3 * a + 1], uz = ue[3 * a + 2];
    vol += (cx[a] - dx[a]) / 3 * ux + (cy[a] - dy[a]) / 3 * uy + (cz[a] - dz[a]) / 3 * uz;
    e0 += dx[a] * ux;
    e1 += dy[a] * uy;
    e2 += dz[a] * uz;
    g01 += dy[a] * ux + dx[a] * uy;
    g12 += dz[a] * uy + dy[a] * uz;
    g20 += dz[a] * ux + dx[a] * uz;
  }
  eps[0] = e0 + vol; eps[1] = e1 + vol; eps[2] = e2 + vol;
  eps[3] = g01; eps[4] = g12; eps[5] = g20;
}

#if FCBGA_GPU
// One thread per solder Gauss point: the point's B-bar rows at its own
// position, the mechanical strain B u_e - F, the Anand trial state from the
// committed one, the stress and (wantK) the consistent tangent (the secant
// one where it is not finite).  Every lane of a wavefront runs a return map.
__global__ void __launch_bounds__(256) k_vp_points(VpView v, fe::AnandParams P, double T, double dt, double E,
                                                   double fsol, int wantK, const double *__restrict__ u, double *st,
                                                   double *tr, double *__restrict__ sigOut, double *__restrict__ Dout,
                                                   double *__restrict__ depOut) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= 8 * (int64_t)v.ns) return;
  const int64_t e = q >> 3;
  const int gq = (int)(q & 7);
  const double hx = v.h[3 * e], hy = v.h[3 * e + 1], hz = v.h[3 * e + 2];
  double ue[24];
  for (int a = 0; a < 8; a++) {
    const int64_t gn = v.nodes[8 * e + a];
    for (int c = 0; c < 3; c++) ue[3 * a + c] = u[3 * gn + c];
  }
  double xi, eta, zeta;
  vp_gauss(gq, xi, eta, zeta);
  double eps[6], sig[6], D[36];
  bbar_strain(hx, hy, hz, xi, eta, zeta, ue, eps);
  eps[0] -= fsol; eps[1] -= fsol; eps[2] -= fsol;
  const double d = fe::anand_step(P, E, T, dt, eps, st + 11 * q, tr + 11 * q, sig, wantK ? D : nullptr, true);
  if (wantK) {
    if (!all_finite(D, 36)) vp_secant_tangent(P, E, d, tr[11 * q + 9], D);
    for (int c = 0; c < 36; c++) Dout[36 * q + c] = D[c];
  }
  for (int r = 0; r < 6; r++) sigOut[6 * q + r] = sig[r];
  depOut[q] = all_finite(sig, 6) ? d : -1.0;  // -1: non-finite stress
}

// One wavefront per solder element, from its points' stresses and tangents:
// the internal force f = sum w B^T sigma, the tangent K = sum w B^T D B (300
// upper entries over the lanes, mirrored; only when wantK) and c = K u_e - f,
// the element's term of the residual correction; the largest plastic
// increment and a flag for non-finite stresses.
__global__ void __launch_bounds__(64 * VP_EW) k_vp_elements(VpView v, int wantK, const double *__restrict__ u,
                                                           const double *__restrict__ gpSig,
                                                           const double *__restrict__ gpD,
                                                           const double *__restrict__ gpDep, double *Kout, double *fe_,
                                                           double *ce, double *dmaxE, int32_t *badE) {
  __shared__ double Bs[VP_EW][8 * 144];
  __shared__ double Ds[VP_EW][8 * 36];
  __shared__ double Ss[VP_EW][8 * 6];
  __shared__ double Us[VP_EW][24];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t e = (int64_t)blockIdx.x * VP_EW + wv;
  const bool valid = e < v.ns;
  const double hx = valid ? v.h[3 * e] : 1, hy = valid ? v.h[3 * e + 1] : 1, hz = valid ? v.h[3 * e + 2] : 1;
  const double w = hx * hy * hz / 8;
  double *B = Bs[wv], *Dl = Ds[wv], *Sg = Ss[wv], *ue = Us[wv];
  double d = 0;
  int bad = 0;
  if (valid) {
    if (lane < 24) ue[lane] = u[3 * v.nodes[8 * e + lane / 3] + lane % 3];
    if (lane < 8) {
      double xi, eta, zeta;
      vp_gauss(lane, xi, eta, zeta);
      fe::bbar_matrix(hx, hy, hz, xi, eta, zeta, B + 144 * lane);
      d = gpDep[8 * e + lane];
      if (d < 0) bad = 1, d = 0;
    }
    if (lane < 48) Sg[lane] = gpSig[48 * e + lane];
    if (wantK)
      for (int c = lane; c < 288; c += 64) Dl[c] = gpD[288 * e + c];
  }
  __syncthreads();
  double *Ke = Kout + 576 * e;
  if (valid && wantK) {
#pragma unroll 1
    for (int idx = lane; idx < 300; idx += 64) {
      int a, bc;
      upper_entry(idx, a, bc);
      double acc = 0;
#pragma unroll 1
      for (int gq = 0; gq < 8; gq++) {
        const double *Bg = B + 144 * gq, *Dg = Dl + 36 * gq;
        double s = 0;
        for (int r = 0; r < 6; r++) {
          double db = 0;
          for (int t = 0; t < 6; t++) db += Dg[r * 6 + t] * Bg[t * 24 + bc];
          s += Bg[r * 24 + a] * db;
        }
        acc += w * s;
      }
      Ke[a * 24 + bc] = acc;
      Ke[bc * 24 + a] = acc;
    }
  }
  __syncthreads();
  if (valid && lane < 24) {
    double f = 0;
    for (int gq = 0; gq < 8; gq++) {
      const double *Bg = B + 144 * gq;
      double s = 0;
      for (int r = 0; r < 6; r++) s += Bg[r * 24 + lane] * Sg[6 * gq + r];
      f += w * s;
    }
    double ku = 0;
    for (int c = 0; c < 24; c++) ku += Ke[lane * 24 + c] * ue[c];
    fe_[24 * e + lane] = f;
    ce[24 * e + lane] = ku - f;
  }
  double m = d;
  int bb = bad;
  for (int o = 32; o >= 1; o >>= 1) {
    m = fmax(m, __shfl_xor(m, o));
    bb |= __shfl_xor(bb, o);
  }
  if (valid && lane == 0) {
    dmaxE[e] = m;
    badE[e] = bb;
  }
}
#endif

// host form of the element pass (host-emulation build): the same operations
// element by element
FE_HD void vp_element_host(const VpView &v, const fe::AnandParams &P, int64_t e, double T, double dt, double E,
                           double fsol, int wantK, const double *u, double *st, double *tr, double *Kout,
                           double *fe_, double *ce, double *dmaxE, int32_t *badE) {
  double B[8 * 144], Dl[8 * 36], Sg[8 * 6], ue[24];
  const double hx = v.h[3 * e], hy = v.h[3 * e + 1], hz = v.h[3 * e + 2], w = hx * hy * hz / 8;
  for (int q = 0; q < 24; q++) ue[q] = u[3 * v.nodes[8 * e + q / 3] + q % 3];
  double dm = 0;
  int bad = 0;
  for (int gq = 0; gq < 8; gq++) {
    double xi, eta, zeta;
    vp_gauss(gq, xi, eta, zeta);
    fe::bbar_matrix(hx, hy, hz, xi, eta, zeta, B + 144 * gq);
    double eps[6], sig[6];
    bbar_strain(hx, hy, hz, xi, eta, zeta, ue, eps);
    eps[0] -= fsol; eps[1] -= fsol; eps[2] -= fsol;
    const int64_t g8 = 8 * e + gq;
    double *Dg = Dl + 36 * gq;
    const double d = fe::anand_step(P, E, T, dt, eps, st + 11 * g8, tr + 11 * g8, sig, wantK ? Dg : nullptr, true);
    if (wantK && !all_finite(Dg, 36)) vp_secant_tangent(P, E, d, tr[11 * g8 + 9], Dg);
    if (!all_finite(sig, 6)) bad = 1;
    for (int r = 0; r < 6; r++) Sg[6 * gq + r] = sig[r];
    dm = fmax(dm, d);
  }
  double *Ke = Kout + 576 * e;
  if (wantK)
    for (int idx = 0; idx < 300; idx++) {
      int a, bc;
      upper_entry(idx, a, bc);
      double acc = 0;
      for (int gq = 0; gq < 8; gq++) {
        const double *Bg = B + 144 * gq, *Dg = Dl + 36 * gq;
        double s = 0;
        for (int r = 0; r < 6; r++) {
          double db = 0;
          for (int t = 0; t < 6; t++) db += Dg[r * 6 + t] * Bg[t * 24 + bc];
          s += Bg[r * 24 + a] * db;
        }
        acc += w * s;
      }
      Ke[a * 24 + bc] = acc;
      Ke[bc * 24 + a] = acc;
    }
  for (int a = 0; a < 24; a++) {
    double f = 0;
    for (int gq = 0; gq < 8; gq++) {
      const double *Bg = B + 144 * gq;
      double s = 0;
      for (int r = 0; r < 6; r++) s += Bg[r * 24 + a] * Sg[6 * gq + r];
      f += w * s;
    }
    double ku = 0;
    for (int c = 0; c < 24; c++) ku += Ke[a * 24 + c] * ue[c];
    fe_[24 * e + a] = f;
    ce[24 * e + a] = ku - f;
  }
  dmaxE[e] = dm;
  badE[e] = bad;
}

void vp_elements(fcbga_vp *V, double T, double dt, int wantK) {
  fcbga_model *M = V->M;
  const VpView v{V->ns, V->eNodes.p, V->eH.p};
  const fe::AnandParams P = V->P;
  const double E = vp_E(V, T), fsol = V->cte * (T - V->Tref);
  if (V->ns == 0) return;
#if FCBGA_GPU
  const int64_t npt = 8 * (int64_t)V->ns;
  hipLaunchKernelGGL(k_vp_points, dim3((unsigned)((npt + 255) / 256)), dim3(256), 0, M->stream, v, P, T, dt, E, fsol,
                     wantK, (const double *)V->u.p, V->st.p, V->tr.p, V->gpSig.p, V->gpD.p, V->gpDep.p);
  BE_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_vp_elements, dim3((V->ns + VP_EW - 1) / VP_EW), dim3(64 * VP_EW), 0, M->stream, v, wantK,
                     (const double *)V->u.p, (const double *)V->gpSig.p, (const double *)V->gpD.p,
                     (const double *)V->gpDep.p, V->Ke.p, V->fe.p, V->ce.p, V->dmaxE.p, V->badE.p);
  BE_CHECK(hipGetLastError());
#else
  const double *u = V->u.p;
  double *st = V->st.p, *tr = V->tr.p, *K = V->Ke.p, *f = V->fe.p, *c = V->ce.p, *dm = V->dmaxE.p;
  int32_t *bd = V->badE.p;
  be::pfor(V->ns, M->stream, [=] __host__ __device__(int64_t e) {
    vp_element_host(v, P, e, T, dt, E, fsol, wantK, u, st, tr, K, f, c, dm, bd);
  });
#endif
}

// largest plastic increment and the non-finite flag over the solder elements
void vp_element_summary(fcbga_vp *V, double &dmax, bool &bad) {
  std::vector<double> dm(V->ns);
  std::vector<int32_t> bd(V->ns);
  be::d2h(dm.data(), V->dmaxE.p, sizeof(double) * V->ns, V->M->stream);
  be::d2h(bd.data(), V->badE.p, sizeof(int32_t) * V->ns, V->M->stream);
  dmax = 0;
  bad = false;
  for (int32_t e = 0; e < V->ns; e++) {
    dmax = std::max(dmax, dm[e]);
    if (bd[e] || !std::isfinite(dm[e])) bad = true;
  }
}

// Out-of-balance force r = f_th - K_el u - f_solder(u) on the grid (zero on
// constrained and absent DOFs), from the prepared step: the operator applies
// K_t = K_el + sum K_e (own types hold K_e), so r = bload - K_t u + sum (K_e u_e
// - f_e), the last sum gathered at the solder nodes in a fixed order.  u holds
// the prescribed values on the constrained DOFs, which the unmasked apply
// carries into the free rows.  For the force scale |f_th| + |f_solder| +
// |K_el u|: fsv receives the gathered solder forces sum f_e, xs_ the elastic
// internal force on the free DOFs.
void vp_residual(fcbga_vp *V) {
  fcbga_model *M = V->M;
  be::stream_t s = M->stream;
  const fe::Grid g = M->g;
  const int64_t n3 = 3 * g.G;
  const double *u = V->u.p, *si = M->sinv64.p, *bl = M->bload.p, *wv = M->w64.p;
  double *xs = V->xs_.p, *q = M->q.p, *r = V->rg.p, *fs = V->fsv.p;
  // K_t u through the scaled operator with the unmasked scale on both sides
  // (s (K_t (s (u / s)))), so prescribed boundary values enter the free rows
  be::pfor(n3, s, [=] __host__ __device__(int64_t i) { xs[i] = u[i] * si[i]; });
  {
    FineOps<double> full = fine_ops<double>(M);
    full.w = M->s64.p;
    fine_apply<double>(full, xs, q, s);
  }
  // xs now receives the elastic elements' internal force K_el u on the free
  // DOFs (the solder elements' K_e u_e is taken off at their nodes below)
  be::pfor(n3, s, [=] __host__ __device__(int64_t i) {
    const double ku = q[i] * si[i];
    r[i] = wv[i] != 0 ? bl[i] - ku : 0.0;
    xs[i] = wv[i] != 0 ? ku : 0.0;
    fs[i] = 0.0;
  });
  const int32_t *cs = V->cellSolder.p, *sn = V->sNodes.p;
  const double *ce = V->ce.p, *fe_ = V->fe.p;
  const int64_t nsn = (int64_t)V->sNodes.n;
  be::pfor(nsn, s, [=] __host__ __device__(int64_t t) {
    const int64_t gi = sn[t];
    const int k = (int)(gi % g.nz);
    const int64_t cj = gi / g.nz;
    const int j = (int)(cj % g.ny), i = (int)(cj / g.ny);
    double add[3] = {0, 0, 0}, fadd[3] = {0, 0, 0};
    for (int a = 0; a < 8; a++) {
      // the cell that has this node as its corner a
      const int ci = i - fe::corner_li(a), cjj = j - fe::corner_lj(a), ck = k - fe::corner_lk(a);
      if (ci < 0 || cjj < 0 || ck < 0 || ci >= g.ncx || cjj >= g.ncy || ck >= g.ncz) continue;
      const int32_t e = cs[((int64_t)ci * g.ncy + cjj) * g.ncz + ck];
      if (e < 0) continue;
      for (int p = 0; p < 3; p++) {
        add[p] += ce[24 * (int64_t)e + 3 * a + p];
        fadd[p] += fe_[24 * (int64_t)e + 3 * a + p];
      }
    }
    for (int p = 0; p < 3; p++) {
      if (wv[3 * gi + p] != 0) {
        r[3 * gi + p] += add[p];
        xs[3 * gi + p] -= add[p] + fadd[p];  // K_e u_e = c_e + f_e
      }
      fs[3 * gi + p] = fadd[p];
    }
  });
}

double vp_norm(fcbga_vp *V, const double *a) {
  return std::sqrt(std::max(0.0, dev_dot(V->M, 3 * V->M->g.G, a, a)));
}

// the model's step data with the solder elements' tangents as their tables;
// element tables, scaling, column factors and multigrid for that operator
void vp_prepare(fcbga_vp *V, const fcbga_step *st, const fcbga_solve_opts &o, fcbga_solve_stats &S) {
  V->M->ownK64 = V->Ke.p;
  prepare_step(V->M, st, o, S);
}

fcbga_vp_stats vp_newton(fcbga_vp *V, const fcbga_step *st, double T, double dt, const fcbga_vp_opts &vo) {
  fcbga_model *M = V->M;
  be::stream_t s = M->stream;
  const double t0 = now_ms();
  fcbga_vp_stats R{};
  fcbga_solve_opts o = vo.inner;
  o.warm = 0;
  const int64_t n3 = 3 * M->g.G;
  bool conv = false;
  int it;
  double rn = 0, fscale = 0;
  // the prescribed displacement of the constrained DOFs at this step
  if (V->bcGrid.n) {
    double *u = V->u.p;
    const int64_t *bg = V->bcGrid.p;
    const double *bv = V->bcVal.p;
    be::pfor((int64_t)V->bcGrid.n, s, [=] __host__ __device__(int64_t q) { u[bg[q]] = bv[q]; });
  }
  for (it = 1; it <= vo.maxit; it++) {
    vp_elements(V, T, dt, 1);
    double dm;
    bool bad;
    vp_element_summary(V, dm, bad);
    if (bad) { R.nonfinite = 1; break; }
    fcbga_solve_stats S{};
    const double ts = now_ms();
    vp_prepare(V, st, o, S);
    R.ms_setup += now_ms() - ts;
    vp_residual(V);
    rn = vp_norm(V, V->rg.p);
    fscale = vp_norm(V, M->bload.p) + vp_norm(V, V->fsv.p) + vp_norm(V, V->xs_.p);
    if (!std::isfinite(rn)) { R.nonfinite = 1; break; }
    if (rn <= vo.res_tol * std::max(fscale, 1e-300)) { conv = true; break; }
    // inexact Newton: the inner solve only has to beat the Newton residual by a margin
    fcbga_solve_opts oi = o;
    oi.tol = std::max(vo.inner_floor, std::min(1e-3, 0.01 * vo.res_tol * fscale / std::max(rn, 1e-300)));
    fcbga_solve_stats Si{};
    const double tp = now_ms();
    pcg_run(M, oi, V->rg.p, V->duNode.p, nullptr, Si);
    R.ms_pcg += now_ms() - tp;
    R.pcg_iters += Si.iters;
    // u += du (grid), du = x w
    double *u = V->u.p, *xs = V->xs_.p;
    const double *x = M->x.p, *wv = M->w64.p;
    be::pfor(n3, s, [=] __host__ __device__(int64_t i) {
      const double d = x[i] * wv[i];
      u[i] += d;
      xs[i] = d;
    });
    const double dun = vp_norm(V, V->xs_.p), un = vp_norm(V, V->u.p);
    if (!std::isfinite(dun)) { R.nonfinite = 1; break; }
    if (dun < vo.du_tol * std::max(un, 1e-30)) { conv = true; it++; break; }
  }
  R.newton = it > vo.maxit ? vo.maxit : it;
  R.res = fscale > 0 ? rn / fscale : rn;
  if (conv) {
    // final state at the converged displacement, then commit the points
    vp_elements(V, T, dt, 0);
    double dm;
    bool bad;
    vp_element_summary(V, dm, bad);
    if (bad) conv = false, R.nonfinite = 1;
    R.dmax = dm;
    if (conv) {
      double *stp = V->st.p;
      const double *trp = V->tr.p;
      be::pfor(8 * (int64_t)V->ns, s, [=] __host__ __device__(int64_t q) {
        for (int c = 0; c < 10; c++) stp[11 * q + c] = trp[11 * q + c];
      });
      V->Tn = T;
    }
  }
  be::sync(s);
  R.converged = conv ? 1 : 0;
  R.ms = now_ms() - t0;
  return R;
}

void vp_create(const fcbga_mesh *mesh, const fcbga_vp_solder *sol, fcbga_vp *V) {
  if (!sol || sol->nE < 1 || !sol->E_T || !sol->E_v) throw be::Error("solder E(T) table missing");
  if (!(sol->anand.nu < 0.5) || !(sol->anand.A > 0) || !(sol->anand.m > 0)) throw be::Error("Anand constants out of range");
  for (int q = 0; q + 1 < sol->nE; q++)
    if (!(sol->E_T[q + 1] > sol->E_T[q])) throw be::Error("solder E(T) temperatures must increase");
  V->M = new fcbga_model();
  {
    const char *ev = getenv("FCBGA_FULL_TABLES");
    V->M->sym_allowed = !(ev && ev[0] == '1');
  }
  model_create(mesh, V->M, sol->group);
  fcbga_model *M = V->M;
  const fcbga_anand_params &a = sol->anand;
  V->P = fe::AnandParams{a.A, a.QR, a.xi, a.m, a.shat, a.n, a.h0, a.a, a.s0, a.nu};
  V->ET.assign(sol->E_T, sol->E_T + sol->nE);
  V->Ev.assign(sol->E_v, sol->E_v + sol->nE);
  V->cte = sol->cte;
  V->Tref = V->Tn = sol->Tref;
  V->ns = (int32_t)M->ownElems.size();
  V->elemOf = M->ownElems;
  const fe::Grid g = M->g;
  const int64_t ncells = (int64_t)g.ncx * g.ncy * g.ncz;
  std::vector<int64_t> nodes(8 * (size_t)V->ns);
  std::vector<double> h(3 * (size_t)V->ns);
  std::vector<int32_t> cellSolder(ncells, -1);
  std::vector<uint8_t> touch(g.G, 0);
  for (int32_t e = 0; e < V->ns; e++) {
    if (M->e_scale[V->elemOf[e]] != 1.0) throw be::Error("solder elements must have element scale 1");
    const int64_t c = M->ownCells[e];
    const int k = (int)(c % g.ncz), j = (int)((c / g.ncz) % g.ncy), i = (int)(c / ((int64_t)g.ncz * g.ncy));
    if (cellSolder[c] >= 0) throw be::Error("two solder elements in one cell");
    cellSolder[c] = e;
    h[3 * e] = M->xs[i + 1] - M->xs[i];
    h[3 * e + 1] = M->ys[j + 1] - M->ys[j];
    h[3 * e + 2] = M->zs[k + 1] - M->zs[k];
    for (int q = 0; q < 8; q++) {
      const int64_t gn = ((int64_t)(i + fe::corner_li(q)) * g.ny + j + fe::corner_lj(q)) * g.nz + k + fe::corner_lk(q);
      nodes[8 * e + q] = gn;
      touch[gn] = 1;
    }
  }
  std::vector<int32_t> sn;
  for (int64_t gi = 0; gi < g.G; gi++)
    if (touch[gi]) sn.push_back((int32_t)gi);
  // constrained DOFs on the grid (prescribed values, zero until set)
  std::vector<int64_t> gridOfNode(M->nn), bc(M->constraints.size());
  for (int64_t q = 0; q < g.G; q++)
    if (M->node_id[q] >= 0) gridOfNode[M->node_id[q]] = q;
  for (size_t q = 0; q < bc.size(); q++) bc[q] = 3 * gridOfNode[M->constraints[q] / 3] + M->constraints[q] % 3;
  std::vector<double> bz(bc.size(), 0.0);
  if (g.G > INT32_MAX) throw be::Error("grid too large for the solder node list");
  be::stream_t s = M->stream;
  V->eNodes.upload(nodes, s);
  V->eH.upload(h, s);
  V->cellSolder.upload(cellSolder, s);
  V->sNodes.upload(sn, s);
  std::vector<double> st0(11 * 8 * (size_t)V->ns, 0.0);
  for (size_t q = 0; q < 8 * (size_t)V->ns; q++) st0[11 * q + 6] = a.s0;
  V->st.upload(st0, s);
  V->tr.upload(st0, s);
  V->Ke.alloc(std::max<size_t>(1, 576 * (size_t)V->ns));
  V->fe.alloc(std::max<size_t>(1, 24 * (size_t)V->ns));
  V->ce.alloc(std::max<size_t>(1, 24 * (size_t)V->ns));
  V->dmaxE.alloc(std::max<size_t>(1, V->ns));
  V->badE.alloc(std::max<size_t>(1, V->ns));
  V->gpSig.alloc(std::max<size_t>(1, 48 * (size_t)V->ns));
  V->gpD.alloc(std::max<size_t>(1, 288 * (size_t)V->ns));
  V->gpDep.alloc(std::max<size_t>(1, 8 * (size_t)V->ns));
  V->u.alloc(3 * g.G);
  V->u.zero(s);
  if (!bc.empty()) {
    V->bcGrid.upload(bc, s);
    V->bcVal.upload(bz, s);
  }
  V->rg.alloc(3 * g.G);
  V->fsv.alloc(3 * g.G);
  V->xs_.alloc(3 * g.G);
  V->duNode.alloc(3 * (size_t)M->nn);
  be::sync(s);
}

void vp_save(fcbga_vp *V) {
  be::stream_t s = V->M->stream;
  V->snapU.alloc(V->u.n);
  V->snapSt.alloc(V->st.n);
  be::d2d(V->snapU.p, V->u.p, sizeof(double) * V->u.n, s);
  be::d2d(V->snapSt.p, V->st.p, sizeof(double) * V->st.n, s);
  V->snapTn = V->Tn;
  V->haveSnap = true;
  be::sync(s);
}

// restore the saved state (the points' Newton warm starts are kept)
void vp_restore(fcbga_vp *V) {
  if (!V->haveSnap) throw be::Error("no saved viscoplastic state");
  be::stream_t s = V->M->stream;
  be::d2d(V->u.p, V->snapU.p, sizeof(double) * V->u.n, s);
  double *st = V->st.p;
  const double *ss = V->snapSt.p;
  be::pfor(8 * (int64_t)V->ns, s, [=] __host__ __device__(int64_t q) {
    for (int c = 0; c < 10; c++) st[11 * q + c] = ss[11 * q + c];
  });
  V->Tn = V->snapTn;
  be::sync(s);
}

// grid layout <-> the reference's node order
void vp_to_nodes(fcbga_vp *V, const double *grid, double *nodes) {
  const int64_t *gof = V->M->gridOf.p;
  be::pfor(V->M->nn, V->M->stream, [=] __host__ __device__(int64_t n) {
    const int64_t gi = gof[n];
    for (int c = 0; c < 3; c++) nodes[3 * n + c] = grid[3 * gi + c];
  });
}

void vp_from_nodes(fcbga_vp *V, const double *nodes, double *grid) {
  const int64_t *gof = V->M->gridOf.p;
  const int64_t n3 = 3 * V->M->g.G;
  be::pfor(n3, V->M->stream, [=] __host__ __device__(int64_t i) { grid[i] = 0.0; });
  be::pfor(V->M->nn, V->M->stream, [=] __host__ __device__(int64_t n) {
    const int64_t gi = gof[n];
    for (int c = 0; c < 3; c++) grid[3 * gi + c] = nodes[3 * n + c];
  });
}

}  // namespace
