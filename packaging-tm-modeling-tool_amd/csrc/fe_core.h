// fe_core.h -- per-node and per-element arithmetic of the hex8 solver, written
// once as __host__ __device__ functions: the HIP kernels wrap them, and a host
// build runs the same code for CPU-side index checks.
//
// Layout conventions (device):
//   grid nodes g = (i * ny + j) * nz + k (z fastest, the reference's node order
//   with holes for absent nodes); vectors hold 3 DOFs per grid node.
//   cells c = (i * ncy + j) * ncz + k, two element slots per cell
//   (src/core/06-mesher.js:27-31); a slot holds a type id (-1 none) and a scale
//   (element scale x 1 or x 1e-6 for an unborn group, src/core/09-assembly.js:403).
//   Corner order of a hex: (0,0,0) (1,0,0) (1,1,0) (0,1,0) then the same at +z
//   (src/core/06-mesher.js:46).
#pragma once

#include <math.h>
#include <stdint.h>

#if !defined(__HIPCC__)
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#ifndef __forceinline__
#define __forceinline__ inline
#endif
#endif

#define FE_HD __host__ __device__ __forceinline__

namespace fe {

struct Grid {
  int nx, ny, nz;     // node lines
  int ncx, ncy, ncz;  // cells
  int64_t G;          // grid nodes
};

// corner index of the local offset (li, lj, lk)
FE_HD int corner_of(int li, int lj, int lk) { return lk * 4 + (lj ? (li ? 2 : 3) : (li ? 1 : 0)); }
FE_HD int corner_li(int a) { return (a & 3) == 1 || (a & 3) == 2; }
FE_HD int corner_lj(int a) { return (a & 3) >= 2; }
FE_HD int corner_lk(int a) { return a >= 4; }

// ---------------------------------------------------------------------------
// operator row of one grid node:
//   y_i = w_i * sum_j K_ij (w_j x_j) + (w_i == 0 ? x_i : 0)
// with w = s * mask (s the Jacobi scaling diag(K)^-1/2, mask 0 on constrained
// DOFs and absent nodes).  This is the Jacobi-scaled, constraint-reduced
// operator of src/core/09-assembly.js:445-464 and src/core/10-solver.js:852-866
// without the assembled matrix: K is gathered from the condensed element
// matrices of the (up to 8 x 2) elements around the node.
// ---------------------------------------------------------------------------
// PRESCALED: x already holds w*x (the cycle iterate y of the fine level); the
// row is then w * (A y) on free DOFs and 0 on constrained ones
template <typename T, typename TK, bool PRESCALED = false>
FE_HD void node_apply(const Grid &g, int i, int j, int k, const int32_t *__restrict__ cellType,
                      const TK *__restrict__ cellScale, const TK *__restrict__ Kc, const T *__restrict__ w,
                      const T *__restrict__ x, T out[3]) {
  // 27-point neighbourhood of w*x
  T nb[3][3][3][3];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int di = 0; di < 3; di++)
    for (int dj = 0; dj < 3; dj++)
      for (int dk = 0; dk < 3; dk++) {
        int ii = i + di - 1, jj = j + dj - 1, kk = k + dk - 1;
        bool in = ii >= 0 && ii < g.nx && jj >= 0 && jj < g.ny && kk >= 0 && kk < g.nz;
        int64_t gn = in ? ((int64_t)ii * g.ny + jj) * g.nz + kk : 0;
        for (int p = 0; p < 3; p++) nb[di][dj][dk][p] = in ? (PRESCALED ? x[3 * gn + p] : w[3 * gn + p] * x[3 * gn + p]) : T(0);
      }
  T a0 = 0, a1 = 0, a2 = 0;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int ci = 0; ci < 2; ci++)
    for (int cj = 0; cj < 2; cj++)
      for (int ck = 0; ck < 2; ck++) {
        int ii = i - 1 + ci, jj = j - 1 + cj, kk = k - 1 + ck;
        if (ii < 0 || ii >= g.ncx || jj < 0 || jj >= g.ncy || kk < 0 || kk >= g.ncz) continue;
        int64_t c = ((int64_t)ii * g.ncy + jj) * g.ncz + kk;
        const int a = corner_of(1 - ci, 1 - cj, 1 - ck);
        for (int q = 0; q < 2; q++) {
          int32_t t = cellType[2 * c + q];
          if (t < 0) continue;
          T sc = (T)cellScale[2 * c + q];
          const TK *K = Kc + (int64_t)t * 576 + (3 * a) * 24;
          T s0 = 0, s1 = 0, s2 = 0;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
          for (int b = 0; b < 8; b++) {
            const T *xv = nb[ci + corner_li(b)][cj + corner_lj(b)][ck + corner_lk(b)];
            s0 += (T)K[3 * b] * xv[0] + (T)K[3 * b + 1] * xv[1] + (T)K[3 * b + 2] * xv[2];
            s1 += (T)K[24 + 3 * b] * xv[0] + (T)K[24 + 3 * b + 1] * xv[1] + (T)K[24 + 3 * b + 2] * xv[2];
            s2 += (T)K[48 + 3 * b] * xv[0] + (T)K[48 + 3 * b + 1] * xv[1] + (T)K[48 + 3 * b + 2] * xv[2];
          }
          a0 += sc * s0; a1 += sc * s1; a2 += sc * s2;
        }
      }
  int64_t gi = ((int64_t)i * g.ny + j) * g.nz + k;
  T acc[3] = {a0, a1, a2};
  for (int p = 0; p < 3; p++) {
    T wi = w[3 * gi + p];
    out[p] = wi != T(0) ? wi * acc[p] : (PRESCALED ? T(0) : x[3 * gi + p]);
  }
}

// thermal load of a node: sum over its elements of scale * fc (src/core/09-assembly.js:416-420)
FE_HD void node_load(const Grid &g, int i, int j, int k, const int32_t *__restrict__ cellType,
                     const double *__restrict__ cellScale, const double *__restrict__ fc,
                     const uint8_t *__restrict__ hasF, double out[3]) {
  double a0 = 0, a1 = 0, a2 = 0;
  for (int ci = 0; ci < 2; ci++)
    for (int cj = 0; cj < 2; cj++)
      for (int ck = 0; ck < 2; ck++) {
        int ii = i - 1 + ci, jj = j - 1 + cj, kk = k - 1 + ck;
        if (ii < 0 || ii >= g.ncx || jj < 0 || jj >= g.ncy || kk < 0 || kk >= g.ncz) continue;
        int64_t c = ((int64_t)ii * g.ncy + jj) * g.ncz + kk;
        const int a = corner_of(1 - ci, 1 - cj, 1 - ck);
        for (int q = 0; q < 2; q++) {
          int32_t t = cellType[2 * c + q];
          if (t < 0 || !hasF[t]) continue;
          double sc = cellScale[2 * c + q];
          a0 += sc * fc[24 * (int64_t)t + 3 * a];
          a1 += sc * fc[24 * (int64_t)t + 3 * a + 1];
          a2 += sc * fc[24 * (int64_t)t + 3 * a + 2];
        }
      }
  out[0] = a0; out[1] = a1; out[2] = a2;
}

// Diagonal block K_nn and the column block K_{n,n-1} (node below) of the
// unscaled matrix, gathered from the elements (both 3x3 row-major).
FE_HD void node_blocks(const Grid &g, int i, int j, int k, const int32_t *__restrict__ cellType,
                       const double *__restrict__ cellScale, const double *__restrict__ Kc, double D[9],
                       double L[9]) {
  for (int q = 0; q < 9; q++) { D[q] = 0; L[q] = 0; }
  for (int ci = 0; ci < 2; ci++)
    for (int cj = 0; cj < 2; cj++)
      for (int ck = 0; ck < 2; ck++) {
        int ii = i - 1 + ci, jj = j - 1 + cj, kk = k - 1 + ck;
        if (ii < 0 || ii >= g.ncx || jj < 0 || jj >= g.ncy || kk < 0 || kk >= g.ncz) continue;
        int64_t c = ((int64_t)ii * g.ncy + jj) * g.ncz + kk;
        const int a = corner_of(1 - ci, 1 - cj, 1 - ck);
        const int ab = corner_of(1 - ci, 1 - cj, 0);  // same column, one plane down (valid when ck == 0)
        for (int q = 0; q < 2; q++) {
          int32_t t = cellType[2 * c + q];
          if (t < 0) continue;
          double sc = cellScale[2 * c + q];
          const double *K = Kc + (int64_t)t * 576;
          for (int p = 0; p < 3; p++)
            for (int r = 0; r < 3; r++) D[3 * p + r] += sc * K[(3 * a + p) * 24 + 3 * a + r];
          if (ck == 0)
            for (int p = 0; p < 3; p++)
              for (int r = 0; r < 3; r++) L[3 * p + r] += sc * K[(3 * a + p) * 24 + 3 * ab + r];
        }
      }
}

// chol3 with the jitter rule of src/core/10-solver.js:12-25, 60-64
FE_HD bool chol3(const double *A, double *L, double jit) {
  double a00 = A[0] + jit, a10 = A[3], a11 = A[4] + jit, a20 = A[6], a21 = A[7], a22 = A[8] + jit;
  if (!(a00 > 0)) return false;
  double l00 = sqrt(a00), l10 = a10 / l00, l20 = a20 / l00;
  double d11 = a11 - l10 * l10;
  if (!(d11 > 0)) return false;
  double l11 = sqrt(d11), l21 = (a21 - l20 * l10) / l11;
  double d22 = a22 - l20 * l20 - l21 * l21;
  if (!(d22 > 0)) return false;
  L[0] = l00; L[1] = 0; L[2] = 0; L[3] = l10; L[4] = l11; L[5] = 0; L[6] = l20; L[7] = l21; L[8] = sqrt(d22);
  return true;
}

// Block-tridiagonal Cholesky of one z-column of the scaled, constraint-reduced
// matrix (factorColumns, src/core/10-solver.js:31-68).  Db/Lb hold the scaled
// blocks per grid node of the column (k = 0..nz-1); F receives 18 values per
// node: L_kk (lower 3x3) and L_{k,k-1}.
FE_HD void column_factor(int nz, const double *Db, const double *Lb, double *F) {
  double tmp[9], D[9];
  for (int k = 0; k < nz; k++) {
    for (int q = 0; q < 9; q++) D[q] = Db[9 * k + q];
    double *fo = F + 18 * k;
    bool haveDown = false;
    if (k > 0) {
      for (int q = 0; q < 9; q++) haveDown |= Lb[9 * k + q] != 0.0;
    }
    if (haveDown) {
      const double *av = Lb + 9 * k, *lp = F + 18 * (k - 1);
      for (int r = 0; r < 3; r++) {
        double b0 = av[3 * r], b1 = av[3 * r + 1], b2 = av[3 * r + 2];
        double y0 = b0 / lp[0];
        double y1 = (b1 - lp[3] * y0) / lp[4];
        double y2 = (b2 - lp[6] * y0 - lp[7] * y1) / lp[8];
        tmp[3 * r] = y0; tmp[3 * r + 1] = y1; tmp[3 * r + 2] = y2;
      }
      for (int q = 0; q < 9; q++) fo[9 + q] = tmp[q];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++)
          D[3 * r + c] -= tmp[3 * r] * tmp[3 * c] + tmp[3 * r + 1] * tmp[3 * c + 1] + tmp[3 * r + 2] * tmp[3 * c + 2];
    } else {
      for (int q = 0; q < 9; q++) fo[9 + q] = 0;
    }
    double jit = 0;
    bool ok = chol3(D, fo, jit);
    for (int t = 0; t < 8 && !ok; t++) {
      double mx = fabs(D[0]);
      if (fabs(D[4]) > mx) mx = fabs(D[4]);
      if (fabs(D[8]) > mx) mx = fabs(D[8]);
      if (mx < 1e-300) mx = 1e-300;
      jit = jit != 0 ? jit * 10 : 1e-10 * mx;
      ok = chol3(D, fo, jit);
    }
    if (!ok) {  // an all-zero block: identity factor keeps the solve finite
      for (int q = 0; q < 9; q++) fo[q] = (q % 4 == 0) ? 1.0 : 0.0;
    }
  }
}

// Forward/back substitution of one column (solveColumn, src/core/10-solver.js:71-104):
// x = (L L^T)^-1 r over nodes 0..nz-1 of the column (stride between nodes 3).
template <typename T, typename TF>
FE_HD void column_solve(int nz, const TF *__restrict__ F, const T *__restrict__ r, T *__restrict__ x) {
  for (int n = 0; n < nz; n++) {
    const TF *fo = F + 18 * n;
    T b0 = r[3 * n], b1 = r[3 * n + 1], b2 = r[3 * n + 2];
    if (n > 0) {
      T p0 = x[3 * n - 3], p1 = x[3 * n - 2], p2 = x[3 * n - 1];
      b0 -= (T)fo[9] * p0 + (T)fo[10] * p1 + (T)fo[11] * p2;
      b1 -= (T)fo[12] * p0 + (T)fo[13] * p1 + (T)fo[14] * p2;
      b2 -= (T)fo[15] * p0 + (T)fo[16] * p1 + (T)fo[17] * p2;
    }
    T y0 = b0 / (T)fo[0];
    T y1 = (b1 - (T)fo[3] * y0) / (T)fo[4];
    T y2 = (b2 - (T)fo[6] * y0 - (T)fo[7] * y1) / (T)fo[8];
    x[3 * n] = y0; x[3 * n + 1] = y1; x[3 * n + 2] = y2;
  }
  for (int n = nz - 1; n >= 0; n--) {
    const TF *fo = F + 18 * n;
    T b0 = x[3 * n], b1 = x[3 * n + 1], b2 = x[3 * n + 2];
    if (n + 1 < nz) {
      const TF *fn = F + 18 * (n + 1);
      T q0 = x[3 * n + 3], q1 = x[3 * n + 4], q2 = x[3 * n + 5];
      b0 -= (T)fn[9] * q0 + (T)fn[12] * q1 + (T)fn[15] * q2;
      b1 -= (T)fn[10] * q0 + (T)fn[13] * q1 + (T)fn[16] * q2;
      b2 -= (T)fn[11] * q0 + (T)fn[14] * q1 + (T)fn[17] * q2;
    }
    T z2 = b2 / (T)fo[8];
    T z1 = (b1 - (T)fo[7] * z2) / (T)fo[4];
    T z0 = (b0 - (T)fo[3] * z1 - (T)fo[6] * z2) / (T)fo[0];
    x[3 * n] = z0; x[3 * n + 1] = z1; x[3 * n + 2] = z2;
  }
}

// 3x3 block K_ij between grid nodes (i,j,k) and (ii,jj,kk) (|offset| <= 1 per
// axis), gathered from the elements of the cells that contain both.
FE_HD void pair_block(const Grid &g, int i, int j, int k, int ii, int jj, int kk, const int32_t *__restrict__ cellType,
                      const double *__restrict__ cellScale, const double *__restrict__ Kc, double B[9]) {
  for (int q = 0; q < 9; q++) B[q] = 0;
  int cx0 = (i > ii ? i : ii) - 1, cx1 = i < ii ? i : ii;
  int cy0 = (j > jj ? j : jj) - 1, cy1 = j < jj ? j : jj;
  int cz0 = (k > kk ? k : kk) - 1, cz1 = k < kk ? k : kk;
  for (int ci = cx0; ci <= cx1; ci++) {
    if (ci < 0 || ci >= g.ncx) continue;
    for (int cj = cy0; cj <= cy1; cj++) {
      if (cj < 0 || cj >= g.ncy) continue;
      for (int ck = cz0; ck <= cz1; ck++) {
        if (ck < 0 || ck >= g.ncz) continue;
        int64_t c = ((int64_t)ci * g.ncy + cj) * g.ncz + ck;
        int a = corner_of(i - ci, j - cj, k - ck), b = corner_of(ii - ci, jj - cj, kk - ck);
        for (int q = 0; q < 2; q++) {
          int32_t t = cellType[2 * c + q];
          if (t < 0) continue;
          double sc = cellScale[2 * c + q];
          const double *K = Kc + (int64_t)t * 576;
          for (int p = 0; p < 3; p++)
            for (int r = 0; r < 3; r++) B[3 * p + r] += sc * K[(3 * a + p) * 24 + 3 * b + r];
        }
      }
    }
  }
}

// Dense Cholesky with the relative jitter rule of cholDense (src/core/01-util.js:70-91)
FE_HD int chol_dense(double *A, int n, double rel) {
  double dmax = 0;
  for (int i = 0; i < n; i++) dmax = fmax(dmax, fabs(A[i * n + i]));
  double jitter = rel * (dmax != 0 ? dmax : 1);
  int fixes = 0;
  for (int j = 0; j < n; j++) {
    double *rj = A + j * n;
    double d = rj[j];
    for (int k = 0; k < j; k++) d -= rj[k] * rj[k];
    if (!(d > jitter)) { d = fmax(fabs(d), jitter); fixes++; }
    double ljj = sqrt(d);
    rj[j] = ljj;
    double inv = 1 / ljj;
    for (int i = j + 1; i < n; i++) {
      double *ri = A + i * n;
      double s = ri[j];
      for (int k = 0; k < j; k++) s -= ri[k] * rj[k];
      ri[j] = s * inv;
    }
  }
  return fixes;
}

// L L^T x = b in place (cholSolveDense, src/core/01-util.js:94-107)
template <typename T, typename TL>
FE_HD void chol_solve(const TL *L, int n, T *b) {
  for (int i = 0; i < n; i++) {
    T s = b[i];
    for (int k = 0; k < i; k++) s -= (T)L[i * n + k] * b[k];
    b[i] = s / (T)L[i * n + i];
  }
  for (int i = n - 1; i >= 0; i--) {
    T s = b[i];
    for (int k = i + 1; k < n; k++) s -= (T)L[k * n + i] * b[k];
    b[i] = s / (T)L[i * n + i];
  }
}

// Condensed element of a box (hx, hy, hz): K33 = sum_m c_m M^m
// (elementK33Fast, src/core/09-assembly.js:96-108), thermal load from the
// natural-coordinate integrals (elementFth, src/core/09-assembly.js:115-133),
// static condensation of the 9 incompatible DOFs with a 1e-14 jittered
// Cholesky and exact symmetrisation (condense33, src/core/07-element.js:88-120).
// layered != 0: V holds 3 zeta-station stress weights (18) and gz the per-station
// integrals g[3][3][11]; else s0 (6) with the full-box integrals gH[3][11].
FE_HD void element_condense(double hx, double hy, double hz, const double *__restrict__ M, int hasF, int layered,
                            const double *__restrict__ load, const double *__restrict__ gz,
                            const double *__restrict__ gH, double *__restrict__ Kc, double *__restrict__ fc) {
  double K[1089], f[33], Kaa[81], X[216];
  const double V8 = hx * hy * hz / 8;
  const double c[6] = {V8 * 4 / (hx * hx), V8 * 4 / (hy * hy), V8 * 4 / (hz * hz),
                       V8 * 4 / (hx * hy), V8 * 4 / (hy * hz), V8 * 4 / (hz * hx)};
  for (int i = 0; i < 33; i++)
    for (int j = i; j < 33; j++) {
      int o = i * 33 + j;
      double v = c[0] * M[o] + c[1] * M[1089 + o] + c[2] * M[2178 + o] + c[3] * M[3267 + o] + c[4] * M[4356 + o] +
                 c[5] * M[5445 + o];
      K[o] = v;
      K[j * 33 + i] = v;
    }
  for (int q = 0; q < 33; q++) f[q] = 0;
  if (hasF) {
    const int VI[3][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}};
    const double h[3] = {hx, hy, hz};
    int nst = layered ? 3 : 1;
    for (int st = 0; st < nst; st++) {
      const double *S = load + 6 * st;
      for (int r = 0; r < 3; r++) {
        double cr = V8 * (2 / h[r]);
        const double *gb = layered ? gz + (st * 3 + r) * 11 : gH + r * 11;
        for (int p = 0; p < 3; p++) {
          double sv = cr * S[VI[p][r]];
          if (sv == 0) continue;
          for (int a = 0; a < 11; a++) f[3 * a + p] += sv * gb[a];
        }
      }
    }
  }
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 9; j++) Kaa[i * 9 + j] = K[(24 + i) * 33 + 24 + j];
  chol_dense(Kaa, 9, 1e-14);
  double col[9];
  for (int j = 0; j < 24; j++) {
    for (int i = 0; i < 9; i++) col[i] = K[(24 + i) * 33 + j];
    chol_solve<double, double>(Kaa, 9, col);
    for (int i = 0; i < 9; i++) X[i * 24 + j] = col[i];
  }
  for (int i = 0; i < 24; i++)
    for (int j = 0; j < 24; j++) {
      double s = K[i * 33 + j];
      for (int q = 0; q < 9; q++) s -= K[i * 33 + 24 + q] * X[q * 24 + j];
      Kc[i * 24 + j] = s;
    }
  for (int i = 0; i < 24; i++)
    for (int j = i + 1; j < 24; j++) {
      double v = 0.5 * (Kc[i * 24 + j] + Kc[j * 24 + i]);
      Kc[i * 24 + j] = v;
      Kc[j * 24 + i] = v;
    }
  if (hasF) {
    for (int i = 0; i < 9; i++) col[i] = f[24 + i];
    chol_solve<double, double>(Kaa, 9, col);
    for (int i = 0; i < 24; i++) {
      double s = f[i];
      for (int q = 0; q < 9; q++) s -= K[i * 33 + 24 + q] * col[q];
      fc[i] = s;
    }
  } else {
    for (int i = 0; i < 24; i++) fc[i] = 0;
  }
}

// ---------------------------------------------------------------------------
// Column solve as two linear recurrences (the same solve as column_solve):
//   forward  L y = r :   y_k = P_k y_{k-1} + Li_k r_k,     P_k = -Li_k L_{k,k-1}
//   backward L^T x = y:  x_k = Q_k x_{k+1} + Li_k^T y_k,   Q_k = -Li_k^T L_{k+1,k}^T
// with Li_k = L_kk^-1.  Per node 27 coefficients: Li (9), P (9), Q (9).  A
// recurrence of affine maps is solved by an inclusive scan in log2(nz) rounds.
// ---------------------------------------------------------------------------
template <typename T>
FE_HD void inv_lower3(const T *L, T *Li) {
  // inverse of a lower-triangular 3x3 (row-major)
  const T a = L[0], b = L[3], c = L[4], d = L[6], e = L[7], f = L[8];
  Li[0] = T(1) / a; Li[1] = 0; Li[2] = 0;
  Li[4] = T(1) / c; Li[5] = 0;
  Li[8] = T(1) / f;
  Li[3] = -b / (a * c);
  Li[7] = -e / (c * f);
  Li[6] = (b * e - c * d) / (a * c * f);
}

// Scan coefficients of node k from its factor Fk (18) and the sub-diagonal
// block of node k+1 (Ld1, the 9 values after Fk+1's diagonal factor; unused
// at the top of the column): out = Li (9), P (9), Q (9).
template <typename T>
FE_HD void node_scan_coefs(const T *Fk, bool hasBelow, bool hasAbove, const T *Ld1, T out[27]) {
  T Li[9];
  inv_lower3<T>(Fk, Li);
  for (int q = 0; q < 9; q++) out[q] = Li[q];
  const T *Ld = Fk + 9;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      T v = 0;
      if (hasBelow)
        for (int t = 0; t < 3; t++) v -= Li[3 * r + t] * Ld[3 * t + c];
      out[9 + 3 * r + c] = v;
      T w = 0;
      if (hasAbove)
        for (int t = 0; t < 3; t++) w -= Li[3 * t + r] * Ld1[3 * c + t];
      out[18 + 3 * r + c] = w;
    }
}

// Scan coefficients of a node from the column's factor (18 per node),
// SCAN_COEFS per node: the lower triangle of Li by rows (Li00, Li10, Li11,
// Li20, Li21, Li22; its upper entries are zero), then P (9) and Q (9): 24
// values, so a node's record is 16-byte aligned (six float4 loads in FP32).
// Formed once per factorisation in FP64; the sweeps read them instead of
// inverting the diagonal factors.
constexpr int SCAN_COEFS = 24;

// node k of a column of nz nodes: Fk its factor record (the next node's
// follows it), out its SCAN_COEFS coefficients
FE_HD void node_scan_coefs_at(int k, int nz, const double *Fk, double *out) {
  double Ld1[9], full[27];
  for (int q = 0; q < 9; q++) Ld1[q] = k + 1 < nz ? Fk[18 + 9 + q] : 0.0;
  node_scan_coefs<double>(Fk, k > 0, k + 1 < nz, Ld1, full);
  out[0] = full[0]; out[1] = full[3]; out[2] = full[4];
  out[3] = full[6]; out[4] = full[7]; out[5] = full[8];
  for (int q = 0; q < 18; q++) out[6 + q] = full[9 + q];
}

// Device layout of the records: per grid column, the records' 16-byte
// vectors (four FP32 or two FP64 values) are stored plane-major within each
// vector slot, so that the lanes of a wavefront (one plane each) read one
// slot of a column as one contiguous run: value q of node (column c, plane k)
// sits at ((c * NV + q / W) * nz + k) * W + q % W, W = 16 / sizeof(T),
// NV = SCAN_COEFS / W.
template <typename T>
FE_HD int64_t scan_coef_index(int64_t c, int k, int nz, int q) {
  constexpr int W = 16 / (int)sizeof(T), NV = SCAN_COEFS / W;
  return ((c * NV + q / W) * nz + k) * W + q % W;
}

// the 27 coefficients (Li 9 with its zero upper triangle, P, Q) of node k of
// column c from the device layout
template <typename T, typename TF>
FE_HD void unpack_scan_coefs(const TF *C, int64_t c, int k, int nz, T out[27]) {
  TF v[SCAN_COEFS];
  for (int q = 0; q < SCAN_COEFS; q++) v[q] = C[scan_coef_index<TF>(c, k, nz, q)];
  out[0] = (T)v[0]; out[1] = 0; out[2] = 0;
  out[3] = (T)v[1]; out[4] = (T)v[2]; out[5] = 0;
  out[6] = (T)v[3]; out[7] = (T)v[4]; out[8] = (T)v[5];
  for (int q = 0; q < 18; q++) out[9 + q] = (T)v[6 + q];
}

// Host form of the scan solve (the GPU kernel does the same rounds with lane
// shuffles): x = (L L^T)^-1 r for column c from the scan coefficients Cd
// (device layout), nz <= 128.
template <typename T, typename TF>
FE_HD void column_solve_scan(int nz, const TF *Cd, int64_t col, const T *r, T *x) {
  T A[128][9], b[128][3], A2[128][9], b2[128][3], C[128][27];
  for (int k = 0; k < nz; k++) unpack_scan_coefs<T, TF>(Cd, col, k, nz, C[k]);
  // forward: maps y_k = P_k y_{k-1} + Li_k r_k
  for (int k = 0; k < nz; k++) {
    const T *o = C[k];
    for (int q = 0; q < 9; q++) A[k][q] = o[9 + q];
    for (int p = 0; p < 3; p++) b[k][p] = o[3 * p] * r[3 * k] + o[3 * p + 1] * r[3 * k + 1] + o[3 * p + 2] * r[3 * k + 2];
  }
  for (int s = 1; s < nz; s <<= 1) {
    for (int k = 0; k < nz; k++) {
      if (k < s) { for (int q = 0; q < 9; q++) A2[k][q] = A[k][q]; for (int p = 0; p < 3; p++) b2[k][p] = b[k][p]; continue; }
      const T *Ak = A[k], *Ap = A[k - s], *bp = b[k - s];
      for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) A2[k][3 * i + j] = Ak[3 * i] * Ap[j] + Ak[3 * i + 1] * Ap[3 + j] + Ak[3 * i + 2] * Ap[6 + j];
        b2[k][i] = Ak[3 * i] * bp[0] + Ak[3 * i + 1] * bp[1] + Ak[3 * i + 2] * bp[2] + b[k][i];
      }
    }
    for (int k = 0; k < nz; k++) { for (int q = 0; q < 9; q++) A[k][q] = A2[k][q]; for (int p = 0; p < 3; p++) b[k][p] = b2[k][p]; }
  }
  // backward: maps x_k = Q_k x_{k+1} + Li_k^T y_k
  for (int k = 0; k < nz; k++) {
    const T *o = C[k];
    T y0 = b[k][0], y1 = b[k][1], y2 = b[k][2];
    for (int q = 0; q < 9; q++) A[k][q] = o[18 + q];
    for (int p = 0; p < 3; p++) b[k][p] = o[p] * y0 + o[3 + p] * y1 + o[6 + p] * y2;
  }
  for (int s = 1; s < nz; s <<= 1) {
    for (int k = 0; k < nz; k++) {
      if (k + s >= nz) { for (int q = 0; q < 9; q++) A2[k][q] = A[k][q]; for (int p = 0; p < 3; p++) b2[k][p] = b[k][p]; continue; }
      const T *Ak = A[k], *An = A[k + s], *bn = b[k + s];
      for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) A2[k][3 * i + j] = Ak[3 * i] * An[j] + Ak[3 * i + 1] * An[3 + j] + Ak[3 * i + 2] * An[6 + j];
        b2[k][i] = Ak[3 * i] * bn[0] + Ak[3 * i + 1] * bn[1] + Ak[3 * i + 2] * bn[2] + b[k][i];
      }
    }
    for (int k = 0; k < nz; k++) { for (int q = 0; q < 9; q++) A[k][q] = A2[k][q]; for (int p = 0; p < 3; p++) b[k][p] = b2[k][p]; }
  }
  for (int k = 0; k < nz; k++)
    for (int p = 0; p < 3; p++) x[3 * k + p] = b[k][p];
}

// ---------------------------------------------------------------------------
// Anand viscoplastic material point, backward Euler with J2 radial return
// (AnandPoint.step, src/core/13-fatigue.js:32-143, 209-218):
//   ep_dot = A sinh(xi sigma / s)^(1/m) exp(-Q/RT),  s_dot = h0 |B|^a sign(B) ep_dot,
//   B = 1 - s/s*,  s* = s_hat [(ep_dot/A) exp(Q/RT)]^n.
// Outer Newton (with bisection safeguard) on the equivalent plastic increment d,
// inner Newton/bisection for s_{n+1}(d); consistent tangent
//   D = K 1x1 + 2G theta I_dev - 2G theta_bar n x n  (Voigt, engineering shear).
// State per point (11): ep[6] (tensor shear), s, ebar, W, sigEq, lastD.
// ---------------------------------------------------------------------------
struct AnandParams {
  double A, QR, xi, m, shat, n, h0, a, s0, nu;
};

// The return map relies on exact cancellations (sigma_trial - 3 G d is exactly 0
// at the bracket end d = sigma_trial / 3G); fused multiply-adds would leave a
// tiny negative value there and sinh(x)^(1/m) would turn NaN.  The Anand
// functions are therefore compiled without floating-point contraction.
#if defined(__clang__)
#define FE_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define FE_NO_CONTRACT
#endif

constexpr int ANAND_MAXIT = 60;       // src/core/00-constants.js:123
constexpr double ANAND_TOL = 1e-11;   // src/core/00-constants.js:124
constexpr double KELVIN = 273.15;

FE_HD double sgn(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }

FE_HD double anand_solve_s(const AnandParams &P, double sn, double d, double dt, double expQ) {
  FE_NO_CONTRACT
  if (d <= 0) return sn;
  const double sstar = P.shat * pow(d / (dt * P.A) * expQ, P.n);
  double lo = fmin(sn, sstar), hi = fmax(sn, sstar);
  if (hi - lo < 1e-14) return sn;
  double s = 0.5 * (lo + hi);
  for (int it = 0; it < ANAND_MAXIT; it++) {
    const double B0 = 1 - s / sstar;
    const double f = s - sn - P.h0 * pow(fabs(B0), P.a) * sgn(B0) * d;
    if (fabs(f) < ANAND_TOL * fmax(1.0, sn)) break;
    if (f > 0) hi = s; else lo = s;
    const double dphi = 1 + P.h0 * P.a * pow(fabs(B0), P.a - 1) * d / sstar;
    double s1 = s - f / dphi;
    if (!(s1 > lo && s1 < hi)) s1 = 0.5 * (lo + hi);
    if (fabs(s1 - s) < 1e-15 * fmax(1.0, s)) { s = s1; break; }
    s = s1;
  }
  return s;
}

FE_HD double anand_dsdd(const AnandParams &P, double d, double s, double dt, double expQ) {
  FE_NO_CONTRACT
  if (d <= 0) return 0;
  const double sstar = P.shat * pow(d / (dt * P.A) * expQ, P.n);
  const double B = 1 - s / sstar, absB = fabs(B), sg = sgn(B);
  const double dphi_ds = 1 + P.h0 * P.a * pow(absB, P.a - 1) * d / sstar;
  const double dphi_dd = -P.h0 * pow(absB, P.a) * sg - P.h0 * P.a * pow(absB, P.a - 1) * s * P.n / sstar;
  return -dphi_dd / dphi_ds;
}

// One trial step from the committed state st (st[10], the warm start, is
// updated in place as the reference does); the trial state goes to out (11).
// Returns the equivalent plastic strain increment.
// robust = false is the reference's iteration, kept bit for bit where the
// reference runs it (screening, submodel): there an iterate at the bracket end
// d = sigma_trial / 3G, where sigma_trial - 3 G d rounds to a tiny negative,
// makes sinh(x)^(1/m) NaN and the NaN branch raises the lower bracket to the
// end, so the iteration stays on the spurious root d = dmax (sigma_eq = 0).
// robust = true (the global viscoplastic Newton, which has no reference
// counterpart) takes a non-positive effective stress as no flow: g(d) = d > 0,
// an upper bracket, so the iteration returns to the interior root.
FE_HD double anand_step(const AnandParams &P, double E, double T, double dt, const double eps[6], double *st,
                        double *out, double sig[6], double *D, bool robust = false) {
  FE_NO_CONTRACT
  const double G = E / (2 * (1 + P.nu)), K = E / (3 * (1 - 2 * P.nu));
  const double expQ = exp(P.QR / (T + KELVIN));
  const double rate0 = dt * P.A / expQ;
  const double vol = (eps[0] + eps[1] + eps[2]) / 3;
  double e[6], str[6];
  e[0] = eps[0] - vol - st[0]; e[1] = eps[1] - vol - st[1]; e[2] = eps[2] - vol - st[2];
  e[3] = 0.5 * eps[3] - st[3]; e[4] = 0.5 * eps[4] - st[4]; e[5] = 0.5 * eps[5] - st[5];
  for (int i = 0; i < 6; i++) str[i] = 2 * G * e[i];
  const double nrm = sqrt(str[0] * str[0] + str[1] * str[1] + str[2] * str[2] +
                          2 * (str[3] * str[3] + str[4] * str[4] + str[5] * str[5]));
  const double sigTr = sqrt(1.5) * nrm;
  const double sn = st[6];
  double d = 0, s1 = sn, dDdSig = 0;
  if (sigTr > 1e-14 && dt > 0) {
    const double dmax = sigTr / (3 * G);
    double lo = 0, hi = dmax;
    const double lastD = st[10];
    double dd = (lastD > 0 && lastD < dmax) ? lastD : fmin(dmax, rate0 * pow(sinh(P.xi * sigTr / sn), 1 / P.m));
    if (!isfinite(dd) || dd <= 0) dd = 0.5 * dmax;
    double rs = sn, rx = 0, rsh = 0, rg = 0;
    for (int it = 0; it < ANAND_MAXIT; it++) {
      rs = anand_solve_s(P, sn, dd, dt, expQ);
      rx = P.xi * (sigTr - 3 * G * dd) / rs;
      if (robust && !(rx > 0)) {
        rsh = 0;
        rg = dd;
      } else {
        rsh = sinh(rx);
        rg = dd - rate0 * pow(rsh, 1 / P.m);
      }
      if (!isfinite(rg)) { lo = dd; dd = 0.5 * (lo + hi); continue; }
      if (fabs(rg) < ANAND_TOL * fmax(dmax, 1e-30)) break;
      if (rg > 0) hi = dd; else lo = dd;
      const double fx = rate0 * (1 / P.m) * pow(rsh, 1 / P.m - 1) * cosh(rx);
      const double dsdd = anand_dsdd(P, dd, rs, dt, expQ);
      const double sg = sigTr - 3 * G * dd;
      const double dg = 1 + fx * P.xi * (3 * G / rs + sg * dsdd / (rs * rs));
      double dn = dd - rg / dg;
      if (!(dn > lo && dn < hi) || !isfinite(dn)) dn = 0.5 * (lo + hi);
      if (fabs(dn - dd) < 1e-16 * fmax(dmax, 1e-30)) { dd = dn; break; }
      dd = dn;
    }
    d = dd;
    s1 = rs;
    st[10] = d;
    if (D) {
      const double fx = rate0 * (1 / P.m) * pow(rsh, 1 / P.m - 1) * cosh(rx);
      const double dsdd = anand_dsdd(P, dd, rs, dt, expQ);
      const double sg = sigTr - 3 * G * dd;
      const double dg = 1 + fx * P.xi * (3 * G / rs + sg * dsdd / (rs * rs));
      dDdSig = (fx * P.xi / rs) / dg;
    }
  }
  const double theta = sigTr > 1e-14 ? 1 - 3 * G * d / sigTr : 1;
  double nv[6];
  for (int i = 0; i < 6; i++) nv[i] = nrm > 0 ? str[i] / nrm : 0.0;
  const double dgm = sqrt(1.5) * d;
  for (int i = 0; i < 6; i++) out[i] = st[i] + dgm * nv[i];
  const double sm = 3 * K * vol;
  sig[0] = theta * str[0] + sm; sig[1] = theta * str[1] + sm; sig[2] = theta * str[2] + sm;
  sig[3] = theta * str[3]; sig[4] = theta * str[4]; sig[5] = theta * str[5];
  const double sigEq = sigTr - 3 * G * d;
  out[6] = s1;
  out[7] = st[7] + d;
  out[8] = st[8] + 0.5 * (st[9] + sigEq) * d;
  out[9] = sigEq;
  out[10] = st[10];
  if (D) {
    const double thetaBar = 2 * G * 1.5 * dDdSig - (1 - theta);
    const double c1 = 2 * G * theta, c2 = 2 * G * thetaBar;
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < 6; j++) {
        double v = -c2 * nv[i] * nv[j];
        if (i < 3 && j < 3) v += K + c1 * ((i == j ? 1.0 : 0.0) - 1.0 / 3.0);
        else if (i == j) v += 0.5 * c1;
        D[i * 6 + j] = v;
      }
  }
  return d;
}

// ---------------------------------------------------------------------------
// B-bar (mean dilatation) strain-displacement matrix of a box element at a
// point (bbarMatrix, src/core/13-fatigue.js:474-490): 6 x 24 row-major, rows
// xx yy zz xy yz zx (engineering shear), columns 3a + {x, y, z} for corners
// a = corner_of(li, lj, lk); the normal rows take the volumetric part of the
// element centre's derivatives.
// ---------------------------------------------------------------------------
FE_HD void box_corner_derivs(double hx, double hy, double hz, double xi, double eta, double zeta, double dx[8],
                             double dy[8], double dz[8]) {
  const double sx = 2 / hx, sy = 2 / hy, sz = 2 / hz;
  for (int a = 0; a < 8; a++) {
    const double xa = 2 * corner_li(a) - 1, ya = 2 * corner_lj(a) - 1, za = 2 * corner_lk(a) - 1;
    dx[a] = xa * (1 + ya * eta) * (1 + za * zeta) * 0.125 * sx;
    dy[a] = ya * (1 + xa * xi) * (1 + za * zeta) * 0.125 * sy;
    dz[a] = za * (1 + xa * xi) * (1 + ya * eta) * 0.125 * sz;
  }
}

FE_HD void bbar_matrix(double hx, double hy, double hz, double xi, double eta, double zeta, double *B) {
  FE_NO_CONTRACT
  double dx[8], dy[8], dz[8], cx[8], cy[8], cz[8];
  box_corner_derivs(hx, hy, hz, xi, eta, zeta, dx, dy, dz);
  box_corner_derivs(hx, hy, hz, 0, 0, 0, cx, cy, cz);
  for (int q = 0; q < 144; q++) B[q] = 0;
  for (int a = 0; a < 8; a++) {
    const int c = 3 * a;
    const double ddx = (cx[a] - dx[a]) / 3, ddy = (cy[a] - dy[a]) / 3, ddz = (cz[a] - dz[a]) / 3;
    B[c] = dx[a] + ddx; B[c + 1] = ddy; B[c + 2] = ddz;
    B[24 + c] = ddx; B[24 + c + 1] = dy[a] + ddy; B[24 + c + 2] = ddz;
    B[48 + c] = ddx; B[48 + c + 1] = ddy; B[48 + c + 2] = dz[a] + ddz;
    B[72 + c] = dy[a]; B[72 + c + 1] = dx[a];
    B[96 + c + 1] = dz[a]; B[96 + c + 2] = dy[a];
    B[120 + c] = dz[a]; B[120 + c + 2] = dx[a];
  }
}

// ---------------------------------------------------------------------------
// Bump screening (screenJoints, src/core/13-fatigue.js:375-455): each joint a
// free-standing solder column whose axial and transverse shear strains follow
// the bump layer while its in-plane stresses vanish (AnandPoint.stepColumn,
// :153-166), or a solder cap on an elastic pillar in series
// (stepSeriesColumn, :178-206), stepped through the cycle history; per joint
// the last cycle's inelastic work density dW and equivalent strain range dE,
// and the change of dW from the cycle before relative to dW (stab).
// ---------------------------------------------------------------------------
struct ScreenSteps {
  int32_t nsteps, cycles;
  const double *T, *dt, *E, *Fth, *wt, *Cp, *Fp;  // Cp, Fp: NULL without a pillar
  const int32_t *wi, *cycle;
  const uint8_t *same;
  double fs;
};

constexpr int COLUMN_NEWTON_MAXIT = 20;     // src/core/00-constants.js:121
constexpr double COLUMN_STRESS_TOL = 1e-3;  // src/core/00-constants.js:122

// 3x3 inverse by cofactors (inv3, src/core/07-element.js:146-153)
FE_HD void inv3_cof(const double *J, double *o) {
  FE_NO_CONTRACT
  const double a = J[0], b = J[1], c = J[2], d = J[3], e = J[4], f = J[5], g = J[6], h = J[7], i = J[8];
  const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
  o[0] = (e * i - f * h) / det; o[1] = (c * h - b * i) / det; o[2] = (b * f - c * e) / det;
  o[3] = (f * g - d * i) / det; o[4] = (a * i - c * g) / det; o[5] = (c * d - a * f) / det;
  o[6] = (d * h - e * g) / det; o[7] = (b * g - a * h) / det; o[8] = (a * e - b * d) / det;
}

// Newton on the in-plane strains (xx, yy, xy) with the consistent tangent
// until the in-plane stresses are below 1e-3 of the axial/shear stress
FE_HD void step_column(const AnandParams &P, double E, double T, double dt, double eps[6], double *st, double *tr) {
  FE_NO_CONTRACT
  const int ip[3] = {0, 1, 3};
  double sig[6], D[36], A[9], Ai[9];
  for (int it = 0; it < COLUMN_NEWTON_MAXIT; it++) {
    anand_step(P, E, T, dt, eps, st, tr, sig, D);
    const double r0 = sig[0], r1 = sig[1], r2 = sig[3];
    const double ref = fmax(1.0, fmax(fabs(sig[2]), fmax(fabs(sig[4]), fabs(sig[5]))));
    if (fmax(fabs(r0), fmax(fabs(r1), fabs(r2))) < COLUMN_STRESS_TOL * ref) break;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) A[3 * r + c] = D[6 * ip[r] + ip[c]];
    inv3_cof(A, Ai);
    for (int r = 0; r < 3; r++) eps[ip[r]] -= Ai[3 * r] * r0 + Ai[3 * r + 1] * r1 + Ai[3 * r + 2] * r2;
  }
}

// cap in series with a pillar of compliances Cp: the column-average strains tz
// are prescribed (fs eps_Z + (1 - fs) Cp sigma_Z = tz) and the cap's in-plane
// stresses vanish; Newton on the six cap strains, Gaussian elimination with
// partial pivoting
FE_HD void step_series_column(const AnandParams &P, double E, double T, double dt, double eps[6], const double tz[3],
                              double fs, const double *Cp, double *st, double *tr) {
  FE_NO_CONTRACT
  const int ip[3] = {0, 1, 3}, iz[3] = {2, 4, 5}, ord[6] = {0, 1, 3, 2, 4, 5};
  const double fp = 1 - fs;
  double sig[6], D[36], J[36], R[6];
  for (int it = 0; it < COLUMN_NEWTON_MAXIT; it++) {
    anand_step(P, E, T, dt, eps, st, tr, sig, D);
    for (int r = 0; r < 3; r++) {
      R[r] = sig[ip[r]];
      R[3 + r] = fs * eps[iz[r]] + fp * Cp[r] * sig[iz[r]] - tz[r];
    }
    const double ref = fmax(1.0, fmax(fabs(sig[2]), fmax(fabs(sig[4]), fabs(sig[5]))));
    double worst = 0;
    for (int r = 0; r < 3; r++) worst = fmax(worst, fmax(fabs(R[r]), fabs(R[3 + r]) * E));
    if (worst < COLUMN_STRESS_TOL * ref) break;
    for (int r = 0; r < 6; r++)
      for (int c = 0; c < 6; c++) {
        const double Dv = D[6 * ord[r] + ord[c]];
        J[6 * r + c] = r < 3 ? Dv : fp * Cp[r - 3] * Dv + (ord[r] == ord[c] ? fs : 0.0);
      }
    for (int c = 0; c < 6; c++) {
      int pr = c;
      for (int r = c + 1; r < 6; r++)
        if (fabs(J[6 * r + c]) > fabs(J[6 * pr + c])) pr = r;
      if (pr != c) {
        for (int q = 0; q < 6; q++) {
          const double t = J[6 * c + q];
          J[6 * c + q] = J[6 * pr + q];
          J[6 * pr + q] = t;
        }
        const double t = R[c];
        R[c] = R[pr];
        R[pr] = t;
      }
      const double piv = J[7 * c];
      for (int r = c + 1; r < 6; r++) {
        const double f = J[6 * r + c] / piv;
        if (f == 0) continue;
        for (int q = c; q < 6; q++) J[6 * r + q] -= f * J[6 * c + q];
        R[r] -= f * R[c];
      }
    }
    for (int c = 5; c >= 0; c--) {
      double v = R[c];
      for (int q = c + 1; q < 6; q++) v -= J[6 * c + q] * R[q];
      R[c] = v / J[7 * c];
    }
    for (int c = 0; c < 6; c++) eps[ord[c]] -= R[c];
  }
}

// One joint through the cycle history: base = its strain-table rows (ntemps x 6,
// NaN when the site has no bump element), i25 the row of 25 C (-1: none).
// out = {dW, dE, stab} (NaN for a site without data; stab NaN below 2 cycles).
FE_HD void screen_site(const AnandParams &P, const ScreenSteps &h, int32_t i25, const double *base, double out[3]) {
  FE_NO_CONTRACT
  const double nan = __builtin_nan("");
  out[0] = out[1] = out[2] = nan;
  if (!isfinite(base[0])) return;
  const int ip[3] = {0, 1, 3};
  const int nc = h.cycles;
  double st[11] = {0, 0, 0, 0, 0, 0, P.s0, 0, 0, 0, 0}, tr[11];
  double eps[6] = {0, 0, 0, 0, 0, 0}, e25[6], e1[3] = {0, 0, 0}, e2[3] = {0, 0, 0};
  double s1[6] = {0, 0, 0, 0, 0, 0}, s2[6] = {0, 0, 0, 0, 0, 0};
  double W3[3] = {0, 0, 0}, E3[3] = {0, 0, 0};  // W and ebar at the end of cycles nc-2, nc-1, nc
  for (int c = 0; c < 6; c++) e25[c] = i25 >= 0 ? base[6 * i25 + c] : 0.0;
  for (int k = 1; k < h.nsteps; k++) {
    const double *a = base + 6 * h.wi[k], *b = a + 6;
    const double t = h.wt[k];
    double lay[6];
    for (int c = 0; c < 6; c++) lay[c] = a[c] + t * (b[c] - a[c]) - e25[c];
    const bool ext = k > 2 && h.same[k];
    if (h.Cp) {
      const double tz[3] = {lay[2] - h.fs * h.Fth[k] - (1 - h.fs) * h.Fp[k], lay[4], lay[5]};
      if (ext)
        for (int c = 0; c < 6; c++) eps[c] = 2 * s1[c] - s2[c];
      step_series_column(P, h.E[k], h.T[k], h.dt[k], eps, tz, h.fs, h.Cp + 3 * k, st, tr);
      for (int c = 0; c < 6; c++) { s2[c] = s1[c]; s1[c] = eps[c]; }
    } else {
      const double ez = lay[2] - h.Fth[k];
      if (ext) {
        for (int r = 0; r < 3; r++) eps[ip[r]] = 2 * e1[r] - e2[r];
      } else {
        // the last strains plus the elastic lateral contraction of the axial increment
        eps[0] -= P.nu * (ez - eps[2]);
        eps[1] -= P.nu * (ez - eps[2]);
      }
      eps[2] = ez; eps[4] = lay[4]; eps[5] = lay[5];
      step_column(P, h.E[k], h.T[k], h.dt[k], eps, st, tr);
    }
    for (int c = 0; c < 10; c++) st[c] = tr[c];  // commit (st[10], the Newton warm start, is the step's)
    for (int r = 0; r < 3; r++) { e2[r] = e1[r]; e1[r] = eps[ip[r]]; }
    const int slot = h.cycle[k] - (nc - 2);
    if (slot >= 0 && slot < 3) { W3[slot] = st[8]; E3[slot] = st[7]; }
  }
  out[0] = W3[2] - W3[1];
  out[1] = E3[2] - E3[1];
  out[2] = nc >= 2 ? fabs((W3[2] - W3[1]) - (W3[1] - W3[0])) / fmax(W3[2] - W3[1], 1e-30) : nan;
}

// ---------------------------------------------------------------------------
// Operator rows of one z-column of nodes from a staged neighbourhood.
// nb holds w*x of the 3 x 3 node columns around column (i, j) at levels
// -1 .. nz (zero padded), layout nb[(col * nzp + k + 1) * 3 + p] (nzp >= nz + 2) with
// col = 3 * (di + 1) + (dj + 1).  acc = sum_j K_ij (w x)_j for node k
// (unscaled row; the caller applies w_i and the identity rows).
// ---------------------------------------------------------------------------
constexpr int NZP = 66;  // largest stride: 64 planes + 2 padding levels

// PRESCALED: x already holds w*x (w unused)
template <typename T, bool PRESCALED = false>
FE_HD void fill_neighbourhood(const Grid &g, int i, int j, int k, const T *__restrict__ w, const T *__restrict__ x,
                              T *nb, int nzp) {
  // every load is issued unconditionally (columns outside the grid read the
  // centre column and are zeroed afterwards), so that their latencies overlap
  T v[9][3];
  bool in[9];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int q = 0; q < 9; q++) {
    const int ii = i + q / 3 - 1, jj = j + q % 3 - 1;
    in[q] = ii >= 0 && ii < g.nx && jj >= 0 && jj < g.ny;
    const int gn = ((in[q] ? ii * g.ny + jj : i * g.ny + j) * g.nz + k) * 3;
    for (int p = 0; p < 3; p++) v[q][p] = PRESCALED ? x[gn + p] : w[gn + p] * x[gn + p];
  }
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int q = 0; q < 9; q++) {
    T *o = nb + (q * nzp + k + 1) * 3;
    for (int p = 0; p < 3; p++) o[p] = in[q] ? v[q][p] : T(0);
  }
}

template <typename T>
FE_HD void node_rows_nb(const Grid &g, int i, int j, int k, const int32_t *__restrict__ cellType,
                        const T *__restrict__ cellScale, const T *__restrict__ Kc, const T *nb, T acc[3], int nzp) {
  T a0 = 0, a1 = 0, a2 = 0;
  for (int ci = 0; ci < 2; ci++) {
    const int ii = i - 1 + ci;
    if (ii < 0 || ii >= g.ncx) continue;
    for (int cj = 0; cj < 2; cj++) {
      const int jj = j - 1 + cj;
      if (jj < 0 || jj >= g.ncy) continue;
      for (int ck = 0; ck < 2; ck++) {
        const int kk = k - 1 + ck;
        if (kk < 0 || kk >= g.ncz) continue;
        const int c = (ii * g.ncy + jj) * g.ncz + kk;
        const int a = corner_of(1 - ci, 1 - cj, 1 - ck);
        for (int q = 0; q < 2; q++) {
          const int32_t t = cellType[2 * c + q];
          if (t < 0) continue;
          const T sc = cellScale[2 * c + q];
          const T *K = Kc + t * 576 + a * 72;
          T s0 = 0, s1 = 0, s2 = 0;
          for (int b = 0; b < 8; b++) {
            const int col = (ci + corner_li(b)) * 3 + (cj + corner_lj(b));
            const T *xv = nb + (col * nzp + k + ck + corner_lk(b)) * 3;
            s0 += K[3 * b] * xv[0] + K[3 * b + 1] * xv[1] + K[3 * b + 2] * xv[2];
            s1 += K[24 + 3 * b] * xv[0] + K[24 + 3 * b + 1] * xv[1] + K[24 + 3 * b + 2] * xv[2];
            s2 += K[48 + 3 * b] * xv[0] + K[48 + 3 * b + 1] * xv[1] + K[48 + 3 * b + 2] * xv[2];
          }
          a0 += sc * s0; a1 += sc * s1; a2 += sc * s2;
        }
      }
    }
  }
  acc[0] = a0; acc[1] = a1; acc[2] = a2;
}

// ---------------------------------------------------------------------------
// Reflection-symmetric element tables.  A box element whose material has no
// normal/shear and no shear/shear coupling is invariant under the mirror
// images through its mid-planes, so with mirror mask m (bits x, y, z) and S_m
// the matching sign flips of the displacement components,
//   K_ab = S_a K_{0, a^b} S_a       (a^b: the corner with bits(a) xor bits(b)).
// The corner-0 rows R_c = K_{0,c} (3 x 24) then carry the whole element, and
// each 3 x 3 block R_c satisfies R_c[q][p] = s_c[p] s_c[q] R_c[p][q] (s_c the
// sign flips of corner c), so its 6 entries on and above the diagonal carry
// it: the compact table holds, per corner c, [R_c00, R_c11, R_c22, R_c01,
// R_c02, R_c12] (SYM_TABLE = 48 values per element type).  cell_sym_pair
// forms, for one cell, the rows of corner (MX, MY, 0) ("own") and of corner
// (MX, MY, 1) ("up") applied to the cell's corner vectors xb, scaled by sc.
constexpr int SYM_TABLE = 48;

template <typename T, int MX, int MY>
FE_HD void cell_sym_pair(const T *__restrict__ R, const T (&xb)[8][3], T sc, T own[3], T up[3]) {
  const T sx = MX ? T(-1) : T(1), sy = MY ? T(-1) : T(1);
  // corner by corner (each corner's 6 table values serve the three rows);
  // every row still sums its terms in corner order, one multiply-add per
  // table entry (the sign flips are exact)
  T zo[3] = {0, 0, 0}, zu[3] = {0, 0, 0};
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int c = 0; c < 8; c++) {
    const int co = corner_of(corner_li(c) ^ MX, corner_lj(c) ^ MY, corner_lk(c));
    const int cu = corner_of(corner_li(c) ^ MX, corner_lj(c) ^ MY, corner_lk(c) ^ 1);
    const T *Rc = R + 6 * c;
    const T s01 = (corner_li(c) ^ corner_lj(c)) ? T(-1) : T(1);
    const T s02 = (corner_li(c) ^ corner_lk(c)) ? T(-1) : T(1);
    const T s12 = (corner_lj(c) ^ corner_lk(c)) ? T(-1) : T(1);
    const T d0 = Rc[0], d1 = Rc[1], d2 = Rc[2], o01 = Rc[3], o02 = Rc[4], o12 = Rc[5];
    const T rows[3][3] = {{d0, o01, o02}, {s01 * o01, d1, o12}, {s02 * o02, s12 * o12, d2}};
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int p = 0; p < 3; p++) {
      zo[p] += rows[p][0] * (sx * xb[co][0]);
      zo[p] += rows[p][1] * (sy * xb[co][1]);
      zo[p] += rows[p][2] * xb[co][2];
      zu[p] += rows[p][0] * (sx * xb[cu][0]);
      zu[p] += rows[p][1] * (sy * xb[cu][1]);
      zu[p] -= rows[p][2] * xb[cu][2];
    }
  }
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int p = 0; p < 3; p++) {
    const T sp = p == 0 ? sx : p == 1 ? sy : T(1);
    own[p] += sc * sp * zo[p];
    up[p] += sc * (p == 2 ? -sp : sp) * zu[p];
  }
}

// The same pair of rows from a full 24 x 24 table (an element type without
// the mirror symmetry: the viscoplastic solder's own tangents)
template <typename T, int MX, int MY>
FE_HD void cell_full_pair(const T *__restrict__ K, const T (&xb)[8][3], T sc, T own[3], T up[3]) {
  constexpr int ao = MX ? (MY ? 2 : 1) : (MY ? 3 : 0), au = ao + 4;  // corners (MX, MY, 0) and (MX, MY, 1)
  for (int p = 0; p < 3; p++) {
    const T *Ko = K + (3 * ao + p) * 24, *Ku = K + (3 * au + p) * 24;
    T so = 0, su = 0;
    for (int b = 0; b < 8; b++)
      for (int q = 0; q < 3; q++) {
        so += Ko[3 * b + q] * xb[b][q];
        su += Ku[3 * b + q] * xb[b][q];
      }
    own[p] += sc * so;
    up[p] += sc * su;
  }
}

// Element tables of the compact row forms: R (SYM_TABLE per type) for the
// types below own0, the full tables K (576 per type) from own0 on
template <typename T>
struct SymTabs {
  const T *R;
  const T *K;
  int own0;
};

// The four cells of level k around node column (i, j) (quadrant q = 2 CI +
// CJ, cell (i - 1 + CI, j - 1 + CJ, k)): both element slots' types and scales,
// -1 where the cell is outside the grid (or k is the top plane).  The loads
// are issued together, with a clamped index for absent cells, so that a
// wavefront waits on them once instead of once per cell.
template <typename T>
struct CellQuad {
  int32_t t[4][2];
  T sc[4][2];
};

template <typename T>
FE_HD void load_cell_quad(const Grid &g, int i, int j, int k, bool valid, const int32_t *__restrict__ cellType,
                          const T *__restrict__ cellScale, CellQuad<T> &cq) {
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int q = 0; q < 4; q++) {
    const int ii = i - 1 + (q >> 1), jj = j - 1 + (q & 1);
    const bool in = valid && k < g.ncz && ii >= 0 && ii < g.ncx && jj >= 0 && jj < g.ncy;
    const int64_t c = in ? ((int64_t)ii * g.ncy + jj) * g.ncz + k : 0;
    const int32_t t0 = cellType[2 * c], t1 = cellType[2 * c + 1];
    const T s0 = cellScale[2 * c], s1 = cellScale[2 * c + 1];
    cq.t[q][0] = in ? t0 : -1;
    cq.t[q][1] = in ? t1 : -1;
    cq.sc[q][0] = s0;
    cq.sc[q][1] = s1;
  }
}

// Column form on the LDS neighbourhood: for node (i, j, k) the cell of
// quadrant (CI, CJ) at level k (above the node).  own = its rows of node
// (i, j, k), up = its rows of node (i, j, k+1); the node's full row is own_k +
// up_{k-1}.  HYB: some types (from tb.own0 on) take the full tables.
template <typename T, int CI, int CJ, bool HYB>
FE_HD void cell_sym_nb(int k, const CellQuad<T> &cq, const SymTabs<T> &tb, const T *nb, T own[3], T up[3], int nzp) {
  constexpr int q4 = 2 * CI + CJ;
  for (int q = 0; q < 2; q++) {
    const int32_t t = cq.t[q4][q];
    if (t < 0) continue;
    const T sc = cq.sc[q4][q];
    T xb[8][3];
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
    for (int b = 0; b < 8; b++) {
      const int col = (CI + corner_li(b)) * 3 + (CJ + corner_lj(b));
      const T *xv = nb + (col * nzp + k + corner_lk(b) + 1) * 3;
      xb[b][0] = xv[0]; xb[b][1] = xv[1]; xb[b][2] = xv[2];
    }
    if (HYB && t >= tb.own0)
      cell_full_pair<T, 1 - CI, 1 - CJ>(tb.K + 576 * (int64_t)t, xb, sc, own, up);
    else
      cell_sym_pair<T, 1 - CI, 1 - CJ>(tb.R + SYM_TABLE * t, xb, sc, own, up);
  }
}

// own and up rows of node (i, j, k) from the four cells of level k (cq from
// load_cell_quad)
template <typename T, bool HYB = false>
FE_HD void node_rows_nb_sym_q(int k, const CellQuad<T> &cq, const SymTabs<T> &tb, const T *nb, T own[3], T up[3],
                              int nzp) {
  for (int p = 0; p < 3; p++) { own[p] = 0; up[p] = 0; }
  cell_sym_nb<T, 0, 0, HYB>(k, cq, tb, nb, own, up, nzp);
  cell_sym_nb<T, 0, 1, HYB>(k, cq, tb, nb, own, up, nzp);
  cell_sym_nb<T, 1, 0, HYB>(k, cq, tb, nb, own, up, nzp);
  cell_sym_nb<T, 1, 1, HYB>(k, cq, tb, nb, own, up, nzp);
}

// true when the four cells hold one element each, of one compact-table type
// (scales may differ), or no element at all: their rows then share one table
template <typename T, bool HYB = false>
FE_HD bool cell_quad_uniform(const CellQuad<T> &cq, const SymTabs<T> &tb) {
  const int32_t t = cq.t[0][0];
  bool u = !(HYB && t >= tb.own0);
  for (int q = 0; q < 4; q++) u = u && cq.t[q][0] == t && cq.t[q][1] < 0;
  return u;
}

// node_rows_nb_sym_q for a uniform quad (cell_quad_uniform): the compact
// table is loaded once for the four cells; the rows are summed in the same
// order with the same terms
// the compact table of the quad's type (zeros for an empty quad)
template <typename T>
FE_HD void load_quad_table(const CellQuad<T> &cq, const SymTabs<T> &tb, T R[SYM_TABLE]) {
  const int32_t t = cq.t[0][0];
  const T *Rg = tb.R + SYM_TABLE * (t < 0 ? 0 : t);
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int e = 0; e < SYM_TABLE; e++) R[e] = Rg[e];
}

// rows of a uniform quad from its table R (load_quad_table)
template <typename T>
FE_HD void node_rows_nb_sym_uni_R(int k, const CellQuad<T> &cq, const T R[SYM_TABLE], const T *nb, T own[3], T up[3],
                                  int nzp) {
  for (int p = 0; p < 3; p++) { own[p] = 0; up[p] = 0; }
  if (cq.t[0][0] < 0) return;
  T xb[8][3];
#define FE_UNI_QUAD(CI, CJ)                                                      \
  for (int b = 0; b < 8; b++) {                                                  \
    const int col = (CI + corner_li(b)) * 3 + (CJ + corner_lj(b));               \
    const T *xv = nb + (col * nzp + k + corner_lk(b) + 1) * 3;                   \
    xb[b][0] = xv[0]; xb[b][1] = xv[1]; xb[b][2] = xv[2];                        \
  }                                                                              \
  cell_sym_pair<T, 1 - CI, 1 - CJ>(R, xb, cq.sc[2 * CI + CJ][0], own, up);
  FE_UNI_QUAD(0, 0)
  FE_UNI_QUAD(0, 1)
  FE_UNI_QUAD(1, 0)
  FE_UNI_QUAD(1, 1)
#undef FE_UNI_QUAD
}

template <typename T>
FE_HD void node_rows_nb_sym_uni(int k, const CellQuad<T> &cq, const SymTabs<T> &tb, const T *nb, T own[3], T up[3],
                                int nzp) {
  T R[SYM_TABLE];
  load_quad_table<T>(cq, tb, R);
  node_rows_nb_sym_uni_R<T>(k, cq, R, nb, own, up, nzp);
}

// Quads whose cells share their types in pairs: J (the two cells of one
// column of cells, i.e. equal i - 1 + CI, hold one type each: quadrants 0, 1
// and 2, 3) or I (equal j - 1 + CJ: quadrants 0, 2 and 1, 3); one element per
// cell, compact tables.  Absent cells (type -1) pair with absent cells.
template <typename T, bool HYB = false>
FE_HD int cell_quad_pairs(const CellQuad<T> &cq, const SymTabs<T> &tb) {
  bool ok = true;
  for (int q = 0; q < 4; q++) ok = ok && cq.t[q][1] < 0 && !(HYB && cq.t[q][0] >= tb.own0);
  if (!ok) return 0;
  if (cq.t[0][0] == cq.t[1][0] && cq.t[2][0] == cq.t[3][0]) return 1;  // J
  if (cq.t[0][0] == cq.t[2][0] && cq.t[1][0] == cq.t[3][0]) return 2;  // I
  return 0;
}

// rows of a paired quad (cell_quad_pairs = PAIR): two table loads instead of
// four.  J keeps the quadrant order of node_rows_nb_sym_q; I takes the
// quadrants sharing a table together (0, 2, 1, 3), which changes the order in
// which the four cells' terms are added (rounding only).
template <typename T, int PAIR>
FE_HD void node_rows_nb_sym_pairs(int k, const CellQuad<T> &cq, const SymTabs<T> &tb, const T *nb, T own[3],
                                  T up[3], int nzp) {
  for (int p = 0; p < 3; p++) { own[p] = 0; up[p] = 0; }
  T R[SYM_TABLE];
  T xb[8][3];
#define FE_PAIR_QUAD(CI, CJ)                                                     \
  if (cq.t[2 * CI + CJ][0] >= 0) {                                               \
    for (int b = 0; b < 8; b++) {                                                \
      const int col = (CI + corner_li(b)) * 3 + (CJ + corner_lj(b));             \
      const T *xv = nb + (col * nzp + k + corner_lk(b) + 1) * 3;                 \
      xb[b][0] = xv[0]; xb[b][1] = xv[1]; xb[b][2] = xv[2];                      \
    }                                                                            \
    cell_sym_pair<T, 1 - CI, 1 - CJ>(R, xb, cq.sc[2 * CI + CJ][0], own, up);     \
  }
#define FE_PAIR_TABLE(Q)                                                         \
  {                                                                              \
    const int32_t t = cq.t[Q][0];                                                \
    const T *Rg = tb.R + SYM_TABLE * (t < 0 ? 0 : t);                            \
    for (int e = 0; e < SYM_TABLE; e++) R[e] = Rg[e];                            \
  }
  if (PAIR == 1) {
    FE_PAIR_TABLE(0)
    FE_PAIR_QUAD(0, 0)
    FE_PAIR_QUAD(0, 1)
    FE_PAIR_TABLE(2)
    FE_PAIR_QUAD(1, 0)
    FE_PAIR_QUAD(1, 1)
  } else {
    FE_PAIR_TABLE(0)
    FE_PAIR_QUAD(0, 0)
    FE_PAIR_QUAD(1, 0)
    FE_PAIR_TABLE(1)
    FE_PAIR_QUAD(0, 1)
    FE_PAIR_QUAD(1, 1)
  }
#undef FE_PAIR_QUAD
#undef FE_PAIR_TABLE
}

template <typename T, bool HYB = false>
FE_HD void node_rows_nb_sym(const Grid &g, int i, int j, int k, const int32_t *__restrict__ cellType,
                            const T *__restrict__ cellScale, const SymTabs<T> &tb, const T *nb, T own[3],
                            T up[3], int nzp) {
  CellQuad<T> cq;
  load_cell_quad<T>(g, i, j, k, true, cellType, cellScale, cq);
  const int pr = cell_quad_pairs<T, HYB>(cq, tb);
  if (cell_quad_uniform<T, HYB>(cq, tb))
    node_rows_nb_sym_uni<T>(k, cq, tb, nb, own, up, nzp);
  else if (pr == 1)
    node_rows_nb_sym_pairs<T, 1>(k, cq, tb, nb, own, up, nzp);
  else if (pr == 2)
    node_rows_nb_sym_pairs<T, 2>(k, cq, tb, nb, own, up, nzp);
  else
    node_rows_nb_sym_q<T, HYB>(k, cq, tb, nb, own, up, nzp);
}

// Host form of one node row through the symmetric tables (same output rule as
// node_apply: w * (A w x) on free DOFs, x on constrained ones).
template <typename T, int CI, int CJ, bool PRESCALED = false>
FE_HD void cell_sym_direct(const Grid &g, int i, int j, int kk, bool wantUp, const int32_t *__restrict__ cellType,
                           const T *__restrict__ cellScale, const SymTabs<T> &tb, const T *__restrict__ w,
                           const T *__restrict__ x, T acc[3]) {
  const int ii = i - 1 + CI, jj = j - 1 + CJ;
  if (ii < 0 || ii >= g.ncx || jj < 0 || jj >= g.ncy || kk < 0 || kk >= g.ncz) return;
  const int64_t c = ((int64_t)ii * g.ncy + jj) * g.ncz + kk;
  for (int q = 0; q < 2; q++) {
    const int32_t t = cellType[2 * c + q];
    if (t < 0) continue;
    const T sc = cellScale[2 * c + q];
    T xb[8][3];
    for (int b = 0; b < 8; b++) {
      const int64_t gn = ((int64_t)(ii + corner_li(b)) * g.ny + jj + corner_lj(b)) * g.nz + kk + corner_lk(b);
      for (int p = 0; p < 3; p++) xb[b][p] = PRESCALED ? x[3 * gn + p] : w[3 * gn + p] * x[3 * gn + p];
    }
    T own[3] = {0, 0, 0}, up[3] = {0, 0, 0};
    if (t >= tb.own0)
      cell_full_pair<T, 1 - CI, 1 - CJ>(tb.K + 576 * (int64_t)t, xb, sc, own, up);
    else
      cell_sym_pair<T, 1 - CI, 1 - CJ>(tb.R + SYM_TABLE * t, xb, sc, own, up);
    for (int p = 0; p < 3; p++) acc[p] += wantUp ? up[p] : own[p];
  }
}

template <typename T, bool PRESCALED = false>
FE_HD void node_apply_sym(const Grid &g, int i, int j, int k, const int32_t *__restrict__ cellType,
                          const T *__restrict__ cellScale, const SymTabs<T> &R, const T *__restrict__ w,
                          const T *__restrict__ x, T out[3]) {
  T own[3] = {0, 0, 0}, up[3] = {0, 0, 0};
  cell_sym_direct<T, 0, 0, PRESCALED>(g, i, j, k, false, cellType, cellScale, R, w, x, own);
  cell_sym_direct<T, 0, 1, PRESCALED>(g, i, j, k, false, cellType, cellScale, R, w, x, own);
  cell_sym_direct<T, 1, 0, PRESCALED>(g, i, j, k, false, cellType, cellScale, R, w, x, own);
  cell_sym_direct<T, 1, 1, PRESCALED>(g, i, j, k, false, cellType, cellScale, R, w, x, own);
  cell_sym_direct<T, 0, 0, PRESCALED>(g, i, j, k - 1, true, cellType, cellScale, R, w, x, up);
  cell_sym_direct<T, 0, 1, PRESCALED>(g, i, j, k - 1, true, cellType, cellScale, R, w, x, up);
  cell_sym_direct<T, 1, 0, PRESCALED>(g, i, j, k - 1, true, cellType, cellScale, R, w, x, up);
  cell_sym_direct<T, 1, 1, PRESCALED>(g, i, j, k - 1, true, cellType, cellScale, R, w, x, up);
  const int64_t gi = ((int64_t)i * g.ny + j) * g.nz + k;
  for (int p = 0; p < 3; p++) {
    const T wi = w[3 * gi + p];
    out[p] = wi != T(0) ? wi * (own[p] + up[p]) : (PRESCALED ? T(0) : x[3 * gi + p]);
  }
}

// Corner-0 rows of one full table (576), made to satisfy the mirror relation
// R_c[q][p] = s_c[p] s_c[q] R_c[p][q] exactly (the operator stays exactly
// symmetric), and the largest deviation of the full table from the mirror
// form, relative to its largest entry.
FE_HD double sym_table(const double *__restrict__ K, double *__restrict__ R) {
  for (int c = 0; c < 8; c++) {
    const double s[3] = {corner_li(c) ? -1.0 : 1.0, corner_lj(c) ? -1.0 : 1.0, corner_lk(c) ? -1.0 : 1.0};
    for (int p = 0; p < 3; p++) {
      R[24 * p + 3 * c + p] = K[24 * p + 3 * c + p];
      for (int q = p + 1; q < 3; q++) {
        const double v = 0.5 * (K[24 * p + 3 * c + q] + s[p] * s[q] * K[24 * q + 3 * c + p]);
        R[24 * p + 3 * c + q] = v;
        R[24 * q + 3 * c + p] = s[p] * s[q] * v;
      }
    }
  }
  double kmax = 0, err = 0;
  for (int e = 0; e < 576; e++) kmax = fmax(kmax, fabs(K[e]));
  for (int a = 0; a < 8; a++) {
    const double sa[3] = {corner_li(a) ? -1.0 : 1.0, corner_lj(a) ? -1.0 : 1.0, corner_lk(a) ? -1.0 : 1.0};
    for (int b = 0; b < 8; b++) {
      const int c = corner_of(corner_li(a) ^ corner_li(b), corner_lj(a) ^ corner_lj(b), corner_lk(a) ^ corner_lk(b));
      for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) {
          const double pred = sa[p] * sa[q] * R[24 * p + 3 * c + q];
          err = fmax(err, fabs(K[72 * a + 24 * p + 3 * b + q] - pred));
        }
    }
  }
  return kmax > 0 ? err / kmax : 0.0;
}

// compact table (SYM_TABLE values) of the corner-0 rows R (72, [p][3c+q])
FE_HD void sym_compact(const double *__restrict__ R, double *__restrict__ Rs) {
  for (int c = 0; c < 8; c++) {
    double *o = Rs + 6 * c;
    o[0] = R[3 * c]; o[1] = R[24 + 3 * c + 1]; o[2] = R[48 + 3 * c + 2];
    o[3] = R[3 * c + 1]; o[4] = R[3 * c + 2]; o[5] = R[24 + 3 * c + 2];
  }
}

// ---------------------------------------------------------------------------
// Stress histories of homogeneous box elements.  State (STATE_SIZE doubles):
// q (33: 24 nodal + 9 incompatible-mode DOFs), then per strain monomial
// phi = [1, xi, eta, zeta, xi eta, eta zeta, xi zeta] the stress S_m (6),
// then the thermal term S_th (6).
// ---------------------------------------------------------------------------
constexpr int STATE_SIZE = 33 + 6 * 8;

constexpr int HEX_XI_[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
constexpr int HEX_ETA_[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
constexpr int HEX_ZETA_[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

// strain of the box element as a polynomial in the natural coordinates
// (boxStrainPoly, src/core/09-assembly.js:167-195): E[6 m + r]
FE_HD void box_strain_poly(double hx, double hy, double hz, const double *q, double E[42]) {
  for (int i = 0; i < 42; i++) E[i] = 0;
  const double sx = 2 / hx, sy = 2 / hy, sz = 2 / hz;
  for (int a = 0; a < 8; a++) {
    const double xa = HEX_XI_[a], ya = HEX_ETA_[a], za = HEX_ZETA_[a];
    const double ux = q[3 * a], uy = q[3 * a + 1], uz = q[3 * a + 2];
    const double cx = xa * sx / 8, cy = ya * sy / 8, cz = za * sz / 8;
    // d/dx: monomials 1, eta, zeta, eta zeta; d/dy: 1, xi, zeta, xi zeta; d/dz: 1, xi, eta, xi eta
    const int mx[4] = {0, 2, 3, 5}, my[4] = {0, 1, 3, 6}, mz[4] = {0, 1, 2, 4};
    const double vx[4] = {cx, cx * ya, cx * za, cx * ya * za};
    const double vy[4] = {cy, cy * xa, cy * za, cy * xa * za};
    const double vz[4] = {cz, cz * xa, cz * ya, cz * xa * ya};
    for (int t = 0; t < 4; t++) {
      E[6 * mx[t]] += vx[t] * ux; E[6 * mx[t] + 3] += vx[t] * uy; E[6 * mx[t] + 5] += vx[t] * uz;
    }
    for (int t = 0; t < 4; t++) {
      E[6 * my[t] + 1] += vy[t] * uy; E[6 * my[t] + 3] += vy[t] * ux; E[6 * my[t] + 4] += vy[t] * uz;
    }
    for (int t = 0; t < 4; t++) {
      E[6 * mz[t] + 2] += vz[t] * uz; E[6 * mz[t] + 4] += vz[t] * uy; E[6 * mz[t] + 5] += vz[t] * ux;
    }
  }
  const double *a1 = q + 24, *a2 = q + 27, *a3 = q + 30;
  E[6] += -2 * sx * a1[0]; E[9] += -2 * sx * a1[1]; E[11] += -2 * sx * a1[2];
  E[13] += -2 * sy * a2[1]; E[15] += -2 * sy * a2[0]; E[16] += -2 * sy * a2[2];
  E[20] += -2 * sz * a3[2]; E[22] += -2 * sz * a3[1]; E[23] += -2 * sz * a3[0];
}

// strain at (xi, eta, zeta) from the 33 element DOFs (boxDerivs + boxStrainAt,
// src/core/07-element.js:19-31, 127-142)
FE_HD void box_strain_at(double hx, double hy, double hz, const double *q, double xi, double eta, double zeta,
                         double eps[6]) {
  const double sx = 2 / hx, sy = 2 / hy, sz = 2 / hz;
  double dx[11], dy[11], dz[11];
  for (int a = 0; a < 8; a++) {
    const double xa = HEX_XI_[a], ya = HEX_ETA_[a], za = HEX_ZETA_[a];
    dx[a] = xa * (1 + ya * eta) * (1 + za * zeta) * 0.125 * sx;
    dy[a] = ya * (1 + xa * xi) * (1 + za * zeta) * 0.125 * sy;
    dz[a] = za * (1 + xa * xi) * (1 + ya * eta) * 0.125 * sz;
  }
  dx[8] = -2 * xi * sx; dy[8] = 0; dz[8] = 0;
  dx[9] = 0; dy[9] = -2 * eta * sy; dz[9] = 0;
  dx[10] = 0; dy[10] = 0; dz[10] = -2 * zeta * sz;
  for (int r = 0; r < 6; r++) eps[r] = 0;
  for (int b = 0; b < 11; b++) {
    const double ux = q[3 * b], uy = q[3 * b + 1], uz = q[3 * b + 2];
    eps[0] += dx[b] * ux; eps[1] += dy[b] * uy; eps[2] += dz[b] * uz;
    eps[3] += dy[b] * ux + dx[b] * uy;
    eps[4] += dz[b] * uy + dy[b] * uz;
    eps[5] += dz[b] * ux + dx[b] * uz;
  }
}

// One step added to the state of a homogeneous element (accumulateStep,
// src/core/09-assembly.js:494-545): dq = (du at the nodes, dalpha =
// K_aa^-1 (f_a - K_au du)) with the step's shape matrices M (6 x 1089),
// q += dq; when born (and not qOnly) S_m += D E_m(dq), S_th += s0.
// du8 holds the 24 nodal increments; gH the full-box thermal integrals.
FE_HD void element_recover(double hx, double hy, double hz, const double *__restrict__ M,
                           const double *__restrict__ D, const double *__restrict__ s0, bool born, bool qOnly,
                           const double *__restrict__ gH, const double *__restrict__ du8, double *__restrict__ state) {
  const double V8 = hx * hy * hz / 8;
  const double c[6] = {V8 * 4 / (hx * hx), V8 * 4 / (hy * hy), V8 * 4 / (hz * hz),
                       V8 * 4 / (hx * hy), V8 * 4 / (hy * hz), V8 * 4 / (hz * hx)};
  double Kr[9 * 33];
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 33; j++) {
      const int o = (24 + i) * 33 + j;
      Kr[i * 33 + j] = c[0] * M[o] + c[1] * M[1089 + o] + c[2] * M[2178 + o] + c[3] * M[3267 + o] +
                       c[4] * M[4356 + o] + c[5] * M[5445 + o];
    }
  // incompatible-mode rows of the thermal load (elementFth, src/core/09-assembly.js:115-133)
  double fa[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (born) {
    const int VI[3][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}};
    const double h[3] = {hx, hy, hz};
    for (int r = 0; r < 3; r++) {
      const double cr = V8 * (2 / h[r]);
      for (int p = 0; p < 3; p++) {
        const double sv = cr * s0[VI[p][r]];
        if (sv == 0) continue;
        for (int a = 8; a < 11; a++) fa[3 * (a - 8) + p] += sv * gH[r * 11 + a];
      }
    }
  }
  double dq[33];
  for (int t = 0; t < 24; t++) dq[t] = du8[t];
  double Kaa[81];
  for (int r = 0; r < 9; r++)
    for (int s = 0; s < 9; s++) Kaa[r * 9 + s] = Kr[r * 33 + 24 + s];
  chol_dense(Kaa, 9, 1e-14);
  double col[9];
  for (int r = 0; r < 9; r++) {
    double sum = fa[r];
    for (int t = 0; t < 24; t++) sum -= Kr[r * 33 + t] * dq[t];
    col[r] = sum;
  }
  chol_solve<double, double>(Kaa, 9, col);
  for (int r = 0; r < 9; r++) dq[24 + r] = col[r];
  for (int t = 0; t < 33; t++) state[t] += dq[t];
  if (!born || qOnly) return;
  double E[42];
  box_strain_poly(hx, hy, hz, dq, E);
  for (int mo = 0; mo < 7; mo++)
    for (int r = 0; r < 6; r++) {
      double v = 0;
      for (int t = 0; t < 6; t++) v += D[r * 6 + t] * E[6 * mo + t];
      state[33 + 6 * mo + r] += v;
    }
  for (int r = 0; r < 6; r++) state[33 + 42 + r] += s0[r];
}

// stress and total strain at a natural point from a state (stateStress,
// src/core/09-assembly.js:556-571): sigma = sum_m phi_m S_m - S_th
FE_HD void state_stress(double hx, double hy, double hz, const double *__restrict__ state, double xi, double eta,
                        double zeta, double sig[6], double eps[6]) {
  box_strain_at(hx, hy, hz, state, xi, eta, zeta, eps);
  const double phi[7] = {1, xi, eta, zeta, xi * eta, eta * zeta, xi * zeta};
  for (int r = 0; r < 6; r++) {
    double v = -state[33 + 42 + r];
    for (int mo = 0; mo < 7; mo++) v += phi[mo] * state[33 + 6 * mo + r];
    sig[r] = v;
  }
}

// Adjoint of the recovered point stresses of one element: given dJ/dsigma at
// natural points (npts x 6), the derivative of J with respect to the element's
// 24 nodal displacements at fixed material data.  sigma = D B(point) q with
// q = (u, alpha), alpha = K_aa^-1 (f_a - K_au u) (accumulateStep's recovery), so
// dJ/du = w_u - K_ua K_aa^-1 w_a with w = sum_points B^T D dJ/dsigma (33).
FE_HD void element_stress_pullback(double hx, double hy, double hz, const double *__restrict__ M,
                                   const double *__restrict__ D, int npts, const double *__restrict__ xez,
                                   const double *__restrict__ dJds, double g[24]) {
  const double sx = 2 / hx, sy = 2 / hy, sz = 2 / hz;
  double w[33];
  for (int t = 0; t < 33; t++) w[t] = 0;
  for (int p = 0; p < npts; p++) {
    const double xi = xez[3 * p], eta = xez[3 * p + 1], zeta = xez[3 * p + 2];
    double v[6];
    for (int r = 0; r < 6; r++) {
      double acc = 0;
      for (int c = 0; c < 6; c++) acc += D[6 * r + c] * dJds[6 * p + c];
      v[r] = acc;
    }
    for (int b = 0; b < 11; b++) {
      double dx, dy, dz;
      if (b < 8) {
        const double xa = HEX_XI_[b], ya = HEX_ETA_[b], za = HEX_ZETA_[b];
        dx = xa * (1 + ya * eta) * (1 + za * zeta) * 0.125 * sx;
        dy = ya * (1 + xa * xi) * (1 + za * zeta) * 0.125 * sy;
        dz = za * (1 + xa * xi) * (1 + ya * eta) * 0.125 * sz;
      } else {
        dx = b == 8 ? -2 * xi * sx : 0.0;
        dy = b == 9 ? -2 * eta * sy : 0.0;
        dz = b == 10 ? -2 * zeta * sz : 0.0;
      }
      w[3 * b] += dx * v[0] + dy * v[3] + dz * v[5];
      w[3 * b + 1] += dy * v[1] + dx * v[3] + dz * v[4];
      w[3 * b + 2] += dz * v[2] + dy * v[4] + dx * v[5];
    }
  }
  const double V8 = hx * hy * hz / 8;
  const double c[6] = {V8 * 4 / (hx * hx), V8 * 4 / (hy * hy), V8 * 4 / (hz * hz),
                       V8 * 4 / (hx * hy), V8 * 4 / (hy * hz), V8 * 4 / (hz * hx)};
  double Kr[9 * 33];
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 33; j++) {
      const int o = (24 + i) * 33 + j;
      Kr[i * 33 + j] = c[0] * M[o] + c[1] * M[1089 + o] + c[2] * M[2178 + o] + c[3] * M[3267 + o] +
                       c[4] * M[4356 + o] + c[5] * M[5445 + o];
    }
  double Kaa[81], y[9];
  for (int r = 0; r < 9; r++)
    for (int q = 0; q < 9; q++) Kaa[r * 9 + q] = Kr[r * 33 + 24 + q];
  chol_dense(Kaa, 9, 1e-14);
  for (int r = 0; r < 9; r++) y[r] = w[24 + r];
  chol_solve<double, double>(Kaa, 9, y);
  for (int t = 0; t < 24; t++) {
    double acc = w[t];
    for (int r = 0; r < 9; r++) acc -= Kr[r * 33 + t] * y[r];
    g[t] = acc;
  }
}

// von Mises stress of a Voigt tensor (engineering shear order xx yy zz xy yz zx)
FE_HD double von_mises(const double *s) {
  const double a = s[0] - s[1], b = s[1] - s[2], c = s[2] - s[0];
  const double v = 0.5 * (a * a + b * b + c * c) + 3 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
  return sqrt(v > 0 ? v : 0.0);
}

}  // namespace fe
