/*
 * fcbga_oracle.c -- CPU FP64 restatement of the reference solver path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP solver and the timed CPU
 * baseline ("port").  The product never links or calls this file.
 *
 * Restated from the reference's behaviour (file:line cited per function):
 *   element with Wilson-Taylor incompatible modes and static condensation
 *     (src/core/07-element.js:19-120, src/core/09-assembly.js:26-133),
 *   full 3x3 block-CSR assembly with element birth and constraints
 *     (src/core/08-sparse.js:7-39, src/core/09-assembly.js:136-165, 392-464),
 *   Jacobi-scaled PCG with the semi-coarsening multigrid: column block
 *     Gauss-Seidel, shell aggregates with Hermite in-plane interpolation,
 *     Galerkin coarse operators and a banded Cholesky coarsest level
 *     (src/core/10-solver.js:12-715, 852-866, 1001-1042),
 *   JEITA warpage (src/core/12-post.js:30-99, src/core/01-util.js:176-190).
 * Condensed element matrices are cached per (group or section, hx, hy, hz):
 * the reference recomputes them per element from the same inputs, so the cache
 * changes the time, not the values.
 */
#define _POSIX_C_SOURCE 200809L
#include "fcbga_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_ms(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec * 1e3 + t.tv_nsec * 1e-6;
}

/* ------------------------------------------------------------------------- */
/* element                                                                   */
/* ------------------------------------------------------------------------- */

static const int HXI[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
static const int HETA[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
static const int HZETA[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
/* Voigt index of the (p, r) pair: xx yy zz xy yz zx (src/core/09-assembly.js:24) */
static const int VIDX[3][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}};

static double U_G[3 * 9 * 121], U_g[3 * 3 * 11], U_GH[9 * 121], U_gH[3 * 11];
static int univ_done = 0;

/* natural-coordinate derivatives of the 8 trilinear + 3 incompatible functions */
static void nat_derivs(double xi, double eta, double zeta, double d[3][11]) {
  for (int a = 0; a < 8; a++) {
    double xa = HXI[a], ya = HETA[a], za = HZETA[a];
    d[0][a] = xa * (1 + ya * eta) * (1 + za * zeta) * 0.125;
    d[1][a] = ya * (1 + xa * xi) * (1 + za * zeta) * 0.125;
    d[2][a] = za * (1 + xa * xi) * (1 + ya * eta) * 0.125;
  }
  d[0][8] = -2 * xi; d[1][8] = 0; d[2][8] = 0;
  d[0][9] = 0; d[1][9] = -2 * eta; d[2][9] = 0;
  d[0][10] = 0; d[1][10] = 0; d[2][10] = -2 * zeta;
}

/* UNIV (src/core/09-assembly.js:26-60): 2x2 Gauss in-plane at zeta = -1, 0, 1,
 * Simpson weights 1/3, 4/3, 1/3 for the full-box sums. */
static void univ_init(void) {
  if (univ_done) return;
  memset(U_G, 0, sizeof U_G); memset(U_g, 0, sizeof U_g);
  memset(U_GH, 0, sizeof U_GH); memset(U_gH, 0, sizeof U_gH);
  const double simpson[3] = {1.0 / 3, 4.0 / 3, 1.0 / 3};
  const double gp = 1 / sqrt(3.0);
  double d[3][11];
  for (int j = 0; j < 3; j++) {
    double zeta = j - 1;
    for (int ip = 0; ip < 4; ip++) {
      double xi = (ip & 1) ? gp : -gp, eta = (ip & 2) ? gp : -gp;
      nat_derivs(xi, eta, zeta, d);
      for (int r = 0; r < 3; r++)
        for (int a = 0; a < 11; a++) {
          U_g[(j * 3 + r) * 11 + a] += d[r][a];
          U_gH[r * 11 + a] += simpson[j] * d[r][a];
          for (int s = 0; s < 3; s++)
            for (int b = 0; b < 11; b++) {
              double v = d[r][a] * d[s][b];
              U_G[(j * 9 + r * 3 + s) * 121 + a * 11 + b] += v;
              U_GH[(r * 3 + s) * 121 + a * 11 + b] += simpson[j] * v;
            }
        }
    }
  }
  univ_done = 1;
}

void or_univ(double *G, double *g, double *GH, double *gH) {
  univ_init();
  memcpy(G, U_G, sizeof U_G); memcpy(g, U_g, sizeof U_g);
  memcpy(GH, U_GH, sizeof U_GH); memcpy(gH, U_gH, sizeof U_gH);
}

/* shapeMatrices (src/core/09-assembly.js:69-93): M^{m}, m = xx yy zz xy yz zx,
 * with K33 = sum_m c_m M^{m}. */
void or_shape_matrices(const double *Wst, const double *D, double *M) {
  univ_init();
  memset(M, 0, sizeof(double) * 6 * 1089);
  static const int PAIR[6][3] = {{0, 0, 0}, {1, 1, 1}, {2, 2, 2}, {3, 0, 1}, {4, 1, 2}, {5, 2, 0}};
  int nst = D ? 1 : 3;
  for (int j = 0; j < nst; j++) {
    const double *Wj = D ? D : Wst + 36 * j;
    for (int pi = 0; pi < 6; pi++) {
      int m = PAIR[pi][0], r = PAIR[pi][1], s = PAIR[pi][2];
      double *Mm = M + m * 1089;
      int nrs = (r == s) ? 1 : 2;
      for (int q2 = 0; q2 < nrs; q2++) {
        int rr = q2 ? s : r, ss = q2 ? r : s;
        const double *Gb = D ? U_GH + (rr * 3 + ss) * 121 : U_G + (j * 9 + rr * 3 + ss) * 121;
        for (int p = 0; p < 3; p++)
          for (int q = 0; q < 3; q++) {
            double dv = Wj[VIDX[p][rr] * 6 + VIDX[q][ss]];
            if (dv == 0) continue;
            for (int a = 0; a < 11; a++) {
              int row = (3 * a + p) * 33 + q, ga = a * 11;
              for (int b = 0; b < 11; b++) Mm[row + 3 * b] += dv * Gb[ga + b];
            }
          }
      }
    }
  }
}

/* elementK33Fast (src/core/09-assembly.js:96-108) */
void or_element_k33(double hx, double hy, double hz, const double *M, double *K) {
  double V8 = hx * hy * hz / 8;
  double c[6] = {V8 * 4 / (hx * hx), V8 * 4 / (hy * hy), V8 * 4 / (hz * hz),
                 V8 * 4 / (hx * hy), V8 * 4 / (hy * hz), V8 * 4 / (hz * hx)};
  for (int i = 0; i < 33; i++)
    for (int j = i; j < 33; j++) {
      int o = i * 33 + j;
      K[o] = c[0] * M[o] + c[1] * M[1089 + o] + c[2] * M[2178 + o] + c[3] * M[3267 + o] + c[4] * M[4356 + o] +
             c[5] * M[5445 + o];
    }
  for (int i = 0; i < 33; i++)
    for (int j = i + 1; j < 33; j++) K[j * 33 + i] = K[i * 33 + j];
}

/* elementFth (src/core/09-assembly.js:115-133) */
void or_element_fth(double hx, double hy, double hz, const double *Sst, const double *s0, double *f) {
  univ_init();
  memset(f, 0, sizeof(double) * 33);
  double h[3] = {hx, hy, hz};
  double V8 = hx * hy * hz / 8;
  int nst = s0 ? 1 : 3;
  for (int j = 0; j < nst; j++) {
    const double *Sj = s0 ? s0 : Sst + 6 * j;
    for (int r = 0; r < 3; r++) {
      double c = V8 * (2 / h[r]);
      const double *gb = s0 ? U_gH + r * 11 : U_g + (j * 3 + r) * 11;
      for (int p = 0; p < 3; p++) {
        double sv = c * Sj[VIDX[p][r]];
        if (sv == 0) continue;
        for (int a = 0; a < 11; a++) f[3 * a + p] += sv * gb[a];
      }
    }
  }
}

/* One step of the stress history of a homogeneous box element, evaluated at
 * natural points (accumulateStep + stateStress, src/core/09-assembly.js:494-571).
 * Restated through the strain at each point instead of the monomial sums:
 * sum_m phi_m D E_m(dq) = D eps(dq; point), with eps from the 11 function
 * derivatives (boxDerivs, src/core/07-element.js:19-31).  dq = (du at the nodes,
 * dalpha = K_aa^-1 (f_a - K_au du)); q (33) += dq; when born, sig (6 per point)
 * += D eps - s0 and eps_out (6 per point, optional) += eps. */
void or_element_step_points(double hx, double hy, double hz, const double *M, const double *D, const double *s0,
                            int born, const double *du24, double *q, int npts, const double *xez, double *sig,
                            double *eps_out) {
  double K[1089], f[33], Kaa[81], dq[33], rhs[9];
  or_element_k33(hx, hy, hz, M, K);
  if (born) or_element_fth(hx, hy, hz, NULL, s0, f);
  else memset(f, 0, sizeof f);
  for (int r = 0; r < 9; r++)
    for (int c = 0; c < 9; c++) Kaa[r * 9 + c] = K[(24 + r) * 33 + 24 + c];
  or_chol_dense(Kaa, 9, 1e-14);
  for (int t = 0; t < 24; t++) dq[t] = du24[t];
  for (int r = 0; r < 9; r++) {
    double v = f[24 + r];
    for (int t = 0; t < 24; t++) v -= K[(24 + r) * 33 + t] * dq[t];
    rhs[r] = v;
  }
  or_chol_solve_dense(Kaa, 9, rhs);
  for (int r = 0; r < 9; r++) dq[24 + r] = rhs[r];
  for (int t = 0; t < 33; t++) q[t] += dq[t];
  const double sc[3] = {2 / hx, 2 / hy, 2 / hz};
  for (int p = 0; p < npts; p++) {
    double d[3][11], e[6] = {0, 0, 0, 0, 0, 0};
    nat_derivs(xez[3 * p], xez[3 * p + 1], xez[3 * p + 2], d);
    for (int b = 0; b < 11; b++) {
      const double gx = d[0][b] * sc[0], gy = d[1][b] * sc[1], gz = d[2][b] * sc[2];
      const double ux = dq[3 * b], uy = dq[3 * b + 1], uz = dq[3 * b + 2];
      e[0] += gx * ux; e[1] += gy * uy; e[2] += gz * uz;
      e[3] += gy * ux + gx * uy; e[4] += gz * uy + gy * uz; e[5] += gz * ux + gx * uz;
    }
    if (eps_out)
      for (int r = 0; r < 6; r++) eps_out[6 * p + r] += e[r];
    if (!born) continue;
    for (int r = 0; r < 6; r++) {
      double v = -s0[r];
      for (int c = 0; c < 6; c++) v += D[6 * r + c] * e[c];
      sig[6 * p + r] += v;
    }
  }
}

/* cholDense with jitter (src/core/01-util.js:70-91) */
int or_chol_dense(double *A, int n, double rel) {
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

/* cholSolveDense (src/core/01-util.js:94-107) */
void or_chol_solve_dense(const double *L, int n, double *b) {
  for (int i = 0; i < n; i++) {
    double s = b[i];
    for (int k = 0; k < i; k++) s -= L[i * n + k] * b[k];
    b[i] = s / L[i * n + i];
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = b[i];
    for (int k = i + 1; k < n; k++) s -= L[k * n + i] * b[k];
    b[i] = s / L[i * n + i];
  }
}

/* condense33 (src/core/07-element.js:88-120) */
void or_condense33(const double *K, const double *f, double *Kc, double *fc, double *X, double *xf) {
  double Kaa[81], col[9];
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 9; j++) Kaa[i * 9 + j] = K[(24 + i) * 33 + 24 + j];
  or_chol_dense(Kaa, 9, 1e-14);
  double Xl[216];
  double *Xp = X ? X : Xl;
  for (int j = 0; j < 24; j++) {
    for (int i = 0; i < 9; i++) col[i] = K[(24 + i) * 33 + j];
    or_chol_solve_dense(Kaa, 9, col);
    for (int i = 0; i < 9; i++) Xp[i * 24 + j] = col[i];
  }
  for (int i = 0; i < 24; i++)
    for (int j = 0; j < 24; j++) {
      double s = K[i * 33 + j];
      for (int q = 0; q < 9; q++) s -= K[i * 33 + 24 + q] * Xp[q * 24 + j];
      Kc[i * 24 + j] = s;
    }
  for (int i = 0; i < 24; i++)
    for (int j = i + 1; j < 24; j++) {
      double v = 0.5 * (Kc[i * 24 + j] + Kc[j * 24 + i]);
      Kc[i * 24 + j] = Kc[j * 24 + i] = v;
    }
  if (f) {
    double xfl[9];
    double *xp = xf ? xf : xfl;
    for (int i = 0; i < 9; i++) col[i] = f[24 + i];
    or_chol_solve_dense(Kaa, 9, col);
    for (int i = 0; i < 9; i++) xp[i] = col[i];
    for (int i = 0; i < 24; i++) {
      double s = f[i];
      for (int q = 0; q < 9; q++) s -= K[i * 33 + 24 + q] * xp[q];
      fc[i] = s;
    }
  }
}

/* Independent check: B^T D B with physical derivatives at the 8 Gauss points of
 * a box (the box case of generalElement, src/core/07-element.js:195-205, where
 * the Taylor factor is 1). */
void or_box_gauss_k33(double hx, double hy, double hz, const double *D, const double *s0, double *K, double *f) {
  memset(K, 0, sizeof(double) * 1089);
  memset(f, 0, sizeof(double) * 33);
  const double gp = 1 / sqrt(3.0);
  double sc[3] = {2 / hx, 2 / hy, 2 / hz};
  double w = hx * hy * hz / 8;
  for (int ig = 0; ig < 8; ig++) {
    double xi = (ig & 1) ? gp : -gp, eta = (ig & 2) ? gp : -gp, zeta = (ig & 4) ? gp : -gp;
    double d[3][11];
    nat_derivs(xi, eta, zeta, d);
    /* B: 6 x 33, strain [exx eyy ezz gxy gyz gzx] */
    double B[6][33];
    memset(B, 0, sizeof B);
    for (int a = 0; a < 11; a++) {
      double dx = d[0][a] * sc[0], dy = d[1][a] * sc[1], dz = d[2][a] * sc[2];
      B[0][3 * a] = dx; B[1][3 * a + 1] = dy; B[2][3 * a + 2] = dz;
      B[3][3 * a] = dy; B[3][3 * a + 1] = dx;
      B[4][3 * a + 1] = dz; B[4][3 * a + 2] = dy;
      B[5][3 * a] = dz; B[5][3 * a + 2] = dx;
    }
    double DB[6][33];
    for (int r = 0; r < 6; r++)
      for (int c = 0; c < 33; c++) {
        double s = 0;
        for (int t = 0; t < 6; t++) s += D[r * 6 + t] * B[t][c];
        DB[r][c] = s;
      }
    for (int i = 0; i < 33; i++) {
      for (int j = 0; j < 33; j++) {
        double s = 0;
        for (int r = 0; r < 6; r++) s += B[r][i] * DB[r][j];
        K[i * 33 + j] += w * s;
      }
      if (s0) {
        double s = 0;
        for (int r = 0; r < 6; r++) s += B[r][i] * s0[r];
        f[i] += w * s;
      }
    }
  }
}

/* ------------------------------------------------------------------------- */
/* condensed element cache                                                    */
/* ------------------------------------------------------------------------- */

typedef struct kc_entry {
  int32_t key_mat;
  double hx, hy, hz;
  double Kc[576];
  double fc[24];
  int has_f;
} kc_entry;

typedef struct kc_cache {
  kc_entry *e;
  int32_t *slot;
  int64_t cap, n, ecap;
} kc_cache;

static uint64_t hkey(int32_t m, double hx, double hy, double hz) {
  uint64_t a, b, c;
  memcpy(&a, &hx, 8); memcpy(&b, &hy, 8); memcpy(&c, &hz, 8);
  uint64_t h = (uint64_t)(uint32_t)m * 0x9E3779B97F4A7C15ull;
  h ^= a + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  h ^= b + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  h ^= c + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  return h;
}

static int32_t kc_find(kc_cache *C, int32_t m, double hx, double hy, double hz, int *found) {
  uint64_t h = hkey(m, hx, hy, hz);
  int64_t i = (int64_t)(h & (uint64_t)(C->cap - 1));
  for (;;) {
    int32_t s = C->slot[i];
    if (s < 0) { *found = 0; return (int32_t)i; }
    kc_entry *e = C->e + s;
    if (e->key_mat == m && e->hx == hx && e->hy == hy && e->hz == hz) { *found = 1; return s; }
    i = (i + 1) & (C->cap - 1);
  }
}

/* ------------------------------------------------------------------------- */
/* system: pattern + values                                                   */
/* ------------------------------------------------------------------------- */

typedef struct mg_level mg_level;

struct or_sys {
  const or_model *m;
  int32_t nn;
  int32_t *rowPtr, *col, *diag;
  int64_t nb;
  double *val, *rhs;
  uint8_t *isC;
  double assemble_ms;
  int32_t ntypes;
};

/* bcsrFromElements + fullPatternFromUpper (src/core/08-sparse.js:7-39,
 * src/core/09-assembly.js:136-148): on the structured grid the full pattern of a
 * node is every node sharing an element with it, sorted. */
static void build_pattern(or_sys *S) {
  const or_model *m = S->m;
  int32_t nn = m->nn, ne = m->ne;
  int32_t *cnt = calloc(nn + 1, sizeof(int32_t));
  for (int64_t q = 0; q < (int64_t)ne * 8; q++) cnt[m->conn[q] + 1]++;
  for (int32_t i = 0; i < nn; i++) cnt[i + 1] += cnt[i];
  int32_t *adj = malloc(sizeof(int32_t) * (size_t)ne * 8), *fill = malloc(sizeof(int32_t) * nn);
  memcpy(fill, cnt, sizeof(int32_t) * nn);
  for (int32_t e = 0; e < ne; e++)
    for (int a = 0; a < 8; a++) { int32_t n = m->conn[8 * e + a]; adj[fill[n]++] = e; }
  int32_t *mark = malloc(sizeof(int32_t) * nn);
  for (int32_t i = 0; i < nn; i++) mark[i] = -1;
  S->rowPtr = malloc(sizeof(int32_t) * (nn + 1));
  S->rowPtr[0] = 0;
  for (int32_t i = 0; i < nn; i++) {
    int c = 0;
    for (int32_t q = cnt[i]; q < cnt[i + 1]; q++) {
      int32_t e = adj[q];
      for (int a = 0; a < 8; a++) { int32_t j = m->conn[8 * e + a]; if (mark[j] != i) { mark[j] = i; c++; } }
    }
    S->rowPtr[i + 1] = S->rowPtr[i] + c;
  }
  S->nb = S->rowPtr[nn];
  S->col = malloc(sizeof(int32_t) * S->nb);
  for (int32_t i = 0; i < nn; i++) mark[i] = -1;
  for (int32_t i = 0; i < nn; i++) {
    int32_t w = S->rowPtr[i];
    for (int32_t q = cnt[i]; q < cnt[i + 1]; q++) {
      int32_t e = adj[q];
      for (int a = 0; a < 8; a++) {
        int32_t j = m->conn[8 * e + a];
        if (mark[j] != i) { mark[j] = i; S->col[w++] = j; }
      }
    }
    /* insertion sort (<= 27 entries) */
    int32_t *c = S->col + S->rowPtr[i];
    int32_t len = S->rowPtr[i + 1] - S->rowPtr[i];
    for (int32_t a = 1; a < len; a++) {
      int32_t v = c[a], b = a - 1;
      while (b >= 0 && c[b] > v) { c[b + 1] = c[b]; b--; }
      c[b + 1] = v;
    }
  }
  S->diag = malloc(sizeof(int32_t) * nn);
  for (int32_t i = 0; i < nn; i++)
    for (int32_t k = S->rowPtr[i]; k < S->rowPtr[i + 1]; k++)
      if (S->col[k] == i) { S->diag[i] = k; break; }
  free(cnt); free(adj); free(fill); free(mark);
}

static int32_t find_block(const or_sys *S, int32_t i, int32_t j) {
  int32_t lo = S->rowPtr[i], hi = S->rowPtr[i + 1] - 1;
  while (lo <= hi) {
    int32_t mid = (lo + hi) >> 1, v = S->col[mid];
    if (v == j) return mid;
    if (v < j) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* assembleStep + applyConstraints (src/core/09-assembly.js:392-424, 445-464) */
static void assemble(or_sys *S) {
  const or_model *m = S->m;
  double t0 = now_ms();
  int32_t nn = m->nn;
  S->val = calloc((size_t)S->nb * 9, sizeof(double));
  S->rhs = calloc((size_t)nn * 3, sizeof(double));
  /* per group / section shape matrices */
  int32_t nmat = m->ngroups + (m->e_section ? m->nsec : 0);
  double *Ms = malloc(sizeof(double) * 6 * 1089 * (size_t)nmat);
  for (int32_t g = 0; g < m->ngroups; g++) or_shape_matrices(NULL, m->D + 36 * g, Ms + (size_t)6 * 1089 * g);
  if (m->e_section)
    for (int32_t s = 0; s < m->nsec; s++)
      or_shape_matrices(m->secW + 108 * s, NULL, Ms + (size_t)6 * 1089 * (m->ngroups + s));
  kc_cache C;
  C.cap = 1 << 16;
  while (C.cap < 4 * 4096) C.cap <<= 1;
  C.slot = malloc(sizeof(int32_t) * C.cap);
  for (int64_t i = 0; i < C.cap; i++) C.slot[i] = -1;
  C.ecap = 4096; C.n = 0;
  C.e = malloc(sizeof(kc_entry) * C.ecap);
  double K33[1089], f33[33];
  int32_t *bl = malloc(sizeof(int32_t) * 64);
  for (int32_t e = 0; e < m->ne; e++) {
    int32_t g = m->e_group[e];
    int32_t sec = m->e_section ? m->e_section[e] : -1;
    int born = m->born[g] != 0;
    double scale = m->e_scale[e] * (born ? 1.0 : 1e-6);
    int32_t i = m->e_ijk[3 * e], j = m->e_ijk[3 * e + 1], k = m->e_ijk[3 * e + 2];
    double hx = m->xs[i + 1] - m->xs[i], hy = m->ys[j + 1] - m->ys[j], hz = m->zs[k + 1] - m->zs[k];
    int32_t key = sec >= 0 ? m->ngroups + sec : g;
    int hasF = sec >= 0 ? (m->sec_born[sec] != 0) : born;
    int found;
    int32_t pos = kc_find(&C, key, hx, hy, hz, &found);
    kc_entry *E;
    if (found) E = C.e + pos;
    else {
      if (C.n == C.ecap) { C.ecap *= 2; C.e = realloc(C.e, sizeof(kc_entry) * C.ecap); }
      E = C.e + C.n;
      C.slot[pos] = (int32_t)C.n;
      C.n++;
      if (C.n * 2 > C.cap) { /* grow the table */
        int64_t nc = C.cap * 2;
        int32_t *ns = malloc(sizeof(int32_t) * nc);
        for (int64_t q = 0; q < nc; q++) ns[q] = -1;
        free(C.slot); C.slot = ns; C.cap = nc;
        for (int64_t q = 0; q < C.n; q++) {
          int f2; int32_t p2 = kc_find(&C, C.e[q].key_mat, C.e[q].hx, C.e[q].hy, C.e[q].hz, &f2);
          C.slot[p2] = (int32_t)q;
        }
      }
      E->key_mat = key; E->hx = hx; E->hy = hy; E->hz = hz;
      or_element_k33(hx, hy, hz, Ms + (size_t)6 * 1089 * key, K33);
      E->has_f = hasF;
      if (hasF) {
        if (sec >= 0) or_element_fth(hx, hy, hz, m->secV + 18 * sec, NULL, f33);
        else or_element_fth(hx, hy, hz, NULL, m->s0 + 6 * g, f33);
      }
      or_condense33(K33, hasF ? f33 : NULL, E->Kc, E->fc, NULL, NULL);
    }
    const int32_t *cn = m->conn + 8 * e;
    for (int a = 0; a < 8; a++)
      for (int b = 0; b < 8; b++) bl[a * 8 + b] = find_block(S, cn[a], cn[b]);
    for (int a = 0; a < 8; a++) {
      for (int b = 0; b < 8; b++) {
        double *v = S->val + 9 * (int64_t)bl[a * 8 + b];
        for (int p = 0; p < 3; p++) {
          const double *r = E->Kc + (3 * a + p) * 24 + 3 * b;
          v[3 * p] += scale * r[0]; v[3 * p + 1] += scale * r[1]; v[3 * p + 2] += scale * r[2];
        }
      }
      if (E->has_f) {
        int32_t n = cn[a];
        S->rhs[3 * n] += scale * E->fc[3 * a];
        S->rhs[3 * n + 1] += scale * E->fc[3 * a + 1];
        S->rhs[3 * n + 2] += scale * E->fc[3 * a + 2];
      }
    }
  }
  S->ntypes = (int32_t)C.n;
  free(bl); free(C.e); free(C.slot); free(Ms);
  /* constraints */
  S->isC = calloc((size_t)nn * 3, 1);
  for (int32_t q = 0; q < m->nconstraints; q++) S->isC[m->constraints[q]] = 1;
  for (int32_t q = 0; q < m->nconstraints; q++) {
    int32_t d = m->constraints[q], i = d / 3, p = d % 3;
    for (int32_t k = S->rowPtr[i]; k < S->rowPtr[i + 1]; k++) {
      int32_t j = S->col[k];
      double *v = S->val + 9 * (int64_t)k;
      for (int qq = 0; qq < 3; qq++) {
        if (3 * j + qq != d) v[3 * p + qq] = 0;
        else if (v[3 * p + qq] == 0) v[3 * p + qq] = 1;
      }
      int32_t kk = find_block(S, j, i);
      double *w = S->val + 9 * (int64_t)kk;
      for (int qq = 0; qq < 3; qq++)
        if (3 * j + qq != d) w[3 * qq + p] = 0;
    }
    S->rhs[d] = 0;
  }
  S->assemble_ms = now_ms() - t0;
}

or_sys *or_build(const or_model *m) {
  or_sys *S = calloc(1, sizeof(or_sys));
  S->m = m;
  S->nn = m->nn;
  univ_init();
  build_pattern(S);
  assemble(S);
  return S;
}

void or_free(or_sys *S) {
  if (!S) return;
  free(S->rowPtr); free(S->col); free(S->diag); free(S->val); free(S->rhs); free(S->isC);
  free(S);
}

int32_t or_sys_nn(const or_sys *s) { return s->nn; }
int64_t or_sys_nb(const or_sys *s) { return s->nb; }
const int32_t *or_sys_rowptr(const or_sys *s) { return s->rowPtr; }
const int32_t *or_sys_col(const or_sys *s) { return s->col; }
double *or_sys_val(or_sys *s) { return s->val; }
double *or_sys_rhs(or_sys *s) { return s->rhs; }
double or_sys_assemble_ms(const or_sys *s) { return s->assemble_ms; }
int32_t or_sys_ntypes(const or_sys *s) { return s->ntypes; }

/* fullMatVec / blockMatVec (src/core/09-assembly.js:151-165, src/core/10-solver.js:134-153) */
static void bmatvec(int32_t n, const int32_t *rp, const int32_t *col, const double *val, int b, const double *x,
                    double *y) {
  int bb = b * b;
  for (int32_t i = 0; i < n; i++) {
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int32_t k = rp[i]; k < rp[i + 1]; k++) {
      const double *v = val + (int64_t)bb * k, *xj = x + (int64_t)b * col[k];
      for (int p = 0; p < b; p++) {
        double s = 0;
        for (int q = 0; q < b; q++) s += v[p * b + q] * xj[q];
        acc[p] += s;
      }
    }
    for (int p = 0; p < b; p++) y[(int64_t)b * i + p] = acc[p];
  }
}

void or_matvec(const or_sys *S, const double *x, double *y) { bmatvec(S->nn, S->rowPtr, S->col, S->val, 3, x, y); }

/* ------------------------------------------------------------------------- */
/* multigrid (src/core/10-solver.js:12-715)                                  */
/* ------------------------------------------------------------------------- */

struct mg_level {
  int kind; /* 0 solid, 1 shell */
  int b, nx, ny, nz;
  int32_t nn;
  double *xs, *ys;
  int32_t *colStart;
  /* operator (block CSR, block b) */
  int32_t *rowPtr, *col, *diag;
  double *val;
  int64_t nb;
  int own_matrix;
  /* solid */
  const int32_t *nodeId;
  const double *coords;
  int32_t *down;
  double *colFact;
  /* shell */
  double *zref, *zlo, *zhi;
  int64_t *cfPtr;
  double *cfData;
  /* feature lines */
  int32_t *keepX, *keepY;
  int nkx, nky;
  /* prolongation to the next level: ptr[nn+1], idx, W (b x bc per entry) */
  int32_t *pPtr, *pIdx;
  double *pW;
  int32_t *nodeAgg;
  /* transpose lists (children) of the prolongation INTO this level */
  int32_t *chPtr, *chIdx, *chW;
  /* coarsest */
  double *band;
  int bw, bandW;
  int32_t denseN;
  /* work */
  double *r, *dx, *bv, *xv, *res;
};

typedef struct mg {
  mg_level *L[24];
  int nl;
  double setup_ms;
} mg;

/* chol3 (src/core/10-solver.js:12-25) */
static int chol3(const double *A, double *L, double jit) {
  double a00 = A[0] + jit, a10 = A[3], a11 = A[4] + jit, a20 = A[6], a21 = A[7], a22 = A[8] + jit;
  if (!(a00 > 0)) return 0;
  double l00 = sqrt(a00), l10 = a10 / l00, l20 = a20 / l00;
  double d11 = a11 - l10 * l10;
  if (!(d11 > 0)) return 0;
  double l11 = sqrt(d11), l21 = (a21 - l20 * l10) / l11;
  double d22 = a22 - l20 * l20 - l21 * l21;
  if (!(d22 > 0)) return 0;
  L[0] = l00; L[1] = 0; L[2] = 0; L[3] = l10; L[4] = l11; L[5] = 0; L[6] = l20; L[7] = l21; L[8] = sqrt(d22);
  return 1;
}

/* factorColumns (src/core/10-solver.js:31-68) */
static void factor_columns(mg_level *lv) {
  int32_t nn = lv->nn;
  if (!lv->colFact) lv->colFact = calloc((size_t)nn * 18, sizeof(double));
  double *F = lv->colFact;
  double tmp[9], D[9];
  for (int32_t n = 0; n < nn; n++) {
    const double *dv = lv->val + 9 * (int64_t)lv->diag[n];
    for (int q = 0; q < 9; q++) D[q] = dv[q];
    double *fo = F + 18 * (int64_t)n;
    if (lv->down[n] >= 0) {
      const double *av = lv->val + 9 * (int64_t)lv->down[n];
      const double *lp = F + 18 * (int64_t)(n - 1);
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
    for (int t = 0; t < 8 && !chol3(D, fo, jit); t++) {
      double mx = fmax(fmax(fabs(D[0]), fabs(D[4])), fmax(fabs(D[8]), 1e-300));
      jit = jit != 0 ? jit * 10 : 1e-10 * mx;
    }
  }
}

/* solveColumn (src/core/10-solver.js:71-104) */
static void solve_column(const mg_level *lv, int32_t n0, int32_t n1, const double *r, double *x) {
  const double *F = lv->colFact;
  for (int32_t n = n0; n < n1; n++) {
    const double *fo = F + 18 * (int64_t)n;
    double b0 = r[3 * n], b1 = r[3 * n + 1], b2 = r[3 * n + 2];
    if (n > n0) {
      double p0 = x[3 * n - 3], p1 = x[3 * n - 2], p2 = x[3 * n - 1];
      b0 -= fo[9] * p0 + fo[10] * p1 + fo[11] * p2;
      b1 -= fo[12] * p0 + fo[13] * p1 + fo[14] * p2;
      b2 -= fo[15] * p0 + fo[16] * p1 + fo[17] * p2;
    }
    double y0 = b0 / fo[0];
    double y1 = (b1 - fo[3] * y0) / fo[4];
    double y2 = (b2 - fo[6] * y0 - fo[7] * y1) / fo[8];
    x[3 * n] = y0; x[3 * n + 1] = y1; x[3 * n + 2] = y2;
  }
  for (int32_t n = n1 - 1; n >= n0; n--) {
    const double *fo = F + 18 * (int64_t)n;
    double b0 = x[3 * n], b1 = x[3 * n + 1], b2 = x[3 * n + 2];
    if (n + 1 < n1) {
      const double *fn = F + 18 * (int64_t)(n + 1);
      double q0 = x[3 * n + 3], q1 = x[3 * n + 4], q2 = x[3 * n + 5];
      b0 -= fn[9] * q0 + fn[12] * q1 + fn[15] * q2;
      b1 -= fn[10] * q0 + fn[13] * q1 + fn[16] * q2;
      b2 -= fn[11] * q0 + fn[14] * q1 + fn[17] * q2;
    }
    double z2 = b2 / fo[8];
    double z1 = (b1 - fo[7] * z2) / fo[4];
    double z0 = (b0 - fo[3] * z1 - fo[6] * z2) / fo[0];
    x[3 * n] = z0; x[3 * n + 1] = z1; x[3 * n + 2] = z2;
  }
}

static int32_t bfind(const int32_t *rp, const int32_t *col, int32_t i, int32_t j) {
  int32_t lo = rp[i], hi = rp[i + 1] - 1;
  while (lo <= hi) {
    int32_t mid = (lo + hi) >> 1, v = col[mid];
    if (v == j) return mid;
    if (v < j) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* transposeProl (src/core/10-solver.js:180-187): children of each coarse node */
static void transpose_prol(mg_level *fine, mg_level *coarse) {
  int32_t nnc = coarse->nn, nf = fine->nn;
  int32_t nnz = fine->pPtr[nf];
  coarse->chPtr = calloc(nnc + 1, sizeof(int32_t));
  for (int32_t q = 0; q < nnz; q++) coarse->chPtr[fine->pIdx[q] + 1]++;
  for (int32_t c = 0; c < nnc; c++) coarse->chPtr[c + 1] += coarse->chPtr[c];
  coarse->chIdx = malloc(sizeof(int32_t) * (nnz ? nnz : 1));
  coarse->chW = malloc(sizeof(int32_t) * (nnz ? nnz : 1));
  int32_t *fill = malloc(sizeof(int32_t) * (nnc ? nnc : 1));
  memcpy(fill, coarse->chPtr, sizeof(int32_t) * nnc);
  for (int32_t f = 0; f < nf; f++)
    for (int32_t q = fine->pPtr[f]; q < fine->pPtr[f + 1]; q++) {
      int32_t c = fine->pIdx[q];
      coarse->chIdx[fill[c]] = f;
      coarse->chW[fill[c]] = q;
      fill[c]++;
    }
  free(fill);
}

static int cmp_i32(const void *a, const void *b) {
  int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

/* galerkinProduct (src/core/10-solver.js:193-245) */
static void galerkin(mg_level *fine, mg_level *coarse) {
  int bf = fine->b, bc = coarse->b, bfc = bf * bc, bcc = bc * bc, bff = bf * bf;
  int32_t nnc = coarse->nn;
  int32_t *mark = malloc(sizeof(int32_t) * nnc);
  for (int32_t i = 0; i < nnc; i++) mark[i] = -1;
  double *acc = malloc(sizeof(double) * (size_t)nnc * bcc);
  int32_t *touched = malloc(sizeof(int32_t) * (nnc ? nnc : 1));
  int32_t *rowPtr = calloc(nnc + 1, sizeof(int32_t));
  int64_t cap = 27 * (int64_t)nnc + 64, nb = 0;
  int32_t *cols = malloc(sizeof(int32_t) * cap);
  double *vals = malloc(sizeof(double) * cap * bcc);
  double T[36];
  const double *W = fine->pW;
  for (int32_t I = 0; I < nnc; I++) {
    int nt = 0;
    for (int32_t q = coarse->chPtr[I]; q < coarse->chPtr[I + 1]; q++) {
      int32_t i = coarse->chIdx[q];
      const double *Wi = W + (int64_t)coarse->chW[q] * bfc;
      for (int32_t k = fine->rowPtr[i]; k < fine->rowPtr[i + 1]; k++) {
        int32_t j = fine->col[k];
        const double *v = fine->val + (int64_t)bff * k;
        for (int32_t p = fine->pPtr[j]; p < fine->pPtr[j + 1]; p++) {
          int32_t J = fine->pIdx[p];
          const double *Wj = W + (int64_t)p * bfc;
          if (mark[J] != I) {
            mark[J] = I;
            touched[nt++] = J;
            memset(acc + (int64_t)bcc * J, 0, sizeof(double) * bcc);
          }
          for (int r = 0; r < bf; r++)
            for (int c = 0; c < bc; c++) {
              double s = 0;
              for (int t = 0; t < bf; t++) s += v[r * bf + t] * Wj[t * bc + c];
              T[r * bc + c] = s;
            }
          double *a = acc + (int64_t)bcc * J;
          for (int r = 0; r < bc; r++)
            for (int c = 0; c < bc; c++) {
              double s = 0;
              for (int t = 0; t < bf; t++) s += Wi[t * bc + r] * T[t * bc + c];
              a[r * bc + c] += s;
            }
        }
      }
    }
    qsort(touched, nt, sizeof(int32_t), cmp_i32);
    rowPtr[I + 1] = rowPtr[I] + nt;
    if (nb + nt > cap) {
      cap = (int64_t)(1.5 * (nb + nt)) + 64;
      cols = realloc(cols, sizeof(int32_t) * cap);
      vals = realloc(vals, sizeof(double) * cap * bcc);
    }
    for (int q = 0; q < nt; q++) {
      int32_t J = touched[q];
      cols[nb] = J;
      memcpy(vals + nb * bcc, acc + (int64_t)bcc * J, sizeof(double) * bcc);
      nb++;
    }
  }
  coarse->rowPtr = rowPtr;
  coarse->col = cols;
  coarse->val = vals;
  coarse->nb = nb;
  coarse->own_matrix = 1;
  coarse->diag = malloc(sizeof(int32_t) * (nnc ? nnc : 1));
  for (int32_t I = 0; I < nnc; I++)
    for (int32_t k = rowPtr[I]; k < rowPtr[I + 1]; k++)
      if (cols[k] == I) { coarse->diag[I] = k; break; }
  free(mark); free(acc); free(touched);
}

/* parts that split a column into separate shell aggregates (src/core/10-solver.js:248) */
static int compliant_part(int p) { return p == 1 || p == 2 || p == 3 || p == 5 || p == 7 || p == 9; }

/* buildShellLevel + shellWeights (src/core/10-solver.js:260-353), mergeSingles 'below'.
 * The element class of a node column (i, j) looks at cell index (i*(ny-1)+j)*(nz-1)+k
 * exactly as the reference does (for j = ny-1 that index falls into the next row of
 * cells and for i = nx-1 past the end, where it reads as no element). */
static mg_level *build_shell_level(mg_level *lv, const double *s, const or_model *m) {
  int nx = lv->nx, ny = lv->ny, nz = lv->nz, ncy = ny - 1, ncz = nz - 1;
  int64_t ncells = (int64_t)(nx - 1) * ncy * ncz;
  int32_t ncol = nx * ny;
  mg_level *co = calloc(1, sizeof(mg_level));
  co->kind = 1; co->b = 6; co->nx = nx; co->ny = ny;
  co->xs = malloc(sizeof(double) * nx); memcpy(co->xs, lv->xs, sizeof(double) * nx);
  co->ys = malloc(sizeof(double) * ny); memcpy(co->ys, lv->ys, sizeof(double) * ny);
  co->colStart = malloc(sizeof(int32_t) * (ncol + 1));
  lv->nodeAgg = malloc(sizeof(int32_t) * lv->nn);
  for (int32_t n = 0; n < lv->nn; n++) lv->nodeAgg[n] = -1;
  int64_t acap = 4 * (int64_t)ncol + 16;
  co->zref = malloc(sizeof(double) * acap); co->zlo = malloc(sizeof(double) * acap); co->zhi = malloc(sizeof(double) * acap);
  int32_t na = 0;
  /* per column runs: node lists */
  int32_t *rn = malloc(sizeof(int32_t) * nz);  /* node ids in run order */
  int32_t *rs = malloc(sizeof(int32_t) * (nz + 1)); /* run starts */
  int32_t *tmpn = malloc(sizeof(int32_t) * nz);
  for (int32_t c = 0; c < ncol; c++) {
    co->colStart[c] = na;
    int i = c / ny, j = c % ny;
    int nruns = 0, len = 0, have = 0, prevK = -2, prevType = 0;
    for (int k = 0; k < nz; k++) {
      int32_t n = lv->nodeId[(int64_t)c * nz + k];
      if (n < 0) { have = 0; prevK = -2; continue; }
      int cls[2];
      for (int side = 0; side < 2; side++) {
        int kk = side == 0 ? k - 1 : k;
        if (kk < 0 || kk >= ncz) { cls[side] = 0; continue; }
        int64_t cc = ((int64_t)i * ncy + j) * ncz + kk;
        int any = 0, stiff = 0;
        for (int q = 0; q < 2; q++) {
          int32_t e = (cc < ncells) ? m->cell_elems[2 * cc + q] : -1;
          if (e < 0) continue;
          any = 1;
          if (!compliant_part(m->e_part[e])) stiff = 1;
        }
        cls[side] = !any ? 0 : (stiff ? 2 : 1);
      }
      int below = cls[0], above = cls[1];
      int type = (below == 2 || above == 2) ? 2 : 1;
      int joined = 0;
      if (have && prevK == k - 1 && below != 0) joined = below == 2 || (type == 1 && prevType == 1);
      if (!joined) { rs[nruns++] = len; have = 1; }
      rn[len++] = n;
      prevK = k; prevType = type;
    }
    rs[nruns] = len;
    /* merge single-node runs into the run below (or above for the first) */
    for (int r = 0; r < nruns; r++) {
      int rl = rs[r + 1] - rs[r];
      if (rl != 1 || nruns < 2) continue;
      int32_t node = rn[rs[r]];
      int t = r > 0 ? r - 1 : r + 1;
      /* rebuild rn without run r, inserting node at the end of run t (t < r) or start (t > r) */
      int w = 0;
      int32_t *nrs = malloc(sizeof(int32_t) * (nruns + 1));
      int nr2 = 0;
      for (int q = 0; q < nruns; q++) {
        if (q == r) continue;
        nrs[nr2++] = w;
        if (q == t && t > r) tmpn[w++] = node;
        for (int p = rs[q]; p < rs[q + 1]; p++) tmpn[w++] = rn[p];
        if (q == t && t < r) tmpn[w++] = node;
      }
      nrs[nr2] = w;
      memcpy(rn, tmpn, sizeof(int32_t) * w);
      memcpy(rs, nrs, sizeof(int32_t) * (nr2 + 1));
      free(nrs);
      nruns = nr2;
      r--;
    }
    for (int r = 0; r < nruns; r++) {
      if (na + 1 > acap) {
        acap *= 2;
        co->zref = realloc(co->zref, sizeof(double) * acap);
        co->zlo = realloc(co->zlo, sizeof(double) * acap);
        co->zhi = realloc(co->zhi, sizeof(double) * acap);
      }
      double zs = 0;
      for (int p = rs[r]; p < rs[r + 1]; p++) { lv->nodeAgg[rn[p]] = na; zs += lv->coords[3 * (int64_t)rn[p] + 2]; }
      co->zlo[na] = lv->coords[3 * (int64_t)rn[rs[r]] + 2];
      co->zhi[na] = lv->coords[3 * (int64_t)rn[rs[r + 1] - 1] + 2];
      co->zref[na] = zs / (rs[r + 1] - rs[r]);
      na++;
    }
  }
  co->colStart[ncol] = na;
  co->nn = na;
  free(rn); free(rs); free(tmpn);
  if (lv->nkx) { co->keepX = malloc(sizeof(int32_t) * lv->nkx); memcpy(co->keepX, lv->keepX, sizeof(int32_t) * lv->nkx); co->nkx = lv->nkx; }
  if (lv->nky) { co->keepY = malloc(sizeof(int32_t) * lv->nky); memcpy(co->keepY, lv->keepY, sizeof(int32_t) * lv->nky); co->nky = lv->nky; }
  /* shellWeights */
  int32_t cnt = 0;
  for (int32_t n = 0; n < lv->nn; n++) if (lv->nodeAgg[n] >= 0) cnt++;
  lv->pPtr = calloc(lv->nn + 1, sizeof(int32_t));
  lv->pIdx = malloc(sizeof(int32_t) * (cnt ? cnt : 1));
  lv->pW = calloc((size_t)(cnt ? cnt : 1) * 18, sizeof(double));
  for (int32_t n = 0, q = 0; n < lv->nn; n++) {
    lv->pPtr[n + 1] = lv->pPtr[n];
    if (lv->nodeAgg[n] < 0) continue;
    lv->pPtr[n + 1]++;
    lv->pIdx[q] = lv->nodeAgg[n];
    double zp = lv->coords[3 * (int64_t)n + 2] - co->zref[lv->nodeAgg[n]];
    double *o = lv->pW + 18 * (int64_t)q++;
    o[0] = 1 / s[3 * n]; o[4] = zp / s[3 * n];
    o[7] = 1 / s[3 * n + 1]; o[9] = -zp / s[3 * n + 1];
    o[14] = 1 / s[3 * n + 2]; o[17] = zp / s[3 * n + 2];
  }
  transpose_prol(lv, co);
  galerkin(lv, co);
  return co;
}

/* coarseLines (src/core/10-solver.js:171-177) */
static int coarse_lines(int n, const int32_t *keep, int nk, int32_t *out) {
  char *in = calloc(n, 1);
  for (int i = 0; i < n; i += 2) in[i] = 1;
  in[n - 1] = 1;
  for (int q = 0; q < nk; q++) if (keep[q] >= 0 && keep[q] < n) in[keep[q]] = 1;
  int c = 0;
  for (int i = 0; i < n; i++) if (in[i]) out[c++] = i;
  free(in);
  return c;
}

/* hermitePair (src/core/10-solver.js:356-363) */
static void hermite_pair(double t, double h, int alongY, double *WA, double *WB) {
  memset(WA, 0, 36 * sizeof(double)); memset(WB, 0, 36 * sizeof(double));
  for (int d = 0; d < 6; d++) { WA[d * 6 + d] = 1 - t; WB[d * 6 + d] = t; }
  double h00 = 2 * t * t * t - 3 * t * t + 1, h10 = t * t * t - 2 * t * t + t, h01 = -2 * t * t * t + 3 * t * t,
         h11 = t * t * t - t * t;
  WA[2 * 6 + 2] = h00; WB[2 * 6 + 2] = h01;
  if (alongY) { WA[2 * 6 + 3] = h * h10; WB[2 * 6 + 3] = h * h11; }
  else { WA[2 * 6 + 4] = -h * h10; WB[2 * 6 + 4] = -h * h11; }
}

/* rbmShift (src/core/10-solver.js:366-373) */
static void rbm_shift(double dz, double *M) {
  memset(M, 0, 36 * sizeof(double));
  for (int d = 0; d < 6; d++) M[d * 7] = 1;
  M[0 * 6 + 4] = dz; M[1 * 6 + 3] = -dz; M[2 * 6 + 5] = dz;
}

static void mul66(const double *A, const double *B, double *C) {
  double t[36];
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) {
      double s = 0;
      for (int k = 0; k < 6; k++) s += A[i * 6 + k] * B[k * 6 + j];
      t[i * 6 + j] = s;
    }
  memcpy(C, t, sizeof t);
}

/* matchAggregate (src/core/10-solver.js:381-389) */
static int32_t match_aggregate(const mg_level *co, int32_t C, double zlo, double zhi, double zref) {
  int32_t best = -1;
  double bo = 0;
  for (int32_t a = co->colStart[C]; a < co->colStart[C + 1]; a++) {
    double ov = fmin(zhi, co->zhi[a]) - fmax(zlo, co->zlo[a]);
    double score = ov > 0 ? ov + 1e-9 : ((zref >= co->zlo[a] - 1e-9 && zref <= co->zhi[a] + 1e-9) ? 1e-9 : -1);
    if (score > bo) { bo = score; best = a; }
  }
  return best;
}

typedef struct loc1 { int I0, I1; double t, h; } loc1;

static loc1 locate(const int32_t *lines, int nl, const double *coords, int i, double x) {
  loc1 r;
  int I0 = 0;
  while (I0 + 1 < nl && lines[I0 + 1] <= i) I0++;
  r.I0 = I0;
  if (lines[I0] == i) { r.I1 = -1; r.t = 0; r.h = 0; return r; }
  r.I1 = I0 + 1;
  r.t = (x - coords[I0]) / (coords[I0 + 1] - coords[I0]);
  r.h = coords[I0 + 1] - coords[I0];
  return r;
}

/* buildCoarseShellLevel (src/core/10-solver.js:396-471) */
static mg_level *build_coarse_shell_level(mg_level *lv) {
  int32_t *cx = malloc(sizeof(int32_t) * lv->nx), *cy = malloc(sizeof(int32_t) * lv->ny);
  int nxc = coarse_lines(lv->nx, lv->keepX, lv->nkx, cx), nyc = coarse_lines(lv->ny, lv->keepY, lv->nky, cy);
  mg_level *co = calloc(1, sizeof(mg_level));
  co->kind = 1; co->b = 6; co->nx = nxc; co->ny = nyc;
  co->xs = malloc(sizeof(double) * nxc); co->ys = malloc(sizeof(double) * nyc);
  for (int I = 0; I < nxc; I++) co->xs[I] = lv->xs[cx[I]];
  for (int J = 0; J < nyc; J++) co->ys[J] = lv->ys[cy[J]];
  co->colStart = malloc(sizeof(int32_t) * (nxc * nyc + 1));
  int64_t acap = lv->nn + 16;
  co->zref = malloc(sizeof(double) * acap); co->zlo = malloc(sizeof(double) * acap); co->zhi = malloc(sizeof(double) * acap);
  int32_t na = 0;
  for (int I = 0; I < nxc; I++)
    for (int J = 0; J < nyc; J++) {
      co->colStart[I * nyc + J] = na;
      int32_t c = cx[I] * lv->ny + cy[J];
      for (int32_t a = lv->colStart[c]; a < lv->colStart[c + 1]; a++) {
        co->zlo[na] = lv->zlo[a]; co->zhi[na] = lv->zhi[a]; co->zref[na] = lv->zref[a];
        na++;
      }
    }
  co->colStart[nxc * nyc] = na;
  co->nn = na;
  /* feature lines in the coarse index space */
  if (lv->nkx) {
    co->keepX = malloc(sizeof(int32_t) * lv->nkx);
    for (int q = 0; q < lv->nkx; q++) {
      int f = -1;
      for (int I = 0; I < nxc; I++) if (cx[I] == lv->keepX[q]) { f = I; break; }
      if (f >= 0) co->keepX[co->nkx++] = f;
    }
  }
  if (lv->nky) {
    co->keepY = malloc(sizeof(int32_t) * lv->nky);
    for (int q = 0; q < lv->nky; q++) {
      int f = -1;
      for (int J = 0; J < nyc; J++) if (cy[J] == lv->keepY[q]) { f = J; break; }
      if (f >= 0) co->keepY[co->nky++] = f;
    }
  }
  /* prolongation: at most 4 parents per fine aggregate */
  lv->pPtr = calloc(lv->nn + 1, sizeof(int32_t));
  lv->pIdx = malloc(sizeof(int32_t) * 4 * (size_t)(lv->nn ? lv->nn : 1));
  lv->pW = malloc(sizeof(double) * 36 * 4 * (size_t)(lv->nn ? lv->nn : 1));
  double WL[36], WR[36], WB[36], WT[36], tmp[36], sh[36], I6[36];
  memset(I6, 0, sizeof I6);
  for (int d = 0; d < 6; d++) I6[d * 7] = 1;
  int32_t q = 0;
  for (int i = 0; i < lv->nx; i++) {
    loc1 px = locate(cx, nxc, co->xs, i, lv->xs[i]);
    for (int j = 0; j < lv->ny; j++) {
      loc1 py = locate(cy, nyc, co->ys, j, lv->ys[j]);
      int32_t c = i * lv->ny + j;
      for (int32_t f = lv->colStart[c]; f < lv->colStart[c + 1]; f++) {
        int xC[2] = {px.I0, px.I1}, yC[2] = {py.I0, py.I1};
        int nxC = px.I1 < 0 ? 1 : 2, nyC = py.I1 < 0 ? 1 : 2;
        int32_t par[2][2];
        int complete = 1;
        for (int a = 0; a < nxC; a++)
          for (int b = 0; b < nyC; b++) {
            par[a][b] = match_aggregate(co, xC[a] * nyc + yC[b], lv->zlo[f], lv->zhi[f], lv->zref[f]);
            if (par[a][b] < 0) complete = 0;
          }
        int32_t q0 = q;
#define PUSH(A, B, W)                                              \
  do {                                                             \
    int32_t ag = par[A][B];                                        \
    rbm_shift(lv->zref[f] - co->zref[ag], sh);                     \
    mul66((W), sh, lv->pW + 36 * (int64_t)q);                      \
    lv->pIdx[q] = ag;                                              \
    q++;                                                           \
  } while (0)
        if (complete) {
          if (px.I1 >= 0) hermite_pair(px.t, px.h, 0, WL, WR);
          if (py.I1 >= 0) hermite_pair(py.t, py.h, 1, WB, WT);
          if (px.I1 < 0 && py.I1 < 0) { PUSH(0, 0, I6); }
          else if (py.I1 < 0) { PUSH(0, 0, WL); PUSH(1, 0, WR); }
          else if (px.I1 < 0) { PUSH(0, 0, WB); PUSH(0, 1, WT); }
          else {
            mul66(WB, WL, tmp); PUSH(0, 0, tmp);
            mul66(WB, WR, tmp); PUSH(1, 0, tmp);
            mul66(WT, WL, tmp); PUSH(0, 1, tmp);
            mul66(WT, WR, tmp); PUSH(1, 1, tmp);
          }
        } else {
          double wx[2] = {px.I1 < 0 ? 1 : 1 - px.t, px.t}, wy[2] = {py.I1 < 0 ? 1 : 1 - py.t, py.t};
          double sum = 0;
          for (int a = 0; a < nxC; a++)
            for (int b = 0; b < nyC; b++) {
              double w = wx[a] * wy[b];
              if (par[a][b] >= 0 && w > 0) sum += w;
            }
          for (int a = 0; a < nxC; a++)
            for (int b = 0; b < nyC; b++) {
              double w = wx[a] * wy[b];
              if (par[a][b] >= 0 && w > 0) {
                double Wd[36];
                memset(Wd, 0, sizeof Wd);
                for (int d = 0; d < 6; d++) Wd[d * 7] = w / sum;
                PUSH(a, b, Wd);
              }
            }
        }
#undef PUSH
        lv->pPtr[f + 1] = q - q0;
      }
    }
  }
  for (int32_t f = 0; f < lv->nn; f++) lv->pPtr[f + 1] += lv->pPtr[f];
  free(cx); free(cy);
  transpose_prol(lv, co);
  galerkin(lv, co);
  return co;
}

/* factorShellColumns (src/core/10-solver.js:474-494) */
static void factor_shell_columns(mg_level *lv) {
  int b = lv->b, bb = b * b;
  int32_t ncol = lv->nx * lv->ny;
  lv->cfPtr = calloc(ncol + 1, sizeof(int64_t));
  for (int32_t c = 0; c < ncol; c++) {
    int64_t mm = (int64_t)(lv->colStart[c + 1] - lv->colStart[c]) * b;
    lv->cfPtr[c + 1] = lv->cfPtr[c] + mm * mm;
  }
  lv->cfData = calloc(lv->cfPtr[ncol] ? lv->cfPtr[ncol] : 1, sizeof(double));
  for (int32_t c = 0; c < ncol; c++) {
    int32_t n0 = lv->colStart[c], n1 = lv->colStart[c + 1];
    int mm = (n1 - n0) * b;
    if (!mm) continue;
    double *A = lv->cfData + lv->cfPtr[c];
    for (int32_t n = n0; n < n1; n++)
      for (int32_t k = lv->rowPtr[n]; k < lv->rowPtr[n + 1]; k++) {
        int32_t j = lv->col[k];
        if (j < n0 || j >= n1) continue;
        const double *v = lv->val + (int64_t)bb * k;
        for (int p = 0; p < b; p++)
          for (int qq = 0; qq < b; qq++) A[((n - n0) * b + p) * mm + (j - n0) * b + qq] = v[p * b + qq];
      }
    or_chol_dense(A, mm, 1e-12);
  }
}

/* columnSweep (src/core/10-solver.js:512-555) */
static void column_sweep(mg_level *lv, const double *bvec, double *x, int backward) {
  int b = lv->b, bb = b * b;
  int32_t ncol = lv->nx * lv->ny;
  double *r = lv->r, *dx = lv->dx;
  for (int32_t cc = 0; cc < ncol; cc++) {
    int32_t c = backward ? ncol - 1 - cc : cc;
    int32_t n0 = lv->colStart[c], n1 = lv->colStart[c + 1];
    if (n1 == n0) continue;
    for (int32_t n = n0; n < n1; n++) {
      double s[6];
      for (int p = 0; p < b; p++) s[p] = bvec[(int64_t)b * n + p];
      for (int32_t k = lv->rowPtr[n]; k < lv->rowPtr[n + 1]; k++) {
        const double *v = lv->val + (int64_t)bb * k, *xj = x + (int64_t)b * lv->col[k];
        for (int p = 0; p < b; p++) {
          double t = 0;
          for (int qq = 0; qq < b; qq++) t += v[p * b + qq] * xj[qq];
          s[p] -= t;
        }
      }
      for (int p = 0; p < b; p++) r[(int64_t)b * n + p] = s[p];
    }
    if (lv->kind == 0) {
      solve_column(lv, n0, n1, r, dx);
      for (int64_t q = 3 * (int64_t)n0; q < 3 * (int64_t)n1; q++) x[q] += dx[q];
    } else {
      int mm = (n1 - n0) * b;
      or_chol_solve_dense(lv->cfData + lv->cfPtr[c], mm, r + (int64_t)b * n0);
      for (int64_t q = (int64_t)b * n0; q < (int64_t)b * n1; q++) x[q] += r[q];
    }
  }
}

/* factorCoarsest / solveCoarsest (src/core/10-solver.js:562-617) */
static void factor_coarsest(mg_level *lv) {
  int b = lv->b, bb = b * b;
  int32_t n = b * lv->nn;
  int bw = 0;
  for (int32_t i = 0; i < lv->nn; i++)
    for (int32_t k = lv->rowPtr[i]; k < lv->rowPtr[i + 1]; k++) {
      int d = abs(lv->col[k] - i);
      if (d > bw) bw = d;
    }
  bw = (bw + 1) * b - 1;
  int w = bw + 1;
  double *A = calloc((size_t)n * w, sizeof(double));
  for (int32_t i = 0; i < lv->nn; i++)
    for (int32_t k = lv->rowPtr[i]; k < lv->rowPtr[i + 1]; k++) {
      int32_t j = lv->col[k];
      if (j > i) continue;
      const double *v = lv->val + (int64_t)bb * k;
      for (int p = 0; p < b; p++)
        for (int q = 0; q < b; q++) {
          int32_t r = b * i + p, c = b * j + q;
          if (c <= r) A[(int64_t)r * w + (c - r + bw)] = v[p * b + q];
        }
    }
  double dmax = 0;
  for (int32_t i = 0; i < n; i++) dmax = fmax(dmax, fabs(A[(int64_t)i * w + bw]));
  double jitter = 1e-12 * (dmax != 0 ? dmax : 1);
  for (int32_t j = 0; j < n; j++) {
    double d = A[(int64_t)j * w + bw];
    int32_t k0 = j - bw > 0 ? j - bw : 0;
    for (int32_t k = k0; k < j; k++) { double l = A[(int64_t)j * w + (k - j + bw)]; d -= l * l; }
    if (!(d > jitter)) d = fmax(fabs(d), jitter);
    double ljj = sqrt(d);
    A[(int64_t)j * w + bw] = ljj;
    double inv = 1 / ljj;
    int32_t iend = j + bw + 1 < n ? j + bw + 1 : n;
    for (int32_t i = j + 1; i < iend; i++) {
      double sum = A[(int64_t)i * w + (j - i + bw)];
      int32_t kk0 = k0 > i - bw ? k0 : i - bw;
      for (int32_t k = kk0; k < j; k++) sum -= A[(int64_t)i * w + (k - i + bw)] * A[(int64_t)j * w + (k - j + bw)];
      A[(int64_t)i * w + (j - i + bw)] = sum * inv;
    }
  }
  lv->band = A; lv->bandW = w; lv->bw = bw; lv->denseN = n;
}

static void solve_coarsest(const mg_level *lv, double *x) {
  const double *A = lv->band;
  int w = lv->bandW, bw = lv->bw;
  int32_t n = lv->denseN;
  for (int32_t i = 0; i < n; i++) {
    double sum = x[i];
    int32_t k0 = i - bw > 0 ? i - bw : 0;
    for (int32_t k = k0; k < i; k++) sum -= A[(int64_t)i * w + (k - i + bw)] * x[k];
    x[i] = sum / A[(int64_t)i * w + bw];
  }
  for (int32_t i = n - 1; i >= 0; i--) {
    double sum = x[i];
    int32_t kend = i + bw + 1 < n ? i + bw + 1 : n;
    for (int32_t k = i + 1; k < kend; k++) sum -= A[(int64_t)k * w + (i - k + bw)] * x[k];
    x[i] = sum / A[(int64_t)i * w + bw];
  }
}

/* MultigridPC constructor (src/core/10-solver.js:620-651) */
static mg *mg_build(or_sys *S, const double *s) {
  const or_model *m = S->m;
  double t0 = now_ms();
  mg *M = calloc(1, sizeof(mg));
  mg_level *f = calloc(1, sizeof(mg_level));
  f->kind = 0; f->b = 3; f->nx = m->nx; f->ny = m->ny; f->nz = m->nz; f->nn = m->nn;
  f->xs = malloc(sizeof(double) * m->nx); memcpy(f->xs, m->xs, sizeof(double) * m->nx);
  f->ys = malloc(sizeof(double) * m->ny); memcpy(f->ys, m->ys, sizeof(double) * m->ny);
  f->nodeId = m->node_id; f->coords = m->coords;
  f->colStart = malloc(sizeof(int32_t) * (m->nx * m->ny + 1));
  memcpy(f->colStart, m->col_start, sizeof(int32_t) * (m->nx * m->ny + 1));
  f->rowPtr = S->rowPtr; f->col = S->col; f->diag = S->diag; f->val = S->val; f->nb = S->nb;
  if (m->nfx) { f->keepX = malloc(sizeof(int32_t) * m->nfx); memcpy(f->keepX, m->feature_x, sizeof(int32_t) * m->nfx); f->nkx = m->nfx; }
  if (m->nfy) { f->keepY = malloc(sizeof(int32_t) * m->nfy); memcpy(f->keepY, m->feature_y, sizeof(int32_t) * m->nfy); f->nky = m->nfy; }
  /* columnDownIndex (src/core/10-solver.js:122-131) */
  f->down = malloc(sizeof(int32_t) * f->nn);
  for (int32_t n = 0; n < f->nn; n++) f->down[n] = -1;
  for (int32_t c = 0; c < m->nx * m->ny; c++)
    for (int32_t n = f->colStart[c] + 1; n < f->colStart[c + 1]; n++) f->down[n] = bfind(f->rowPtr, f->col, n, n - 1);
  M->L[0] = f;
  M->nl = 1;
  factor_columns(f);
  mg_level *lv = build_shell_level(f, s, m);
  M->L[M->nl++] = lv;
  for (int guard = 0; guard < 20; guard++) {
    factor_shell_columns(lv);
    int32_t ncol = lv->nx * lv->ny, ndof = lv->b * lv->nn;
    if (ncol <= 9 || ndof <= 900 || (lv->nx <= 2 && lv->ny <= 2)) break;
    int32_t *tmp = malloc(sizeof(int32_t) * (lv->nx > lv->ny ? lv->nx : lv->ny));
    int ncx = coarse_lines(lv->nx, lv->keepX, lv->nkx, tmp), ncy = coarse_lines(lv->ny, lv->keepY, lv->nky, tmp);
    free(tmp);
    if ((double)ncx * ncy > 0.7 * ncol) break;
    lv = build_coarse_shell_level(lv);
    M->L[M->nl++] = lv;
  }
  factor_coarsest(lv);
  for (int l = 0; l < M->nl; l++) {
    mg_level *L = M->L[l];
    int64_t n = (int64_t)L->b * L->nn;
    L->bv = calloc(n ? n : 1, sizeof(double)); L->xv = calloc(n ? n : 1, sizeof(double));
    L->res = calloc(n ? n : 1, sizeof(double)); L->r = calloc(n ? n : 1, sizeof(double));
    L->dx = calloc(n ? n : 1, sizeof(double));
  }
  M->setup_ms = now_ms() - t0;
  return M;
}

static void mg_free(mg *M) {
  for (int l = 0; l < M->nl; l++) {
    mg_level *L = M->L[l];
    free(L->xs); free(L->ys); free(L->colStart); free(L->down); free(L->colFact);
    free(L->zref); free(L->zlo); free(L->zhi); free(L->cfPtr); free(L->cfData);
    free(L->keepX); free(L->keepY); free(L->pPtr); free(L->pIdx); free(L->pW); free(L->nodeAgg);
    free(L->chPtr); free(L->chIdx); free(L->chW); free(L->band);
    free(L->bv); free(L->xv); free(L->res); free(L->r); free(L->dx);
    if (L->own_matrix) { free(L->rowPtr); free(L->col); free(L->val); free(L->diag); }
    free(L);
  }
  free(M);
}

/* _cycle / _cycleOnce (src/core/10-solver.js:666-712): 2 forward pre-sweeps,
 * 2 backward post-sweeps */
static void mg_cycle(mg *M, int li, const double *bvec, double *x) {
  mg_level *lv = M->L[li];
  int64_t n = (int64_t)lv->b * lv->nn;
  if (li == M->nl - 1) {
    memcpy(x, bvec, sizeof(double) * n);
    solve_coarsest(lv, x);
    return;
  }
  memset(x, 0, sizeof(double) * n);
  for (int s = 0; s < 2; s++) column_sweep(lv, bvec, x, 0);
  double *res = lv->res;
  bmatvec(lv->nn, lv->rowPtr, lv->col, lv->val, lv->b, x, res);
  for (int64_t i = 0; i < n; i++) res[i] = bvec[i] - res[i];
  mg_level *co = M->L[li + 1];
  int bf = lv->b, bc = co->b, bfc = bf * bc;
  double *rc = co->bv;
  memset(rc, 0, sizeof(double) * (int64_t)bc * co->nn);
  for (int32_t f = 0; f < lv->nn; f++)
    for (int32_t q = lv->pPtr[f]; q < lv->pPtr[f + 1]; q++) {
      int32_t c = lv->pIdx[q];
      const double *w = lv->pW + (int64_t)bfc * q;
      for (int a = 0; a < bc; a++) {
        double s = 0;
        for (int p = 0; p < bf; p++) s += w[p * bc + a] * res[(int64_t)bf * f + p];
        rc[(int64_t)bc * c + a] += s;
      }
    }
  mg_cycle(M, li + 1, rc, co->xv);
  const double *ec = co->xv;
  for (int32_t f = 0; f < lv->nn; f++)
    for (int32_t q = lv->pPtr[f]; q < lv->pPtr[f + 1]; q++) {
      int32_t c = lv->pIdx[q];
      const double *w = lv->pW + (int64_t)bfc * q;
      for (int p = 0; p < bf; p++) {
        double s = 0;
        for (int a = 0; a < bc; a++) s += w[p * bc + a] * ec[(int64_t)bc * c + a];
        x[(int64_t)bf * f + p] += s;
      }
    }
  for (int s = 0; s < 2; s++) column_sweep(lv, bvec, x, 1);
}

static double dotv(const double *a, const double *b, int64_t n) {
  double s = 0;
  for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* LinearSolver.solve with a cold start (src/core/10-solver.js:933-994) and _pcg (:1001-1042) */
/* ---- IC(0) with the Manteuffel shift (scalarUpperCSR, ICPC: src/core/10-solver.js:718-800) ---- */
typedef struct icpc {
  int64_t n;
  int64_t *rp;
  int32_t *col;
  double *v;
  double shift;
} icpc;

static int ic_factor(icpc *U) {
  const int64_t n = U->n;
  const int64_t *rp = U->rp;
  const int32_t *col = U->col;
  double *v = U->v;
  for (int64_t i = 0; i < n; i++) {
    const double d = v[rp[i]];
    if (!(d > 0)) return 0;
    const double uii = sqrt(d);
    v[rp[i]] = uii;
    for (int64_t k = rp[i] + 1; k < rp[i + 1]; k++) v[k] /= uii;
    for (int64_t k1 = rp[i] + 1; k1 < rp[i + 1]; k1++) {
      const int32_t j1 = col[k1];
      const double u1 = v[k1];
      int64_t kk = rp[j1];
      const int64_t kend = rp[j1 + 1];
      for (int64_t k2 = k1; k2 < rp[i + 1]; k2++) {
        const int32_t j2 = col[k2];
        while (kk < kend && col[kk] < j2) kk++;
        if (kk < kend && col[kk] == j2) v[kk] -= u1 * v[k2];
      }
    }
  }
  return 1;
}

/* scalar upper CSR of the (scaled) block matrix, then the factor with the
 * shift schedule 0, 1e-3, 1e-2, ... capped at 1 (8 attempts); NULL on failure */
static icpc *ic_build(const or_sys *S) {
  const int32_t nn = S->nn;
  icpc *U = calloc(1, sizeof *U);
  U->n = 3 * (int64_t)nn;
  U->rp = malloc(sizeof(int64_t) * (U->n + 1));
  U->rp[0] = 0;
  for (int32_t i = 0; i < nn; i++)
    for (int p = 0; p < 3; p++) {
      int64_t c = 0;
      for (int32_t k = S->rowPtr[i]; k < S->rowPtr[i + 1]; k++) {
        const int32_t j = S->col[k];
        if (j > i) c += 3;
        else if (j == i) c += 3 - p;
      }
      U->rp[3 * i + p + 1] = U->rp[3 * i + p] + c;
    }
  const int64_t nnz = U->rp[U->n];
  U->col = malloc(sizeof(int32_t) * nnz);
  double *v0 = malloc(sizeof(double) * nnz);
  for (int32_t i = 0; i < nn; i++)
    for (int p = 0; p < 3; p++) {
      int64_t w = U->rp[3 * i + p];
      for (int32_t k = S->rowPtr[i]; k < S->rowPtr[i + 1]; k++) {
        const int32_t j = S->col[k];
        if (j < i) continue;
        for (int q = (j == i ? p : 0); q < 3; q++) {
          U->col[w] = 3 * j + q;
          v0[w] = S->val[9 * (int64_t)k + 3 * p + q];
          w++;
        }
      }
    }
  U->v = malloc(sizeof(double) * nnz);
  double shift = 0;
  for (int attempt = 0; attempt < 8; attempt++) {
    memcpy(U->v, v0, sizeof(double) * nnz);
    if (shift > 0)
      for (int64_t i = 0; i < U->n; i++) U->v[U->rp[i]] *= 1 + shift;
    if (ic_factor(U)) {
      U->shift = shift;
      free(v0);
      return U;
    }
    shift = shift == 0 ? 1e-3 : shift * 10; /* IC_SHIFT0, src/core/00-constants.js:83 */
    if (shift > 1) shift = 1;               /* IC_SHIFT_MAX, src/core/00-constants.js:84 */
  }
  free(v0);
  free(U->rp); free(U->col); free(U->v); free(U);
  return NULL;
}

static void ic_free(icpc *U) {
  if (!U) return;
  free(U->rp); free(U->col); free(U->v); free(U);
}

/* z = (U^T U)^-1 r (ICPC.apply) */
static void ic_apply(const icpc *U, const double *r, double *z) {
  const int64_t n = U->n, *rp = U->rp;
  const int32_t *col = U->col;
  const double *v = U->v;
  memcpy(z, r, sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) {
    const double yi = z[i] / v[rp[i]];
    z[i] = yi;
    for (int64_t k = rp[i] + 1; k < rp[i + 1]; k++) z[col[k]] -= v[k] * yi;
  }
  for (int64_t i = n - 1; i >= 0; i--) {
    double sum = z[i];
    for (int64_t k = rp[i] + 1; k < rp[i + 1]; k++) sum -= v[k] * z[col[k]];
    z[i] = sum / v[rp[i]];
  }
}

int or_solve(or_sys *S, double tol, int32_t maxit, int32_t precond, double *u, or_stats *st) {
  double t0 = now_ms();
  memset(st, 0, sizeof *st);
  int32_t nn = S->nn;
  int64_t n = 3 * (int64_t)nn;
  double *s = malloc(sizeof(double) * n);
  /* _scale */
  for (int32_t i = 0; i < nn; i++) {
    const double *v = S->val + 9 * (int64_t)S->diag[i];
    s[3 * i] = 1 / sqrt(fmax(v[0], 1e-300));
    s[3 * i + 1] = 1 / sqrt(fmax(v[4], 1e-300));
    s[3 * i + 2] = 1 / sqrt(fmax(v[8], 1e-300));
  }
  for (int32_t i = 0; i < nn; i++)
    for (int32_t k = S->rowPtr[i]; k < S->rowPtr[i + 1]; k++) {
      int32_t j = S->col[k];
      double *v = S->val + 9 * (int64_t)k;
      for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) v[3 * p + q] *= s[3 * i + p] * s[3 * j + q];
    }
  double *b = malloc(sizeof(double) * n), *x = calloc(n, sizeof(double));
  double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n), *p = malloc(sizeof(double) * n),
         *q = malloc(sizeof(double) * n);
  for (int64_t i = 0; i < n; i++) b[i] = S->rhs[i] * s[i];
  double bnorm = sqrt(dotv(b, b, n));
  st->ms_assemble = S->assemble_ms;
  if (bnorm == 0) {
    memset(u, 0, sizeof(double) * n);
    st->converged = 1;
    st->ms_total = now_ms() - t0;
    free(s); free(b); free(x); free(r); free(z); free(p); free(q);
    return 0;
  }
  double t1 = now_ms();
  mg *M = precond == 0 ? mg_build(S, s) : NULL;
  icpc *IC = precond == 2 ? ic_build(S) : NULL;
  if (precond == 2 && !IC) { st->ms_total = now_ms() - t0; free(s); free(b); free(x); free(r); free(z); free(p); free(q); return -1; }
  st->ms_setup = now_ms() - t1;
  if (M) { st->levels = M->nl; st->coarsest_dof = M->L[M->nl - 1]->denseN; }
  if (IC) st->ic_shift = IC->shift;
  double t2 = now_ms();
  const uint8_t *isC = S->isC;
  int it = 0, conv = 0, total = 0;
  double rnorm = 0;
  /* _pcg from the current x; after a failed run with one preconditioner the
   * reference rebuilds the other once and continues from the same iterate
   * (LinearSolver.solve, src/core/10-solver.js:960-972) */
  for (int attempt = 0; attempt < 2; attempt++) {
    or_matvec(S, x, r);
    for (int64_t i = 0; i < n; i++) r[i] = b[i] - r[i];
    for (int64_t i = 0; i < n; i++) if (isC[i]) { r[i] = 0; x[i] = 0; }
    rnorm = sqrt(dotv(r, r, n));
    if (attempt == 0) st->r0 = rnorm / bnorm;
    it = 0;
    if (rnorm / bnorm < tol) { conv = 1; break; }
#define OR_PC(rr, zz) do { if (M) mg_cycle(M, 0, rr, zz); else if (IC) ic_apply(IC, rr, zz); else memcpy(zz, rr, sizeof(double) * n); } while (0)
    OR_PC(r, z);
    for (int64_t i = 0; i < n; i++) if (isC[i]) z[i] = 0;
    memcpy(p, z, sizeof(double) * n);
    double rz = dotv(r, z, n);
    for (it = 1; it <= maxit; it++) {
      or_matvec(S, p, q);
      double pq = dotv(p, q, n);
      if (!(pq > 0)) break;
      double alpha = rz / pq;
      for (int64_t i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
      rnorm = sqrt(dotv(r, r, n));
      if (rnorm / bnorm < tol) { conv = 1; break; }
      OR_PC(r, z);
      for (int64_t i = 0; i < n; i++) if (isC[i]) z[i] = 0;
      double rz2 = dotv(r, z, n);
      double beta = rz2 / rz;
      rz = rz2;
      for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
#undef OR_PC
    if (it > maxit) it = maxit;
    total += it;
    if (conv || precond == 1 || attempt == 1) break;
    /* fallback: the other preconditioner */
    st->fallback_iters = total;
    st->fallback_relres = rnorm / bnorm;
    if (M) { mg_free(M); M = NULL; IC = ic_build(S); if (IC) st->ic_shift = IC->shift; }
    else { ic_free(IC); IC = NULL; M = mg_build(S, s); }
    if (!M && !IC) break;
  }
  it = total;
  st->ms_pcg = now_ms() - t2;
  st->iters = it;
  st->converged = conv;
  st->relres = rnorm / bnorm;
  for (int64_t i = 0; i < n; i++) u[i] = x[i] * s[i];
  if (M) mg_free(M);
  ic_free(IC);
  free(s); free(b); free(x); free(r); free(z); free(p); free(q);
  st->ms_total = now_ms() - t0;
  return 0;
}

/* ------------------------------------------------------------------------- */
/* warpage                                                                   */
/* ------------------------------------------------------------------------- */

/* fitPlane (src/core/01-util.js:176-190) */
void or_fit_plane(const double *xs, const double *ys, const double *ws, int32_t n, double *c) {
  double sx = 0, sy = 0, sw = 0;
  for (int32_t i = 0; i < n; i++) { sx += xs[i]; sy += ys[i]; sw += ws[i]; }
  double mx = sx / n, my = sy / n, mw = sw / n;
  double sxx = 0, sxy = 0, syy = 0, sxw = 0, syw = 0;
  for (int32_t i = 0; i < n; i++) {
    double dx = xs[i] - mx, dy = ys[i] - my, dw = ws[i] - mw;
    sxx += dx * dx; sxy += dx * dy; syy += dy * dy; sxw += dx * dw; syw += dy * dw;
  }
  double det = sxx * syy - sxy * sxy, c1 = 0, c2 = 0;
  if (fabs(det) > 1e-300) { c1 = (sxw * syy - syw * sxy) / det; c2 = (syw * sxx - sxw * sxy) / det; }
  c[0] = mw - c1 * mx - c2 * my; c[1] = c1; c[2] = c2;
}

static int find_interval(const double *a, int n, double x) {
  if (x <= a[0]) return 0;
  if (x >= a[n - 1]) return n - 2;
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

/* samplePlane on plane 0, component z (src/core/12-post.js:30-39) */
static double sample_w(int nx, int ny, int nz, const double *xs, const double *ys, const int32_t *nid, const double *u,
                       double x, double y) {
  int i = find_interval(xs, nx, x), j = find_interval(ys, ny, y);
  double tx = (x - xs[i]) / (xs[i + 1] - xs[i]), ty = (y - ys[j]) / (ys[j + 1] - ys[j]);
  int32_t n00 = nid[((int64_t)i * ny + j) * nz], n10 = nid[((int64_t)(i + 1) * ny + j) * nz];
  int32_t n01 = nid[((int64_t)i * ny + j + 1) * nz], n11 = nid[((int64_t)(i + 1) * ny + j + 1) * nz];
  if (n00 < 0 || n10 < 0 || n01 < 0 || n11 < 0) return NAN;
  return (1 - tx) * (1 - ty) * u[3 * (int64_t)n00 + 2] + tx * (1 - ty) * u[3 * (int64_t)n10 + 2] +
         (1 - tx) * ty * u[3 * (int64_t)n01 + 2] + tx * ty * u[3 * (int64_t)n11 + 2];
}

/* warpageJEITA magnitude and sign (src/core/12-post.js:60-84) */
int or_warpage_jeita(int32_t nx, int32_t ny, int32_t nz, const double *xs, const double *ys, const int32_t *node_id,
                     const double *u, const double *bx, const double *by, int32_t nb, double zx0, double zx1,
                     double zy0, double zy1, or_warpage *out) {
  double *w = malloc(sizeof(double) * nb);
  for (int32_t q = 0; q < nb; q++) w[q] = sample_w(nx, ny, nz, xs, ys, node_id, u, bx[q], by[q]);
  double c[3];
  or_fit_plane(bx, by, w, nb, c);
  double rmin = INFINITY, rmax = -INFINITY;
  for (int32_t q = 0; q < nb; q++) {
    double r = w[q] - (c[0] + c[1] * bx[q] + c[2] * by[q]);
    rmin = fmin(rmin, r); rmax = fmax(rmax, r);
  }
  double mag = rmax - rmin;
  double sums = 0;
  for (int d = 0; d < 2; d++) {
    double xa = zx0, ya = d ? zy1 : zy0, xb = zx1, yb = d ? zy0 : zy1;
    double wa = sample_w(nx, ny, nz, xs, ys, node_id, u, xa, ya), wb = sample_w(nx, ny, nz, xs, ys, node_id, u, xb, yb);
    double mx = 0, mn = 0;
    const int M = 201;
    for (int q = 0; q < M; q++) {
      double t = (double)q / (M - 1);
      double wv = sample_w(nx, ny, nz, xs, ys, node_id, u, xa + t * (xb - xa), ya + t * (yb - ya));
      double rel = wv - (wa + t * (wb - wa));
      if (rel > mx) mx = rel;
      if (rel < mn) mn = rel;
    }
    sums += mx + mn;
  }
  int sign = sums > 0 ? 1 : (sums < 0 ? -1 : 0);
  out->signed_mag = sign * mag; out->mag = mag; out->sign = sign;
  out->c0 = c[0]; out->c1 = c[1]; out->c2 = c[2]; out->rmin = rmin; out->rmax = rmax;
  free(w);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Anand material point (src/core/13-fatigue.js:32-143, 209-218)             */
/* ------------------------------------------------------------------------- */

static double sgnd(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }

/* s_{n+1} for a plastic increment d: Newton with bisection on
 * phi(s) = s - s_n - h0 |1 - s/s*|^a sign(1 - s/s*) d (src/core/13-fatigue.js:32-52) */
static double an_s(const double *P, double sn, double d, double dt, double expQ) {
  const double A = P[0], shat = P[4], n = P[5], h0 = P[6], a = P[7];
  if (d <= 0) return sn;
  double sstar = shat * pow(d / (dt * A) * expQ, n);
  double lo = fmin(sn, sstar), hi = fmax(sn, sstar);
  if (hi - lo < 1e-14) return sn;
  double s = 0.5 * (lo + hi);
  for (int it = 0; it < 60; it++) {
    double B = 1 - s / sstar;
    double f = s - sn - h0 * pow(fabs(B), a) * sgnd(B) * d;
    if (fabs(f) < 1e-11 * fmax(1, sn)) break;
    if (f > 0) hi = s; else lo = s;
    double dphi = 1 + h0 * a * pow(fabs(B), a - 1) * d / sstar;
    double s2 = s - f / dphi;
    if (!(s2 > lo && s2 < hi)) s2 = 0.5 * (lo + hi);
    if (fabs(s2 - s) < 1e-15 * fmax(1, s)) { s = s2; break; }
    s = s2;
  }
  return s;
}

static double an_dsdd(const double *P, double d, double s, double dt, double expQ) {
  const double A = P[0], shat = P[4], n = P[5], h0 = P[6], a = P[7];
  if (d <= 0) return 0;
  double sstar = shat * pow(d / (dt * A) * expQ, n);
  double B = 1 - s / sstar;
  double dps = 1 + h0 * a * pow(fabs(B), a - 1) * d / sstar;
  double dpd = -h0 * pow(fabs(B), a) * sgnd(B) - h0 * a * pow(fabs(B), a - 1) * s * n / sstar;
  return -dpd / dps;
}

/* P = {A, QR, xi, m, shat, n, h0, a, s0, nu}; st (11) committed state with the
 * warm start st[10] updated in place; tr (11) trial state; returns d. */
/* robust = 0: the reference's iteration; robust = 1: a non-positive effective
 * stress (an iterate at or past the bracket end d = sigma_trial / 3G) counts as
 * no flow, g(d) = d > 0, instead of the NaN branch that raises the lower
 * bracket (the rule of the global viscoplastic Newton) */
static double an_step(const double *P, double E, double T, double dt, const double *eps, double *st, double *tr,
                      double *sig, double *D, int robust);

double or_anand_step(const double *P, double E, double T, double dt, const double *eps, double *st, double *tr,
                     double *sig, double *D) {
  return an_step(P, E, T, dt, eps, st, tr, sig, D, 0);
}

static double an_step(const double *P, double E, double T, double dt, const double *eps, double *st, double *tr,
                      double *sig, double *D, int robust) {
  const double A = P[0], QR = P[1], xi = P[2], m = P[3], nu = P[9];
  double G = E / (2 * (1 + nu)), K = E / (3 * (1 - 2 * nu));
  double expQ = exp(QR / (T + 273.15)), rate0 = dt * A / expQ;
  double vol = (eps[0] + eps[1] + eps[2]) / 3, str[6];
  for (int i = 0; i < 3; i++) str[i] = 2 * G * (eps[i] - vol - st[i]);
  for (int i = 3; i < 6; i++) str[i] = 2 * G * (0.5 * eps[i] - st[i]);
  double nrm = sqrt(str[0] * str[0] + str[1] * str[1] + str[2] * str[2] +
                    2 * (str[3] * str[3] + str[4] * str[4] + str[5] * str[5]));
  double sigTr = sqrt(1.5) * nrm, d = 0, s1 = st[6], dDdSig = 0;
  if (sigTr > 1e-14 && dt > 0) {
    double dmax = sigTr / (3 * G), lo = 0, hi = dmax;
    double dd = (st[10] > 0 && st[10] < dmax) ? st[10] : fmin(dmax, rate0 * pow(sinh(xi * sigTr / st[6]), 1 / m));
    if (!isfinite(dd) || dd <= 0) dd = 0.5 * dmax;
    double rs = st[6], rx = 0, rsh = 0;
    for (int it = 0; it < 60; it++) {
      rs = an_s(P, st[6], dd, dt, expQ);
      rx = xi * (sigTr - 3 * G * dd) / rs;
      double g;
      if (robust && !(rx > 0)) {
        rsh = 0;
        g = dd;
      } else {
        rsh = sinh(rx);
        g = dd - rate0 * pow(rsh, 1 / m);
      }
      if (!isfinite(g)) { lo = dd; dd = 0.5 * (lo + hi); continue; }
      if (fabs(g) < 1e-11 * fmax(dmax, 1e-30)) break;
      if (g > 0) hi = dd; else lo = dd;
      double fx = rate0 / m * pow(rsh, 1 / m - 1) * cosh(rx);
      double dg = 1 + fx * xi * (3 * G / rs + (sigTr - 3 * G * dd) * an_dsdd(P, dd, rs, dt, expQ) / (rs * rs));
      double dn = dd - g / dg;
      if (!(dn > lo && dn < hi) || !isfinite(dn)) dn = 0.5 * (lo + hi);
      if (fabs(dn - dd) < 1e-16 * fmax(dmax, 1e-30)) { dd = dn; break; }
      dd = dn;
    }
    d = dd; s1 = rs; st[10] = d;
    double fx = rate0 / m * pow(rsh, 1 / m - 1) * cosh(rx);
    double dg = 1 + fx * xi * (3 * G / rs + (sigTr - 3 * G * dd) * an_dsdd(P, dd, rs, dt, expQ) / (rs * rs));
    dDdSig = (fx * xi / rs) / dg;
  }
  double theta = sigTr > 1e-14 ? 1 - 3 * G * d / sigTr : 1, nv[6];
  for (int i = 0; i < 6; i++) nv[i] = nrm > 0 ? str[i] / nrm : 0;
  for (int i = 0; i < 6; i++) tr[i] = st[i] + sqrt(1.5) * d * nv[i];
  double sm = 3 * K * vol, sigEq = sigTr - 3 * G * d;
  for (int i = 0; i < 6; i++) sig[i] = theta * str[i] + (i < 3 ? sm : 0);
  tr[6] = s1; tr[7] = st[7] + d; tr[8] = st[8] + 0.5 * (st[9] + sigEq) * d; tr[9] = sigEq; tr[10] = st[10];
  if (D) {
    double tb = 2 * G * 1.5 * dDdSig - (1 - theta), c1 = 2 * G * theta, c2 = 2 * G * tb;
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < 6; j++)
        D[6 * i + j] = -c2 * nv[i] * nv[j] + ((i < 3 && j < 3) ? K + c1 * ((i == j) - 1.0 / 3) : (i == j ? 0.5 * c1 : 0));
  }
  return d;
}

void or_anand_batch(const double *P, int64_t n, double E, double T, double dt, const double *eps, double *st,
                    double *tr, double *sig, double *D, double *dep) {
  for (int64_t q = 0; q < n; q++)
    dep[q] = or_anand_step(P, E, T, dt, eps + 6 * q, st + 11 * q, tr + 11 * q, sig + 6 * q, D ? D + 36 * q : NULL);
}

void or_anand_batch_robust(const double *P, int64_t n, double E, double T, double dt, const double *eps, double *st,
                           double *tr, double *sig, double *D, double *dep) {
  for (int64_t q = 0; q < n; q++)
    dep[q] = an_step(P, E, T, dt, eps + 6 * q, st + 11 * q, tr + 11 * q, sig + 6 * q, D ? D + 36 * q : NULL, 1);
}

/* ------------------------------------------------------------------------- */
/* bump screening (screenJoints, src/core/13-fatigue.js:375-455)             */
/* ------------------------------------------------------------------------- */

/* inverse of a 3x3 by cofactors (inv3, src/core/07-element.js:146-153) */
static void inv3c(const double *J, double *o) {
  const double a = J[0], b = J[1], c = J[2], d = J[3], e = J[4], f = J[5], g = J[6], h = J[7], i = J[8];
  const double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
  o[0] = (e * i - f * h) / det; o[1] = (c * h - b * i) / det; o[2] = (b * f - c * e) / det;
  o[3] = (f * g - d * i) / det; o[4] = (a * i - c * g) / det; o[5] = (c * d - a * f) / det;
  o[6] = (d * h - e * g) / det; o[7] = (b * g - a * h) / det; o[8] = (a * e - b * d) / det;
}

/* AnandPoint.stepColumn (src/core/13-fatigue.js:153-166): Newton on the
 * in-plane strains (xx, yy, xy) until the in-plane stresses vanish */
static void col_step(const double *P, double E, double T, double dt, double *eps, double *st, double *tr) {
  static const int Pi[3] = {0, 1, 3};
  double sig[6], D[36], A[9], Ai[9];
  for (int it = 0; it < 20; it++) { /* COLUMN_NEWTON_MAXIT, src/core/00-constants.js:121 */
    or_anand_step(P, E, T, dt, eps, st, tr, sig, D);
    const double r0 = sig[0], r1 = sig[1], r2 = sig[3];
    const double ref = fmax(1.0, fmax(fabs(sig[2]), fmax(fabs(sig[4]), fabs(sig[5]))));
    if (fmax(fabs(r0), fmax(fabs(r1), fabs(r2))) < 1e-3 * ref) break; /* COLUMN_STRESS_TOL */
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) A[3 * r + c] = D[6 * Pi[r] + Pi[c]];
    inv3c(A, Ai);
    for (int r = 0; r < 3; r++) eps[Pi[r]] -= Ai[3 * r] * r0 + Ai[3 * r + 1] * r1 + Ai[3 * r + 2] * r2;
  }
}

/* AnandPoint.stepSeriesColumn (src/core/13-fatigue.js:178-206): solder cap in
 * series with an elastic pillar; Newton on the six cap strains */
static void series_step(const double *P, double E, double T, double dt, double *eps, const double *tz, double fs,
                        const double *Cp, double *st, double *tr) {
  static const int Pi[3] = {0, 1, 3}, Z[3] = {2, 4, 5}, ord[6] = {0, 1, 3, 2, 4, 5};
  const double fp = 1 - fs;
  double sig[6], D[36], J[36], R[6];
  for (int it = 0; it < 20; it++) {
    or_anand_step(P, E, T, dt, eps, st, tr, sig, D);
    for (int r = 0; r < 3; r++) {
      R[r] = sig[Pi[r]];
      R[3 + r] = fs * eps[Z[r]] + fp * Cp[r] * sig[Z[r]] - tz[r];
    }
    const double ref = fmax(1.0, fmax(fabs(sig[2]), fmax(fabs(sig[4]), fabs(sig[5]))));
    double worst = 0;
    for (int r = 0; r < 3; r++) worst = fmax(worst, fmax(fabs(R[r]), fabs(R[3 + r]) * E));
    if (worst < 1e-3 * ref) break;
    for (int r = 0; r < 6; r++)
      for (int c = 0; c < 6; c++) {
        const double Dv = D[6 * ord[r] + ord[c]];
        J[6 * r + c] = r < 3 ? Dv : fp * Cp[r - 3] * Dv + (ord[r] == ord[c] ? fs : 0);
      }
    for (int c = 0; c < 6; c++) {
      int pr = c;
      for (int r = c + 1; r < 6; r++)
        if (fabs(J[6 * r + c]) > fabs(J[6 * pr + c])) pr = r;
      if (pr != c) {
        for (int q = 0; q < 6; q++) { const double t = J[6 * c + q]; J[6 * c + q] = J[6 * pr + q]; J[6 * pr + q] = t; }
        const double t = R[c]; R[c] = R[pr]; R[pr] = t;
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

void or_screen_joints(const double *P, const or_screen_steps *h, int64_t nsites, int32_t ntemps, int32_t i25,
                      const double *table, double *dW, double *dE, double *stab) {
  static const int IP[3] = {0, 1, 3};
  const int nc = h->cycles;
  double *Wc = malloc(sizeof(double) * (nc + 1)), *Ec = malloc(sizeof(double) * (nc + 1));
  double e1[3] = {0, 0, 0}, e2[3] = {0, 0, 0}, s1[6] = {0}, s2[6] = {0};
  const double nu = P[9];
  for (int64_t q = 0; q < nsites; q++) {
    const double *base = table + q * ntemps * 6;
    dW[q] = dE[q] = stab[q] = NAN;
    if (!isfinite(base[0])) continue;
    double st[11] = {0, 0, 0, 0, 0, 0, P[8], 0, 0, 0, 0}, tr[11], eps[6] = {0, 0, 0, 0, 0, 0}, e25[6];
    for (int c = 0; c < 6; c++) e25[c] = i25 >= 0 ? base[i25 * 6 + c] : 0;
    for (int c = 0; c <= nc; c++) Wc[c] = Ec[c] = 0;
    for (int k = 1; k < h->nsteps; k++) {
      const double *a = base + h->wi[k] * 6, *b = a + 6;
      const double t = h->wt[k];
      double lay[6];
      for (int c = 0; c < 6; c++) lay[c] = a[c] + t * (b[c] - a[c]) - e25[c];
      const int ext = k > 2 && h->same[k];
      if (h->Cp) {
        const double tz[3] = {lay[2] - h->fs * h->Fth[k] - (1 - h->fs) * h->Fp[k], lay[4], lay[5]};
        if (ext)
          for (int c = 0; c < 6; c++) eps[c] = 2 * s1[c] - s2[c];
        series_step(P, h->E[k], h->T[k], h->dt[k], eps, tz, h->fs, h->Cp + 3 * k, st, tr);
        for (int c = 0; c < 6; c++) { s2[c] = s1[c]; s1[c] = eps[c]; }
      } else {
        const double ez = lay[2] - h->Fth[k];
        if (ext) {
          for (int r = 0; r < 3; r++) eps[IP[r]] = 2 * e1[r] - e2[r];
        } else {
          eps[0] -= nu * (ez - eps[2]);
          eps[1] -= nu * (ez - eps[2]);
        }
        eps[2] = ez; eps[4] = lay[4]; eps[5] = lay[5];
        col_step(P, h->E[k], h->T[k], h->dt[k], eps, st, tr);
      }
      for (int c = 0; c < 10; c++) st[c] = tr[c]; /* commit (the warm start st[10] is the step's) */
      for (int r = 0; r < 3; r++) { e2[r] = e1[r]; e1[r] = eps[IP[r]]; }
      Wc[h->cycle[k]] = st[8];
      Ec[h->cycle[k]] = st[7];
    }
    dW[q] = Wc[nc] - Wc[nc - 1];
    dE[q] = Ec[nc] - Ec[nc - 1];
    stab[q] = nc >= 2 ? fabs((Wc[nc] - Wc[nc - 1]) - (Wc[nc - 1] - Wc[nc - 2])) / fmax(Wc[nc] - Wc[nc - 1], 1e-30) : NAN;
  }
  free(Wc); free(Ec);
}
