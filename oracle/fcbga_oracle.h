/*
 * fcbga_oracle.h -- CPU FP64 restatement of the reference solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP solver
 * (tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it);
 * the product never links, loads or calls it.
 *
 * Each function cites the reference file:line whose algorithm it restates
 * (paths relative to the reference repository root).
 */
#ifndef FCBGA_ORACLE_H
#define FCBGA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- element (src/core/07-element.js, src/core/09-assembly.js:26-133) ---- */
void or_univ(double *G /*3*9*121*/, double *g /*3*3*11*/, double *GH /*9*121*/, double *gH /*3*11*/);
void or_shape_matrices(const double *Wst /*108 or NULL*/, const double *D /*36 or NULL*/, double *M /*6*1089*/);
void or_element_k33(double hx, double hy, double hz, const double *M, double *K /*1089*/);
void or_element_fth(double hx, double hy, double hz, const double *Sst /*18 or NULL*/, const double *s0 /*6 or NULL*/,
                    double *f /*33*/);
/* one step of a homogeneous element's stress history at natural points (src/core/09-assembly.js:494-571) */
void or_element_step_points(double hx, double hy, double hz, const double *M, const double *D, const double *s0,
                            int born, const double *du24, double *q, int npts, const double *xez, double *sig,
                            double *eps_out);
int or_chol_dense(double *A, int n, double rel);
void or_chol_solve_dense(const double *L, int n, double *b);
void or_condense33(const double *K, const double *f /*nullable*/, double *Kc, double *fc, double *X, double *xf);
/* brute-force 2x2x2 Gauss integration of the same box element (an independent check of the closed form) */
void or_box_gauss_k33(double hx, double hy, double hz, const double *D, const double *s0, double *K, double *f);

/* ---- model ---- */
typedef struct or_model {
  int32_t nx, ny, nz, nn, ne;
  const double *xs, *ys, *zs;
  const int32_t *node_id;    /* nx*ny*nz, -1 absent */
  const int32_t *cell_elems; /* 2*(nx-1)*(ny-1)*(nz-1) */
  const int32_t *e_group;    /* ne */
  const int8_t *e_part;      /* ne */
  const double *e_scale;     /* ne */
  const int32_t *e_ijk;      /* 3*ne */
  const int32_t *conn;       /* 8*ne */
  const double *coords;      /* 3*nn */
  const int32_t *col_start;  /* nx*ny+1 */
  int32_t ngroups;
  const double *D;           /* ngroups*36 step-mean stiffness */
  const double *s0;          /* ngroups*6 step thermal stress increment */
  const uint8_t *born;       /* ngroups */
  const int32_t *e_section;  /* ne or NULL: layered section of an element (-1 homogeneous) */
  int32_t nsec;
  const double *secW;        /* nsec*108 */
  const double *secV;        /* nsec*18 */
  const uint8_t *sec_born;   /* nsec */
  int32_t nconstraints;
  const int32_t *constraints;
  int32_t nfx, nfy;
  const int32_t *feature_x, *feature_y;
} or_model;

typedef struct or_sys or_sys;

typedef struct or_stats {
  int32_t iters;
  int32_t converged;
  double relres;
  double r0;
  double ms_total;
  double ms_assemble;
  double ms_setup;
  double ms_pcg;
  int32_t levels;
  int32_t coarsest_dof;
  int32_t fallback_iters;   /* iterations of the first preconditioner when it failed (0: no fallback) */
  int32_t pad;
  double fallback_relres;
  double ic_shift;          /* Manteuffel shift of the IC(0) factor */
} or_stats;

or_sys *or_build(const or_model *m);             /* pattern + assembled full BCSR + rhs (constraints applied) */
void or_free(or_sys *s);
int32_t or_sys_nn(const or_sys *s);
int64_t or_sys_nb(const or_sys *s);
const int32_t *or_sys_rowptr(const or_sys *s);
const int32_t *or_sys_col(const or_sys *s);
double *or_sys_val(or_sys *s);
double *or_sys_rhs(or_sys *s);
double or_sys_assemble_ms(const or_sys *s);
int32_t or_sys_ntypes(const or_sys *s);
void or_matvec(const or_sys *s, const double *x, double *y); /* y = K x, full BCSR */
/* Jacobi-scaled PCG (src/core/10-solver.js:852-866, 1001-1042); precond 0 = multigrid
 * (src/core/10-solver.js:620-715), 1 = none (Jacobi only), 2 = IC(0) with the Manteuffel
 * shift (src/core/10-solver.js:718-800).  A multigrid or IC(0) run that does not converge
 * continues once with the other preconditioner from the same iterate (src/core/10-solver.js:
 * 960-972); iters counts both.  The values are scaled in place.  -1: IC(0) failed at every shift. */
int or_solve(or_sys *s, double tol, int32_t maxit, int32_t precond, double *u, or_stats *st);

/* ---- post (src/core/12-post.js:30-99, src/core/01-util.js:176-190) ---- */
typedef struct or_warpage {
  double signed_mag, mag;
  int32_t sign;
  double c0, c1, c2, rmin, rmax;
} or_warpage;
void or_fit_plane(const double *x, const double *y, const double *w, int32_t n, double *c);
int or_warpage_jeita(int32_t nx, int32_t ny, int32_t nz, const double *xs, const double *ys, const int32_t *node_id,
                     const double *u, const double *bx, const double *by, int32_t nb, double zx0, double zx1,
                     double zy0, double zy1, or_warpage *out);

/* ---- Anand material point (src/core/13-fatigue.js:32-143) ----
 * P = {A, QR, xi, m, shat, n, h0, a, s0, nu}; state (11): ep[6], s, ebar, W, sigEq, lastD */
double or_anand_step(const double *P, double E, double T, double dt, const double *eps, double *st, double *tr,
                     double *sig, double *D);
void or_anand_batch(const double *P, int64_t n, double E, double T, double dt, const double *eps, double *st,
                    double *tr, double *sig, double *D, double *dep);
void or_anand_batch_robust(const double *P, int64_t n, double E, double T, double dt, const double *eps, double *st,
                    double *tr, double *sig, double *D, double *dep); /* the global Newton's return map (bracket-end guard) */

/* ---- bump screening (screenJoints, src/core/13-fatigue.js:375-455) ----
 * per-step data of the cycle history (step 0 is the start); Cp/Fp NULL without a pillar */
typedef struct or_screen_steps {
  int32_t nsteps, cycles;
  const double *T, *dt, *E, *Fth, *wt, *Cp, *Fp;
  const int32_t *wi, *cycle;
  const uint8_t *same;
  double fs;
} or_screen_steps;
void or_screen_joints(const double *P, const or_screen_steps *h, int64_t nsites, int32_t ntemps, int32_t i25,
                      const double *table, double *dW, double *dE, double *stab);

#ifdef __cplusplus
}
#endif
#endif
