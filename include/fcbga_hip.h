/*
 * fcbga_hip.h -- C ABI of the MI355X (gfx950) hex8 thermoelastic solver.
 *
 * Drop-in for the reference's step solve: "mesh + materials + load step in,
 * displacements out".  The reference has no FFI; its seam is the JS call
 * surface inside the worker driver:
 *   new FEModel(mesh, cfg)                    src/core/09-assembly.js:208-233
 *   model.assembleStep(Ta, Tb, stage, rhs)    src/core/09-assembly.js:392-424
 *   solver.solve(rhs, label, key)             src/core/10-solver.js:933-994
 *   Analysis.solveStep(k)                     src/core/11-analysis.js:189-206
 * fcbga_model_create replaces the FEModel constructor (mesh contract of
 * buildStructuredMesh, src/core/06-mesher.js:66-71), and fcbga_step_solve fuses
 * assembleStep + solve on the device (no assembled matrix) and returns the
 * step increment du with the reference's solve record fields.
 *
 * All functions return 0 on success and a negative code on failure; the
 * message is available from fcbga_last_error (per calling thread).
 * Arrays are borrowed for the duration of the call only.
 * Each model (and submodel) has its own stream and state: calls on different
 * models may run on different host threads at once, calls on one model may
 * not (fcbga_path_copy also reads its source model, which must then be idle).
 */
#ifndef FCBGA_HIP_H
#define FCBGA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FCBGA_OK 0
#define FCBGA_ERR_ARG -1
#define FCBGA_ERR_DEVICE -2
#define FCBGA_ERR_NOT_CONVERGED -3
#define FCBGA_ERR_INTERNAL -4
#define FCBGA_ERR_CANCELLED -5  /* a progress callback asked to stop (the reference's Error('cancelled')) */

typedef struct fcbga_model fcbga_model;

/* Mesh contract of buildStructuredMesh (src/core/06-mesher.js:14-72) plus the
 * 3-2-1 constraint DOFs (src/core/06-mesher.js:112-116) and feature lines. */
typedef struct fcbga_mesh {
  int32_t nx, ny, nz;          /* grid lines in x, y, z */
  int32_t nn, ne;              /* nodes, elements */
  const double *xs, *ys, *zs;  /* nx, ny, nz coordinates (mm) */
  const int32_t *node_id;      /* nx*ny*nz: compact node id (z fastest, then y, then x) or -1 */
  const int32_t *cell_elems;   /* 2*(nx-1)*(ny-1)*(nz-1): up to two superposed elements per cell, -1 none */
  const int32_t *e_group;      /* ne: material group */
  const int8_t *e_part;        /* ne: part code (src/core/06-mesher.js:3-4) */
  const double *e_scale;       /* ne: element scale (fill fraction) */
  const int32_t *e_section;    /* ne or NULL: layered section of an element, -1 homogeneous */
  int32_t ngroups, nsections;
  int32_t nconstraints;
  const int32_t *constraint_dofs; /* constrained DOFs 3*node+component */
  int32_t nfeature_x, nfeature_y;
  const int32_t *feature_x, *feature_y; /* grid-line indices kept on every coarse level */
  /* Stacked-module stages (FEModel.moduleConstraints, src/core/09-assembly.js:227, and
   * assembleStep/moduleStageSet, :392-440): the package constraints plus each module's 3-2-1
   * support on its base.  With nmodule_constraints > 0 a step of stage < 0 is a module stage:
   * only the elements of born groups are assembled (the others are dropped, not scaled by
   * 1e-6), these DOFs and every DOF of a node that no born element touches are held, and the
   * multigrid aggregates are built over the born elements only.  0 / NULL: no module stages. */
  int32_t nmodule_constraints;
  const int32_t *module_constraint_dofs;
} fcbga_mesh;

/* One step of the thermal path [Ta, Tb] in a stage: the per-group data of
 * FEModel.prepareStep (src/core/09-assembly.js:346-365). */
typedef struct fcbga_step {
  double Ta, Tb;
  int32_t stage;
  const double *D;         /* ngroups*36: mean stiffness over the step (Voigt xx yy zz xy yz zx, MPa) */
  const double *s0;        /* ngroups*6: thermal stress increment, integral of D dF over the step */
  const uint8_t *born;     /* ngroups: 1 when the group is born in this stage */
  const double *secW;      /* nsections*108: zeta-station stiffness weights of layered sections (or NULL) */
  const double *secV;      /* nsections*18: zeta-station thermal stress weights (or NULL) */
  const uint8_t *sec_born; /* nsections (or NULL) */
} fcbga_step;

#define FCBGA_PREC_FP64 0  /* FP64 PCG, FP64 multigrid (the plain comparator) */
#define FCBGA_PREC_MIXED 1 /* FP64 flexible PCG, FP32 multigrid cycle, FP64 coarsest */

#define FCBGA_PC_MG 0      /* shell-aggregate multigrid (src/core/10-solver.js:620-715) */
#define FCBGA_PC_COLUMN 1  /* column block Gauss-Seidel sweeps only */
#define FCBGA_PC_JACOBI 2  /* Jacobi scaling only */
#define FCBGA_PC_IC 3      /* IC(0) with the Manteuffel shift (ICPC, src/core/10-solver.js:718-800), FP64 */

/* MG and IC(0) runs that do not converge continue once with the other one from
 * the same iterate (LinearSolver.solve, src/core/10-solver.js:965-973); the
 * record's precond is then the second one and fallback_* describe the first. */

typedef struct fcbga_solve_opts {
  double tol;         /* relative residual in the Jacobi-scaled norm (src/core/10-solver.js:1014) */
  int32_t maxit;
  int32_t precision;  /* FCBGA_PREC_* */
  int32_t precond;    /* FCBGA_PC_* */
  int32_t sweeps;     /* pre- and post-smoothing sweeps per level; 0 = the reference's 2 */
  int32_t warm;       /* warm start (LinearSolver._warmStart/_storeSolution, src/core/10-solver.js:885-926):
                         size of the basis of previous solutions the load is projected onto, and that a
                         converged solution joins (the reference's default is 20); 0 = cold, the basis is
                         neither used nor extended */
  int32_t smoother;   /* FCBGA_SMOOTH_*: the multigrid's smoother on the coarse shell levels */
} fcbga_solve_opts;

/* Smoothers of the multigrid V-cycle.  GS: the reference's column Gauss-Seidel
 * in four colours on every level (MultigridPC, src/core/10-solver.js:512-555,
 * 620-715).  COARSE_JACOBI: the same on the fine level and the first two shell
 * levels; from the third shell level down, 2 x sweeps damped (0.6) column
 * Jacobi sweeps, each one launch over all columns instead of four colour
 * launches (no counterpart in the reference; same solution to the tolerance,
 * iteration counts within one of GS on the bundled workloads). */
#define FCBGA_SMOOTH_GS 0
#define FCBGA_SMOOTH_COARSE_JACOBI 1

/* Solve record of LinearSolver.solve (src/core/10-solver.js:988-993) plus timings. */
typedef struct fcbga_solve_stats {
  int32_t iters;
  int32_t converged;
  double relres;
  double r0;          /* relative residual of the initial guess (iters0: 1 cold, below 1 warm) */
  double ms_total;    /* whole step: assembly + preconditioner set-up + PCG */
  double ms_assemble; /* element tables, load, diagonal blocks */
  double ms_setup;    /* multigrid set-up (numeric) */
  double ms_pcg;      /* PCG iterations */
  int32_t levels;
  int32_t coarsest_dof;
  int64_t ndof;
  int32_t ntypes;     /* distinct condensed element matrices */
  int32_t sym_tables; /* 1: the fine operator ran on the mirror-symmetric compact element tables */
  int32_t warm;       /* warm-start basis size after the solve (the record's warm field) */
  int32_t precond;    /* preconditioner of the final PCG run (the record's precond) */
  int32_t fallback_first;  /* -1: no fallback; else the preconditioner of the run that did not converge */
  int32_t fallback_iters;  /* iterations before the fallback (the record's fallback.iters) */
  double fallback_relres;  /* relative residual when the fallback started (fallback.relres) */
  double ic_shift;         /* Manteuffel shift of the IC(0) factor in use (ICPC.info.shift), 0 without IC */
} fcbga_solve_stats;

/* Live progress from inside the PCG (LIVE.post / liveReport, src/core/10-solver.js:811-849, 1032) and
 * in-solve cancellation (the UI's cancel, src/ui/00-app.js:234-238; Analysis.checkCancel,
 * src/core/11-analysis.js:99).  The callback runs on the calling thread, during the solve call,
 * every `every` PCG iterations (and at the start and end of each PCG run); a non-zero return stops
 * the solve, which then returns FCBGA_ERR_CANCELLED with the message "cancelled".  Per model
 * (fn NULL: none); the state of the last iteration is complete when the callback runs. */
typedef struct fcbga_progress {
  int32_t event;     /* 0 start of a PCG run, 1 iteration, 2 end of the run */
  int32_t iters;     /* iterations so far in this solve (across a fallback) */
  int32_t maxit;
  int32_t precond;   /* FCBGA_PC_* of the running PCG */
  double relres;     /* relative residual in the scaled norm */
  double r0;         /* relative residual of the initial guess */
  double tol;
  double ms;         /* since the start of the PCG run */
} fcbga_progress;
typedef int (*fcbga_progress_fn)(const fcbga_progress *p, void *user);

int fcbga_model_create(const fcbga_mesh *mesh, fcbga_model **out);
int fcbga_set_progress_cb(fcbga_model *m, fcbga_progress_fn fn, void *user, int32_t every);
void fcbga_model_destroy(fcbga_model *m);
/* Empties the warm-start basis (a new path, or a stage with another constraint set). */
int fcbga_warm_reset(fcbga_model *m);
int64_t fcbga_model_ndof(const fcbga_model *m);

/* Step solve with host output: du_out has 3*nn doubles in the reference's node order. */
int fcbga_step_solve(fcbga_model *m, const fcbga_step *step, const fcbga_solve_opts *opts, double *du_out,
                     fcbga_solve_stats *stats);
/* Same, leaving du in device memory (3*nn doubles at du_dev, written on the library's stream
 * and complete on return). */
int fcbga_step_solve_device(fcbga_model *m, const fcbga_step *step, const fcbga_solve_opts *opts, double *du_dev,
                            fcbga_solve_stats *stats);

/* Test hooks: the assembled load f (3*nn, constraints applied) and y = K x with the
 * step's unconstrained, unscaled stiffness (both in the reference's node order). */
int fcbga_step_load(fcbga_model *m, const fcbga_step *step, double *f_out);
int fcbga_step_apply(fcbga_model *m, const fcbga_step *step, const double *x, double *y_out);

/* Stress histories of homogeneous elements (accumulateStep, src/core/09-assembly.js:494-545;
 * stepRecovery :480-492).  State per element, FCBGA_STATE doubles: q (33: 24 nodal + 9
 * incompatible-mode DOFs), then per strain monomial [1, xi, eta, zeta, xi eta, eta zeta,
 * xi zeta] the stress sum S_m (6 each), then the thermal term S_th (6).  One call adds one
 * step of the path: du (3*nn, the step's increment in the reference's node order) with
 * the step's group data; q_only != 0 adds only the DOFs (strain histories).  Elements of
 * layered sections are rejected. */
#define FCBGA_STATE 81
int fcbga_state_accumulate(fcbga_model *m, const fcbga_step *step, const double *du, int64_t nelem,
                           const int32_t *elems, double *states, int32_t q_only);
/* Stress and total strain at natural points (stateStress, src/core/09-assembly.js:556-571;
 * the point() of StateView, src/core/11-analysis.js:356-364): point p lies in element
 * elems[row[p]] whose state is row row[p] of states; xez (3 per point) are its natural
 * coordinates; sig and eps (6 per point, Voigt, engineering shear) receive the results
 * (eps may be NULL). */
int fcbga_state_stress(fcbga_model *m, int64_t nelem, const int32_t *elems, const double *states, int64_t npts,
                       const int32_t *row, const double *xez, double *sig, double *eps);

/* Adjoint sensitivities of one thermal step: the forward/backward pair (autograd.Function
 * semantics) that replaces the finite-difference tornado of the sensitivity study
 * (runSensitivityCase / tornadoFromCases, src/core/14-sensitivity.js:8-172).  Objective
 * components of the step solution u:
 *   J[0] = sum over the existing nodes of grid plane k_top of (w - mean w)^2  (warpage proxy)
 *   J[1] = (sum over the 8 Gauss points of the elements vm_elems of sigma_vm^p)^(1/p)
 *          (peak von Mises as a p-norm; point stresses as stateStress, src/core/09-assembly.js:556-571)
 * fcbga_adjoint_forward solves the step (as fcbga_step_solve, du_out optional), evaluates J
 * and keeps the solution on the model as the pair's saved context.  fcbga_adjoint_backward
 * takes the upstream gradients gJ (2) and returns, from one adjoint PCG solve with the same
 * operator and preconditioner, the log-derivatives of gJ . J: dlnE and dlnA (ngroups) per
 * group's E and CTE (the stiffness and thermal stress increment scale with E, the thermal
 * stress increment with CTE), dlnA_elem (ne, optional) per element's CTE.  Divide by E or
 * CTE for plain derivatives.  Elements of layered sections are not supported in vm_elems. */
typedef struct fcbga_objective {
  int32_t k_top;             /* grid plane of the warpage nodes */
  int32_t n_vm;              /* elements of the von Mises p-norm (0: J[1] = 0) */
  const int32_t *vm_elems;
  double p;                  /* p-norm exponent, >= 1 */
} fcbga_objective;
int fcbga_adjoint_forward(fcbga_model *m, const fcbga_step *step, const fcbga_solve_opts *opts,
                          const fcbga_objective *obj, double *J, double *du_out, fcbga_solve_stats *stats);
int fcbga_adjoint_backward(fcbga_model *m, const double *gJ, double *dlnE, double *dlnA, double *dlnA_elem,
                           fcbga_solve_stats *stats);
/* The same backward pass for a caller's own objective of the step solution: dJdu (3 nn, node
 * order) is dJ/du, the adjoint load as it is (gJ of the built-in objective taken as 0).  Sums
 * over the steps of a thermal path give the derivatives of an objective of a path field
 * (a field is a sum of step increments: each step's adjoint takes the field's dJ/du). */
int fcbga_adjoint_backward_u(fcbga_model *m, const double *dJdu, double *dlnE, double *dlnA, double *dlnA_elem,
                             fcbga_solve_stats *stats);

/* Anand viscoplastic material points, batched (AnandPoint.step + commit,
 * src/core/13-fatigue.js:67-143, 229-236).  Per point: mechanical strain eps (6,
 * engineering shear) at temperature T (C) over dt (s), modulus E(T) (MPa).
 * state (FCBGA_ANAND_STATE per point: ep[6] tensor shear, s, ebar, W, sigEq,
 * lastD) is the committed state; its lastD (the Newton warm start) is updated in
 * place, as the reference's step does.  trial receives the stepped state (copy it
 * over state to commit), sig the stress (6), D (36, optional) the consistent
 * tangent, dep (optional) the equivalent plastic increment. */
#define FCBGA_ANAND_STATE 11
typedef struct fcbga_anand_params {
  double A, QR, xi, m, shat, n, h0, a, s0, nu;
} fcbga_anand_params;
int fcbga_anand_update(const fcbga_anand_params *params, int64_t n, double T, double dt, double E, const double *eps,
                       double *state, double *trial, double *sig, double *D, double *dep);
/* Same on device arrays, on the given HIP stream (NULL: the null stream); complete on return. */
int fcbga_anand_update_device(const fcbga_anand_params *params, int64_t n, double T, double dt, double E,
                              const double *eps, double *state, double *trial, double *sig, double *D, double *dep,
                              void *stream);

/* Bump screening through the temperature cycle (screenJoints, src/core/13-fatigue.js:375-455):
 * every joint a free-standing solder column (AnandPoint.stepColumn, :153-166) or a solder
 * cap on an elastic pillar in series (stepSeriesColumn, :178-206), stepped through the
 * cycle history (cycleHistory, :275-293), one device thread per joint.  The per-step
 * arrays are the reference's precomputation (:394-415); step 0 is the start of the
 * history and only its cycle index is read. */
typedef struct fcbga_screen_steps {
  int32_t nsteps;         /* history points, >= 2 */
  int32_t cycles;         /* >= 1; cycle[k] in 0..cycles */
  const double *T;        /* temperature of step k (C) */
  const double *dt;       /* t[k] - t[k-1] (s) */
  const double *E;        /* solder modulus at T[k] on the 1 C property grid (MPa) */
  const double *Fth;      /* solder thermal strain F(T[k]) - F(25) */
  const double *wt;       /* weight of T[k] in its table interval, clamped to [0, 1] */
  const double *Cp;       /* pillar compliances (1/E_p, 1/G_p, 1/G_p) per step (3 per step), or NULL */
  const double *Fp;       /* pillar thermal strain F_p(T[k]) - F_p(25), or NULL (with Cp) */
  const int32_t *wi;      /* interval of T[k] in the table temperatures (findInterval) */
  const int32_t *cycle;   /* cycle index of step k */
  const uint8_t *same;    /* 1 when step k continues the ramp or dwell of step k-1 */
  double fs;              /* solder fraction of the joint height (1 without a pillar) */
} fcbga_screen_steps;
/* table: nsites x ntemps x 6 bump-layer strains (engineering, siteStrainTable,
 * src/core/13-fatigue.js:318-345; NaN rows for sites without a bump element); i25 the
 * row of 25 C or -1.  Out per site: dW (last cycle's inelastic work density), dE (its
 * equivalent plastic strain), stab (|dW - dW of the cycle before| / dW, NaN below two
 * cycles); NaN for sites without data. */
int fcbga_screen_joints(const fcbga_anand_params *solder, const fcbga_screen_steps *steps, int64_t nsites,
                        int32_t ntemps, int32_t i25, const double *table, double *dW, double *dE, double *stab);
/* fcbga_screen_joints without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  The caller passes the strain table and the three outputs in device memory
 * (nsites x ntemps x 6 and nsites doubles each) and the HIP stream (NULL: the null stream); the
 * history arrays of steps stay on the host and are staged into a static device array of the library
 * (one per device, held until the process ends), so nsteps may not exceed
 * FCBGA_SCREEN_NOALLOC_MAX_STEPS (FCBGA_ERR_ARG otherwise).  Calls that overlap on one device share
 * that array and race.  The results are complete on return and equal fcbga_screen_joints' bit for bit. */
#define FCBGA_SCREEN_NOALLOC_MAX_STEPS 4096
int fcbga_screen_joints_noalloc(const fcbga_anand_params *solder, const fcbga_screen_steps *steps, int64_t nsites,
                                int32_t ntemps, int32_t i25, const double *table_dev, double *dW_dev,
                                double *dE_dev, double *stab_dev, void *stream);

/* Bump submodel (BumpSubmodel, src/core/13-fatigue.js:514-1105): one pitch cell around a
 * bump on a structured grid of B-bar hexahedra (one element per cell), the nodes on the six
 * faces prescribed.  The engine does the time steps (stepTo, :924-961: Newton with the
 * algorithmic tangent and IC(0)-PCG inner solves), the commit, the cutback snapshot and the
 * layer averages on the device; the caller keeps the cycle and the adaptive sub-stepping
 * (run, :963-1090) and the material data of each step. */
#define FCBGA_SUB_SUBSTRATE 0 /* homogenized top of the substrate (elastic, 36 + 6 per step) */
#define FCBGA_SUB_SOLDER 1    /* Anand points */
#define FCBGA_SUB_UNDERFILL 2
#define FCBGA_SUB_SILICON 3
typedef struct fcbga_submodel fcbga_submodel;
typedef struct fcbga_submodel_mesh {
  int32_t nx, ny, nz;           /* grid lines per axis (nodes), z fastest */
  const double *xs, *ys, *zs;
  const int8_t *cell_mat;       /* per cell ((i*(ny-1) + j)*(nz-1) + k): FCBGA_SUB_* */
  const uint8_t *cell_layer;    /* per cell: bit 0 die-side averaging layer, bit 1 substrate-side layer */
} fcbga_submodel_mesh;
typedef struct fcbga_submodel_step_in {
  double T, dt;                 /* end temperature of the step (C) and its duration (s) */
  double E;                     /* solder modulus E(T) (MPa) */
  double fsolder;               /* solder thermal strain F(T) - F(25) */
  const double *elas_D;         /* 4 x 36: mean stiffness of each material over the step from the committed
                                   temperature to T (the solder row is unused) */
  const double *elas_sth;       /* 4 x 6: thermal stress increment over the same step */
  const double *ub;             /* 3 per boundary node (face nodes in grid order): u(T) - u(25) */
} fcbga_submodel_step_in;
typedef struct fcbga_submodel_step_out {
  int32_t converged, newton;    /* the reference's {converged, newton, dMax} */
  int32_t pcg_iters, ic_builds;
  double dmax;                  /* largest equivalent plastic increment over the solder points */
} fcbga_submodel_step_out;
int fcbga_submodel_create(const fcbga_submodel_mesh *mesh, const fcbga_anand_params *solder, fcbga_submodel **out);
void fcbga_submodel_destroy(fcbga_submodel *s);
int32_t fcbga_submodel_nboundary(const fcbga_submodel *s);
int fcbga_submodel_step(fcbga_submodel *s, const fcbga_submodel_step_in *in, fcbga_submodel_step_out *out);
/* committed state for a cutback (_snapshot / _restore, :1069-1079) */
int fcbga_submodel_save(fcbga_submodel *s);
int fcbga_submodel_restore(fcbga_submodel *s);
/* volume-weighted means over the solder points of the two layers: {W die, ebar die, W sub, ebar sub} */
int fcbga_submodel_layer_means(fcbga_submodel *s, double *out4);
/* hysteresis sample: {tau die, gamma die, tau sub, gamma sub} along (cphi, sphi) */
int fcbga_submodel_loop_sample(fcbga_submodel *s, double T, double E, double fsolder, double cphi, double sphi,
                               double *out4);
/* stress change from 25 C at the committed state, extrapolated to the corners of every element
 * (_stressSnapshot, :1036-1062): 48 floats per element (8 corners x 6), E and fsolder at T */
int fcbga_submodel_section(fcbga_submodel *s, double E, double fsolder, float *out);
/* displacement (3 per grid node) */
int fcbga_submodel_displacement(fcbga_submodel *s, double *u);
/* committed material-point states, 8 x FCBGA_ANAND_STATE doubles per cell in cell order, the
 * cell's Gauss points in the element's point order (the states of AnandPoint: ep[6], s, ebar,
 * W, sigEq, and the Newton warm start; zero for cells without solder): what a host that keeps
 * the reference's BumpSubmodel instance reads back into its own points after a device step */
int fcbga_submodel_points(fcbga_submodel *s, double *st);

/* Global viscoplastic Newton (the bump-array cycling workload): a package mesh whose solder group
 * holds discrete bumps with Anand points at the 2x2x2 Gauss points of B-bar elements
 * (AnandPoint.step, src/core/13-fatigue.js:67-143; bbarMatrix, :474-490), the rest of the mesh the
 * step solve's elastic elements.  Each time step is a Newton solve for the total displacement
 * (the time step of BumpSubmodel.stepTo, :924-961: residual and increment tolerances, inexact
 * inner tolerance, a final pass at the converged displacement, then the points are committed) with
 * the step solve's flexible PCG and multigrid built on the current tangent as the inner solver.
 * The reference has no global model of this kind; its Anand points live in the bump submodel. */
typedef struct fcbga_vp fcbga_vp;
typedef struct fcbga_vp_solder {
  int32_t group;              /* mesh group of the solder elements (one element per cell, scale 1) */
  fcbga_anand_params anand;   /* Anand constants; nu is the solder's Poisson ratio */
  int32_t nE;                 /* E(T) table of the solder: temperatures (C) and moduli (MPa), */
  const double *E_T, *E_v;    /* linear between and beyond the points */
  double cte;                 /* solder CTE (1/K) */
  double Tref;                /* stress-free temperature of the solder and start of the history (C) */
} fcbga_vp_solder;
typedef struct fcbga_vp_opts {
  fcbga_solve_opts inner;     /* inner PCG: precision, preconditioner, maxit (tol and warm are set per iteration) */
  int32_t maxit;              /* Newton iterations (the reference's SUB_NEWTON_MAXIT is 25) */
  double res_tol;             /* |r| / (|f_load| + |f_solder| + |K_el u|), K_el the elastic elements' part
                               * of the operator (the reference's SUB_RES_TOL is 1e-6) */
  double du_tol;              /* |du| / |u| (SUB_DU_TOL 1e-8) */
  double inner_floor;         /* floor of the inexact inner tolerance (SUB_INNER_TOL 1e-8) */
} fcbga_vp_opts;
typedef struct fcbga_vp_stats {
  int32_t converged, newton, pcg_iters;
  int32_t nonfinite;          /* 1: a point's stress turned non-finite (the caller cuts the step back) */
  double dmax;                /* largest equivalent plastic increment of the step */
  double res;                 /* last relative residual */
  double ms, ms_setup, ms_pcg;
} fcbga_vp_stats;
int fcbga_vp_create(const fcbga_mesh *mesh, const fcbga_vp_solder *solder, fcbga_vp **out);
void fcbga_vp_destroy(fcbga_vp *v);
int32_t fcbga_vp_nsolder(const fcbga_vp *v);
/* One time step to temperature T over dt (s).  elastic: the step data of every group for the
 * total load from the stress-free state (Ta = Tref, Tb = T; the solder group's D is its elastic
 * stiffness and its s0 is ignored).  A step that does not converge leaves the points uncommitted;
 * the caller restores the saved state and cuts the step. */
int fcbga_vp_step(fcbga_vp *v, const fcbga_step *elastic, double T, double dt, const fcbga_vp_opts *opts,
                  fcbga_vp_stats *stats);
int fcbga_vp_save(fcbga_vp *v);
int fcbga_vp_restore(fcbga_vp *v);
/* prescribed displacement of the constrained DOFs (nconstraints values in the mesh's constraint_dofs
 * order; zero until set), applied by every later step: the boundary of a bump submodel, whose six
 * faces follow the global field (BumpSubmodel.prescribe, src/core/13-fatigue.js:646-693) */
int fcbga_vp_set_prescribed(fcbga_vp *v, const double *values);
/* committed point states, 8 x FCBGA_ANAND_STATE per solder element in the order of elems (the mesh
 * element of each solder element, nsolder entries), the points of an element in corner order */
int fcbga_vp_points(fcbga_vp *v, double *st, int32_t *elems);
/* total displacement (3*nn, the reference's node order) */
int fcbga_vp_displacement(fcbga_vp *v, double *u);
/* test hooks: replace the displacement; evaluate at the current displacement without committing:
 * the out-of-balance force r (3*nn) and, for x (3*nn, optional), y = K_t x with the tangent there */
int fcbga_vp_set_displacement(fcbga_vp *v, const double *u);
int fcbga_vp_residual(fcbga_vp *v, const fcbga_step *elastic, double T, double dt, double *r, const double *x,
                      double *y);

/* Thermal path on the device (Analysis.run / finishPath, src/core/11-analysis.js:212-261): the
 * increment of each solved step stays in device memory (FP64); fields are sums of increments.
 * fcbga_path_solve solves step k of the plan (as fcbga_step_solve, opts->warm carrying the
 * warm-start basis along the path) and keeps its increment; fcbga_path_sum writes the sum of the
 * increments of steps ids[0..n-1], added in that order (3*nn doubles, host); fcbga_path_du the
 * FP32 increment of step k, the form the reference keeps for the stress histories. */
int fcbga_path_reset(fcbga_model *m);
int fcbga_path_solve(fcbga_model *m, int32_t k, const fcbga_step *step, const fcbga_solve_opts *opts,
                     fcbga_solve_stats *stats);
int fcbga_path_sum(fcbga_model *m, int32_t n, const int32_t *ids, double *out);
int fcbga_path_du(fcbga_model *m, int32_t k, float *out);
/* Analysis.restorePath (src/core/11-analysis.js:220-229): step k's increment from a results file
 * (3*nn FP32, as exportAnalysisState saves them, src/core/16-results-file.js:121-132) replaces its
 * solve; fcbga_path_sum then sums the restored increments as finishPath does. */
int fcbga_path_load(fcbga_model *m, int32_t k, const float *du);
/* The reference's worker farm (runSweep / evalSteps, src/worker.js:76-133) solves blocks of a
 * plan's steps on several models and sends each step's FP64 increment back (the partial {k, du}
 * message, src/worker.js:89), which finishPath then sums.  fcbga_path_load64 keeps such an
 * increment (3*nn FP64, host) as step k; fcbga_path_copy copies step k's kept increment from
 * another model of the same mesh on the same device (device to device). */
int fcbga_path_load64(fcbga_model *m, int32_t k, const double *du);
int fcbga_path_copy(fcbga_model *dst, const fcbga_model *src, int32_t k);

/* Step material data from 1 C property grids: FEModel.prepareStep (src/core/09-assembly.js:346-365)
 * with gridMean / gridStressIncrement (src/core/03-materials.js:370-409) and sectionWmean /
 * sectionVstep (src/core/04-homogenize.js:206-221).  A grid holds FCBGA_TGRID_N points, one per
 * degree C from FCBGA_TGRID_MIN, linear between them and clamped outside.  Host only: it fills
 * the fcbga_step arrays of a step [Ta, Tb] in `stage`. */
#define FCBGA_TGRID_MIN (-80)
#define FCBGA_TGRID_N 381
typedef struct fcbga_group_grids {
  int32_t stage;        /* birth stage: born in every stage >= stage (FEModel.isBorn) */
  int32_t layered;      /* 1: a layered group (its elements read their sections; D, s0 stay 0) */
  const double *Dgrid;  /* FCBGA_TGRID_N x 36 stiffness (MPa), or NULL for a constant one */
  const double *Dconst; /* 36, when Dgrid is NULL */
  const double *F;      /* FCBGA_TGRID_N x 3 normal thermal strains from the strain origin */
} fcbga_group_grids;
typedef struct fcbga_section_grids {
  int32_t ns;           /* sub-layers */
  int32_t group;        /* group whose birth the section follows, -1: always born */
  const double *W;      /* FCBGA_TGRID_N x 108 zeta-station stiffness weights */
  const double *D;      /* ns x FCBGA_TGRID_N x 36 sub-layer stiffness */
  const double *F;      /* ns x FCBGA_TGRID_N x 3 sub-layer thermal strains */
  const double *Ls;     /* ns x 3 Lagrange integrals of the sub-layer's thickness range */
} fcbga_section_grids;
/* D (ngroups*36), s0 (ngroups*6), born (ngroups), secW (nsections*108), secV (nsections*18) and
 * sec_born (nsections) of the step: the mean stiffness (a constant one as it is), the thermal
 * stress increment of born groups (0 for the unborn), and per section the mean weights, the
 * thermal stress weights (0 when unborn) and the birth. */
int fcbga_prepare_step(int32_t ngroups, const fcbga_group_grids *groups, int32_t nsections,
                       const fcbga_section_grids *sections, double Ta, double Tb, int32_t stage, double *D,
                       double *s0, uint8_t *born, double *secW, double *secV, uint8_t *sec_born);

/* meshFingerprint (src/core/16-results-file.js:104-119): the 14-hex-digit hash a results file
 * carries of the mesh its increments belong to, over the bytes of xs, ys, zs (FP64), eGroup
 * (int32), ePart (int8) and eScale (FP64), seeded with nn and ne.  Host only; out receives 14
 * lower-case hex digits and a NUL. */
int fcbga_mesh_fingerprint(int32_t nn, int32_t ne, const double *xs, int32_t nx, const double *ys, int32_t ny,
                           const double *zs, int32_t nz, const int32_t *e_group, const int8_t *e_part,
                           const double *e_scale, char out[15]);

/* Measurement hook: with timing enabled, launches of the kernel classes below are
 * bracketed by HIP events on the library's stream and their device time is summed
 * per class (reset by every fcbga_set_kernel_timing call).  The reference has no
 * counterpart; its timings are the wall-clock fields of the solve record
 * (src/core/10-solver.js:988-993). */
#define FCBGA_KCLASS_APPLY 0        /* FP64 operator apply of the outer PCG (one launch) */
#define FCBGA_KCLASS_FINE_SWEEP 1   /* fine colour-sweep residual (one launch per colour) */
#define FCBGA_KCLASS_FINE_RESID 2   /* fine residual before restriction (one launch) */
#define FCBGA_KCLASS_SHELL1_SWEEP 3 /* first shell level: 2 sweeps x 4 colours (bracketed together) */
#define FCBGA_KCLASS_SHELL_SWEEP 4  /* coarser shell levels: 2 sweeps x 4 colours (bracketed together) */
#define FCBGA_KCLASS_COARSEST 5     /* coarsest dense FP64 product (one launch) */
#define FCBGA_KCLASS_IC_APPLY 6     /* IC(0) forward + backward solve (one launch per level each, bracketed together) */
#define FCBGA_KCLASS_IC_FACTOR 7    /* IC(0) factor, one attempt (one launch per level, bracketed together) */
#define FCBGA_KCLASS_COUNT 8
int fcbga_set_kernel_timing(fcbga_model *m, int enable);
int fcbga_kernel_times(const fcbga_model *m, double *ms_total /*[FCBGA_KCLASS_COUNT]*/,
                       int32_t *launches /*[FCBGA_KCLASS_COUNT]*/);

int fcbga_last_error(char *buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
