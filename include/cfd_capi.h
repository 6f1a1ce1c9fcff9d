/*
 * cfd_capi.h -- C-ABI of the MI355X (gfx950) hot path of the staggered-grid
 * incompressible Navier-Stokes solver: QUICK convection-diffusion predictor,
 * direct-forcing immersed boundary (DFIB) forcing and velocity correction,
 * pressure right-hand side, ELL / matrix-free ("SPE") pressure operator,
 * BiCGSTAB and red-black SOR pressure solves, boundary conditions, and the
 * prediction-correction timestep that sequences them.
 *
 * Plain C: no C++ or torch types cross this boundary.  Every entry point
 * returns an int status (CFD_OK == 0, negative on error); the message of the
 * last failure on the calling thread is available from cfd_last_error().
 *
 * Array layouts follow the reference solver exactly:
 *   ghosted fields (u, v, w, p, eta): length G = nx*ny*nz with nx = nx_cal+4
 *     (two ghost layers per side), index icel(i,j,k) = (i*ny + j)*nz + k
 *     (inc/grid/structureGrid.hpp:159-161); u(i,j,k) lives on the +x face of
 *     cell (i,j,k), v on the +y face, w on the +z face.
 *   interior vectors (pressure solution, RHS, jp): length N = nx_cal*ny_cal*
 *     nz_cal, index icelCal = ((i-2)*ny_cal + (j-2))*nz_cal + (k-2)
 *     (inc/grid/structureGrid.hpp:182-185).
 *   DFIB force f and the accumulated source term: 3*N, component-major
 *     (inc/resize_and_simuSetting.hpp:20, inc/dfib/updateUandF.hpp:112-114).
 *   ELL matrix: slot-major structure of arrays col[s*N + row] (int32),
 *     val[s*N + row] (fp64), s in [0, 7); slot 0 is the diagonal, slots 1..6
 *     hold xm, xp, ym, yp, zm, zp with absent boundary neighbours skipped and
 *     padding (col = N-1, val = 0) at the end of each row, exactly the slot
 *     order mat::ELL_matrix<T>::set produces (inc/matrix/PressureMatrix.hpp:71-89,
 *     inc/matrix/ELL_sparseMatrix.hpp:146-160, 246-320).  This is the layout
 *     of every ELL array crossing the ABI; the library's own device copy is
 *     stored in 64-row slices and cfd_ell_download / cfd_ell_upload convert.
 *
 * Threading: a context owns one HIP stream and its device buffers; calls on
 * one context are stream-ordered and must come from one thread at a time.
 * Calls that return host scalars synchronise the stream.  Distinct contexts
 * are independent.
 */
#ifndef CFD_CAPI_H
#define CFD_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFD_ABI_VERSION 6

/* status codes */
#define CFD_OK 0
#define CFD_ERR_ARG (-1)      /* invalid argument (reference: std::invalid_argument) */
#define CFD_ERR_HIP (-2)      /* HIP runtime failure */
#define CFD_ERR_STATE (-3)    /* call out of order (e.g. solve before assembly) */
#define CFD_ERR_NOMEM (-4)    /* device allocation failed */
#define CFD_ERR_RUNTIME (-5)  /* reference: std::runtime_error */

/* pressure solvers (inc/controlPanel.hpp:17-29) */
#define CFD_SOLVER_BICGSTAB 0 /* solver::bicgstabRe2 */
#define CFD_SOLVER_RBSOR 1    /* red-black SOR on the ELL (or CSR) operator, ||b - Ax|| / ||b|| stop */
#define CFD_SOLVER_SOR_PIPELINE 2 /* SorPipeLine (P_SOLVER_SOR, SorPipeLine.seq.hpp:4-122): the volume-weighted
                                     form on the ghosted pressure in place, stop on max |dp| <= tol
                                     (SOR_ERROR_MAXVAL); red-black order; FP64 */

/* pressure operators */
#define CFD_OP_ELL 0 /* mat::ELL_matrix (P_SOLVER_BICG_ELL) */
#define CFD_OP_SPE 1 /* mat::SPE_matrix1, matrix-free (P_SOLVER_BICG_SPE, the reference default) */
#define CFD_OP_CSR 2 /* mat::CSR_matrix from grid::createPressureMatrixCSR (P_SOLVER_BICG_CSR) */

/* convection schemes of ConvectionDifussion (inc/controlPanel.hpp:125-129) */
#define CFD_CONV_QUICK 0 /* QuickUnitA / QuickUnitB (convection_and_difussion.hpp:20-84), the reference's default */
#define CFD_CONV_LUD 1   /* LudUnitA / LudUnitB, linear upwind (:90-144) */
#define CFD_CONV_UD 2    /* UdUnit, first-order upwind (:148-165) */

/* arithmetic of the pressure solve */
#define CFD_FP64 0
#define CFD_MIXED 1 /* FP32 inner Krylov/SOR solves inside FP64 defect correction */
/* mixed-precision schemes */
#define CFD_MIXED_RELIABLE 0  /* FP32 BiCGSTAB with FP64 residual replacement (reliable updates) */
#define CFD_MIXED_RESTARTED 1 /* FP64 defect correction around restarted FP32 inner solves */
#define CFD_MIXED_VALUES32 2  /* FP64 BiCGSTAB on the FP32-valued ELL operator, FP64 residual replacement */
#define CFD_MIXED_PC32 3      /* with CFD_PRECOND_MG: FP64 BiCGSTAB whose multigrid V-cycle runs in FP32 */

/* preconditioning of the BiCGSTAB pressure solve */
#define CFD_PRECOND_JACOBI 0 /* the reference: Jacobi scaling only (bicgstabRe2 on the jp-scaled operator) */
#define CFD_PRECOND_MG 1     /* right preconditioning by a geometric-multigrid V-cycle (no reference
                                counterpart); with CFD_MIXED the V-cycle and the inner Krylov solve run in
                                FP32 inside the FP64 defect correction */

/* device arrays addressable through upload / download */
#define CFD_T0_U 0 /* old velocity (ghosted) */
#define CFD_T0_V 1
#define CFD_T0_W 2
#define CFD_T1_U 3 /* predicted velocity u* (ghosted) */
#define CFD_T1_V 4
#define CFD_T1_W 5
#define CFD_T3_U 6 /* corrected velocity (ghosted) */
#define CFD_T3_V 7
#define CFD_T3_W 8
#define CFD_P 9      /* pressure (ghosted) */
#define CFD_ETA 10   /* solid fraction (ghosted) */
#define CFD_F 11     /* DFIB force, 3*N */
#define CFD_SRC 12   /* accumulated source term source_termF, 3*N */
#define CFD_RHS 13   /* matB, N */
#define CFD_X 14     /* X_result, N */
#define CFD_JP 15    /* Jacobi scaling 1/diag, N */
#define CFD_VEC_A 16 /* scratch vectors for standalone operator calls, N */
#define CFD_VEC_B 17
#define CFD_NFIELDS 18

typedef struct cfd_grid_desc {
  int nx_cal, ny_cal, nz_cal; /* interior cells per axis (grid::nxCal..) */
  const double *X, *Y, *Z;    /* node coordinates, n+1 entries (grid::X..) */
  const double *Dx, *Dy, *Dz; /* cell widths, n entries (grid::Dx..) */
  const double *Dxs, *Dys, *Dzs; /* centre-to-centre distances, n entries (grid::Dxs..) */
  double lx, ly, lz;          /* domain lengths (grid::lx..) */
} cfd_grid_desc;

typedef struct cfd_params {
  double Re;             /* nu = 1/Re (inc/simuClass.hpp:39-44) */
  double dt;             /* simuClass::dt */
  double bc_dir[6][3];   /* velocity Dirichlet table Dir[face][component] (Boundary_Condition.hpp:4-31) */
  double bc_num[6][3];   /* velocity Neumann table Num[face][component] */
  int p_dirichlet[6];    /* pressure face type: 0 Neumann (P_NEUMANN.hpp), 1 Dirichlet p = 0 */
  int source_index_quirk;/* 1: source_termF[icelCal] added at T1[icelCal] as in
                            inc/run/RUN_CPU_PC_METHOD.hpp:314,465; 0: added at T1[icel] */
  int inner_loops;       /* predictor passes before the corrector: 3 (RUN_CPU_PC_METHOD.hpp:291), 0 (RUN_CPU.hpp) */
  int solver;            /* CFD_SOLVER_* */
  int op;                /* CFD_OP_* */
  int precision;         /* CFD_FP64 / CFD_MIXED */
  double tol;            /* simuClass::p_criteria; relative residual for BiCGSTAB / RB-SOR */
  double tol_inner;      /* CFD_MIXED, restarted mode: tolerance of each FP32 inner solve */
  int mixed_mode;        /* CFD_MIXED: CFD_MIXED_RELIABLE (default), _RESTARTED, _VALUES32 or _PC32 */
  double mixed_delta;    /* CFD_MIXED_RELIABLE: FP64 residual replacement when the FP32
                            residual has dropped by this factor (0.1) */
  int maxit;             /* bicgstabRe2::iters_max_ (5000) / RB-SOR sweep cap */
  double sor_omega;      /* simuClass::p_sor_omega (1.8) */
  int sor_check;         /* RB-SOR: sweeps between true-residual checks */
  int ib_on;             /* DFIB on (simuClass::DfibMethod != "OFF") */
  double ib_cx, ib_cy, ib_r; /* z-cylinder centre and radius (setEta.hpp:145-147) */
  int ib_nsub;           /* sub-cells per cut-cell edge (setEta.hpp:152) */
  double u_init[3];      /* initial velocity (RUN_CPU_PC_METHOD.hpp:77) */
  int precond;           /* CFD_PRECOND_JACOBI (default) or CFD_PRECOND_MG */
  int mg_nu;             /* CFD_PRECOND_MG: red-black Gauss-Seidel sweeps before and after each coarse
                            correction (default 4) */
  int convection;        /* face values of the predictor's convection term: CFD_CONV_* */
  int les;               /* 1: Smagorinsky eddy viscosity (C_s 0.18) added to nu in the predictor's diffusion
                            (smagorinskyModel.hpp:30-79, TERBULENCE_SMAGORINSKY); 0: off, the reference's default */
  /* boundary conditions the reference's Dirichlet/Neumann tables cannot
     express (no reference counterpart; 0 keeps the table semantics):
     periodic_z      1: velocity and pressure periodic across the z faces
                     (faces 4 and 5 of the tables are ignored; nz_cal even,
                     at least 4).  The last z face of w is a real face,
                     predicted and projected like the others; the pressure
                     operator wraps in z (ELL and matrix-free; not CSR).
     outflow_convective 1: face 1 (x+) is a convective outflow
                     d(phi)/dt + U d(phi)/dx = 0 for u, v and w, applied by
                     BC_staggered_main after each step as an implicit upwind
                     step phi_b = (phi_b + s phi_in) / (1 + s), s = U dt / h;
                     the outflow face of u is projected with the Dirichlet
                     pressure p = 0 there (p_dirichlet[1] must be 1) and
                     carries no immersed-boundary forcing.
     outflow_speed   U of the convective outflow (> 0) */
  int periodic_z;
  int outflow_convective;
  double outflow_speed;
} cfd_params;

/* per-timestep report; pass index 0..inner_loops-1 are predictor passes, the last is the corrector */
#define CFD_MAX_PASSES 8
typedef struct cfd_step_stats {
  int loop;                       /* simuClass::loop of the step just taken (1-based) */
  int n_passes;
  int iters[CFD_MAX_PASSES];      /* solver iterations per pass, -1 when no solve (loop == 1) */
  double relres[CFD_MAX_PASSES];  /* final relative residual per pass */
  double cd[CFD_MAX_PASSES];      /* -2 sum(f_x dV)/lz per pass (RUN_CPU_PC_METHOD.hpp:403-435) */
  double cl[CFD_MAX_PASSES];
  double maxdiff[3];              /* CheckSteadyStatebyMaxVal (CheckSteadystate.hpp:165-219) */
  int inner_iters[CFD_MAX_PASSES];/* CFD_MIXED: total FP32 inner iterations per pass */
  /* FP64 post-projection divergence check of the corrector pass, in the
     finite-volume form of the pressure right-hand side (b_matrix.hpp:68-80):
     the corrected velocity on the interior faces and the predicted velocity
     T1 on the boundary faces (index 1 and n-3 along each component's axis,
     the faces the projection saw).  max |div u| and sqrt(sum (div u)^2)
     over every interior cell, and the same over the fluid cells (eta zero
     at the cell and its six neighbours: no face carries immersed-boundary
     forcing, so div u is what the pressure solve left).  0 in a step
     without a pressure solve is not implied: loop 1 reports the predicted
     field's divergence after the forcing. */
  double div_max, div_l2, div_fluid_max, div_fluid_l2;
} cfd_step_stats;

typedef struct cfd_ctx cfd_ctx;

/* ---- library ---------------------------------------------------------- */
int cfd_abi_version(void);
const char *cfd_last_error(void);
/* default parameters: the reference's native settings (src/CFD_MX.cpp:35-57,
   inc/simuClass.hpp:74-80, BC_Cavity_A.hpp table, all-Neumann pressure) */
int cfd_params_default(cfd_params *out);
/* number of visible HIP devices (0 on a host without GPU) */
int cfd_device_count(int *n);

/* ---- grid generation (host only; one axis each; X has n+1 entries, D and Ds
   n entries, n = ncal + 4) ------------------------------------------------- */
/* grid::init_uniform (inc/grid/structureGrid.hpp:409-440) */
int cfd_axis_uniform(int ncal, double L, double *X, double *D, double *Ds);
/* grid::unit_SML + init_gC_dx (structureGrid.hpp:326-386, 506-559); the
   reference's std::invalid_argument cases return CFD_ERR_ARG */
int cfd_axis_sml(int ncal, double L, double lS, int nS, int off_nS, double *X, double *D, double *Ds);
/* stretched axis of the large configuration (no reference counterpart):
   width h0 over a band of width `band` around xc, widths growing by `ratio`
   per cell outside it, capped so that the widths sum to L; then init_gC_dx */
int cfd_axis_tanh(int ncal, double L, double xc, double h0, double band, double ratio, double *X, double *D,
                  double *Ds);

/* ---- context ---------------------------------------------------------- */
/* Allocates device buffers for the grid; replaces the reference's
   resize_variable (inc/resize_and_simuSetting.hpp:5-31) and grid metric upload. */
int cfd_create(const cfd_grid_desc *grid, const cfd_params *params, int device, cfd_ctx **out);
int cfd_destroy(cfd_ctx *ctx);
int cfd_set_params(cfd_ctx *ctx, const cfd_params *params);
int cfd_sizes(const cfd_ctx *ctx, size_t *G, size_t *N);

/* host <-> device copies of the arrays named by CFD_* field ids; len is in elements */
int cfd_upload(cfd_ctx *ctx, int field, const double *host, size_t len);
int cfd_download(cfd_ctx *ctx, int field, double *host, size_t len);

/* ---- set-up ----------------------------------------------------------- */
/* T0 = T1 = T3 = u_init (+ seeded uniform noise of amplitude noise_amp on
   interior cells), p = 0, X = 0, then BC_staggered_main + BC_staggered_copy
   (RUN_CPU_PC_METHOD.hpp:77-119). noise_amp = 0 reproduces the reference. */
int cfd_init_state(cfd_ctx *ctx, double noise_amp, uint64_t seed);
/* DFIB_CylinderZ (inc/dfib/setEta.hpp:136-204); eta = 0 everywhere when ib_on == 0 */
int cfd_set_eta_cylinder_z(cfd_ctx *ctx);
/* createPressureMatrix (inc/matrix/PressureMatrix.hpp:5-117): ELL values, jp,
   plus the FP32 value copy when precision == CFD_MIXED */
int cfd_ell_assemble(cfd_ctx *ctx);
/* SPE_matrix1::setupPressure + set_Jp_pc (inc/matrix/SPE_sparseMatrix1.hpp:40-121) */
int cfd_spe_assemble(cfd_ctx *ctx);
/* grid::createPressureMatrixCSR (inc/grid/structureGrid.hpp:630-728) into
   mat::CSR_matrix::set (inc/matrix/CSR_sparseMatrix.hpp:42-120): row
   pointers, columns and values of the Jacobi-scaled operator with the
   builder's own diagonal summation order and Neumann closure on every face;
   its jp becomes gA.jp when op == CFD_OP_CSR (RUN_CPU_PC_METHOD.hpp:153-160) */
int cfd_csr_assemble(cfd_ctx *ctx);
/* copies of the CSR operator: ptr n_rows + 1 entries, col/val nnz entries
   (nnz = ptr[n_rows]; cfd_csr_nnz reports it) */
int cfd_csr_nnz(cfd_ctx *ctx, size_t *nnz);
int cfd_csr_download(cfd_ctx *ctx, int32_t *ptr, int32_t *col, double *val, size_t n_rows, size_t nnz);
int cfd_ell_download(cfd_ctx *ctx, int32_t *col, double *val, size_t n_rows, int k);
int cfd_ell_upload(cfd_ctx *ctx, const int32_t *col, const double *val, size_t n_rows, int k);

/* ---- operators -------------------------------------------------------- */
/* y = A x on device vectors named by field ids (N-sized), op = CFD_OP_*,
   precision = CFD_FP64 or CFD_MIXED (FP32 values and vectors, result widened)
   (ELL_matrix::multiply_omp, ELL_sparseMatrix.hpp:477-492;
    SPE_matrix1::spmv_ompJP, SPE_sparseMatrix1.hpp:126-144;
    CSR_matrix::multiply_omp, CSR_sparseMatrix.hpp:278-292; FP64 only) */
int cfd_spmv(cfd_ctx *ctx, int op, int precision, int x_field, int y_field);

/* ---- pressure solves (matB -> X_result, warm start from X_result) ------ */
/* solver::bicgstabRe2::solve (inc/solver/bicgstabRe2.hpp:91-189), FP64 */
int cfd_bicgstab(cfd_ctx *ctx, double tol, int maxit, int *iters, double *relres);
/* red-black SOR on the ELL operator (on the CSR operator when params.op is
   CFD_OP_CSR) to ||b-Ax||/||b|| < tol, checked every check_every sweeps;
   iters = sweeps taken */
int cfd_rbsor(cfd_ctx *ctx, double omega, double tol, int maxit, int check_every, int *iters, double *relres);
/* SorPipeLine_seq (inc/pressure/solver/SorPipeLine.seq.hpp:4-122) on the
   context's predicted velocity T1 and ghosted pressure p (in place):
   cf6 = mChange / dt of T1 (written to CFD_RHS), sweeps of the
   volume-weighted update p += omega (pNEW - p) in red-black order until
   max |pNEW - p| over a sweep is at most crit or itmax sweeps ran; iters =
   sweeps - 1 (simu.iters), error = the last sweep's max |pNEW - p|
   (simu.error).  Boundary neighbours read the ghost pressure as the last
   BC left it. */
int cfd_sor_pipeline(cfd_ctx *ctx, double omega, double crit, int itmax, int *iters, double *error);
/* FP64 defect correction around FP32 BiCGSTAB (or RB-SOR) inner solves */
int cfd_defect_correction(cfd_ctx *ctx, double tol, double tol_inner, int max_outer,
                          int *outer, int *inner_total, double *relres);
/* the configured solve (params.solver / precision / op): pSolver(Mx.matB, Mx.X_result) */
int cfd_solve_pressure(cfd_ctx *ctx, int *iters, double *relres);

/* ---- timestep phases (reference semantics, one call = one reference call) */
/* ConvectionDifussion (inc/source/convection_and_difussion.hpp:168-446): T0 -> T1 */
int cfd_predictor(cfd_ctx *ctx);
/* T1 += source_termF (RUN_CPU_PC_METHOD.hpp:309-316), index per source_index_quirk */
int cfd_source_add(cfd_ctx *ctx);
/* createBMatrix_omp (inc/matrix/b_matrix.hpp:51-102): T1 -> matB */
int cfd_build_rhs(cfd_ctx *ctx);
/* Pressure_transform_X_result (inc/pressure/Transform.hpp:8-24): X_result -> p */
int cfd_pressure_scatter(cfd_ctx *ctx);
/* update_UandF_omp (inc/dfib/updateUandF.hpp:95-192): T1, p, eta -> T3, f */
int cfd_dfib_update(cfd_ctx *ctx);
/* source_termF += dt * f (RUN_CPU_PC_METHOD.hpp:393-401) */
int cfd_source_accumulate(cfd_ctx *ctx);
/* BC_staggered_main (Boundary_Condition.hpp:37-175) on T0 and p */
int cfd_bc_main(cfd_ctx *ctx);
/* BC_staggered_copy (updateBC.hpp:6-109): boundary faces of T0 -> T1 */
int cfd_bc_copy(cfd_ctx *ctx);
/* virtualF_Int x and y (virtualForceIntergrator.hpp:5-37): cD, cL of f */
int cfd_drag(cfd_ctx *ctx, double *cd, double *cl);
/* virtualF_Int (virtualForceIntergrator.hpp:5-37) for one direction: the sum
   over the interior of f[which_direction] times the cell volume (the value
   the reference stores in DfibArray::ValSum[whichDirection]) */
int cfd_force_sum(cfd_ctx *ctx, int which_direction, double *sum);
/* CheckSteadyStatebyMaxVal_omp (CheckSteadystate.hpp:165-219): max(T0 - T3) per component */
int cfd_maxdiff(cfd_ctx *ctx, double out[3]);
/* velocity::copy_omp T0 <- T3 (physicalVariables.hpp:84-102) */
int cfd_copy_t3_to_t0(cfd_ctx *ctx);
/* FP64 divergence of the velocity triple starting at vel_field (CFD_T0_U,
   CFD_T1_U or CFD_T3_U), every face from that triple, in the form of
   cfd_step_stats: out[0] max |div u|, out[1] sqrt(sum (div u)^2) over every
   interior cell, out[2] and out[3] the same over the fluid cells */
int cfd_divergence(cfd_ctx *ctx, int vel_field, double out[4]);

/* ---- whole timestep --------------------------------------------------- */
/* one iteration of runSeqOmp_predictionMethod's main loop
   (inc/run/RUN_CPU_PC_METHOD.hpp:287-627), or of runSeqOmp (RUN_CPU.hpp:290-460)
   when inner_loops == 0; advances simuClass::loop and time. */
int cfd_step(cfd_ctx *ctx, cfd_step_stats *stats);
int cfd_loop(const cfd_ctx *ctx, int *loop, double *time);
int cfd_set_loop(cfd_ctx *ctx, int loop, double time);

/* bicgstabRe2's counters over the context's FP64 BiCGSTAB solves
   (inc/solver/bicgstabRe2.hpp:80-84): solves is timestep_ (the restart
   rule's clock, :98, 171), restarts numRestarts_ (shadow-residual restarts,
   :177-182), resets pResets_ (solves ended by the NaN / Inf reset,
   :143-163), iters totalNumIter_ */
/* per-phase device time of cfd_step, the reference's clockstruct stopwatches
   (inc/profiling/STL_clock.hpp:35-47) around the same parts of the loop
   (inc/run/RUN_CPU_PC_METHOD.hpp:296-607), as HIP events on the context
   stream.  checkL2norm, updateT3toT0 and getCdCl are fused into the update
   kernel, so their time is in CFD_PHASE_UPDATE.  Off by default;
   cfd_phase_timing switches it and clears the totals. */
#define CFD_PHASE_CONVECTION 0 /* convectionDifussion: predictor and source add */
#define CFD_PHASE_PRESSURE 1   /* pressure: right-hand side, solve, scatter */
#define CFD_PHASE_UPDATE 2     /* updateT1toT3 (+ checkL2norm, updateT3toT0, getCdCl) */
#define CFD_PHASE_BC 3         /* BC */
#define CFD_NPHASES 4
int cfd_phase_timing(cfd_ctx *ctx, int on);
/* milliseconds per phase summed over the timed steps, and their number */
int cfd_phase_times(cfd_ctx *ctx, double ms[CFD_NPHASES], int *steps);

typedef struct cfd_solver_counters {
  int solves, restarts, resets, pad;
  long long iters;
} cfd_solver_counters;
int cfd_get_solver_counters(cfd_ctx *ctx, cfd_solver_counters *out);

/* ---- discrete adjoint (no reference counterpart) ---------------------- */
/* Objective J = mean of the corrector pass's cD (virtualF_Int,
   virtualForceIntergrator.hpp:5-37) over steps [avg_from, nsteps) of a run
   of cfd_step from the current state.  Parameters: the inflow profile u_in(y)
   -- the face-0 u Dirichlet value of BC_staggered_main per interior row,
   replacing bc_dir[0][0] -- and the DFIB forcing gain, u_new = (1 - gain *
   eta_f) T2 in update_UandF (gain 1 reproduces the reference bit for bit).
   Needs the ELL operator, BiCGSTAB and source_index_quirk = 0. */
typedef struct cfd_adjoint_stats {
  int steps;             /* forward steps of the objective run */
  int checkpoints;       /* states kept by the forward run */
  int recomputed_steps;  /* forward steps re-run by the reverse sweep */
  int adjoint_solves;    /* transposed pressure solves */
  int adjoint_iters;     /* their BiCGSTAB iterations */
  double forward_ms;     /* device time of the objective run */
  double backward_ms;    /* device time of the reverse sweep (recomputation included) */
} cfd_adjoint_stats;
/* u_in has ny_cal values; NULL restores the table value bc_dir[0][0] */
int cfd_set_inflow_profile(cfd_ctx *ctx, const double *u_in, size_t n);
int cfd_set_forcing_gain(cfd_ctx *ctx, double gain);
/* run nsteps timesteps keeping a checkpoint every ckpt_every steps; *J receives the objective */
int cfd_adjoint_forward(cfd_ctx *ctx, int nsteps, int avg_from, int ckpt_every, double *J);
/* dJ/du_in (n = ny_cal values) and dJ/dgain of the last cfd_adjoint_forward
   run, by the transposed timestep with recomputation from the checkpoints;
   leaves the state where the forward run ended */
int cfd_adjoint_backward(cfd_ctx *ctx, double *grad_uin, size_t n, double *grad_gain, cfd_adjoint_stats *stats);
/* the number of the objective run cfd_adjoint_backward would differentiate
   (each cfd_adjoint_forward counts one), 0 when there is none: changing the
   state or the parameters after the run (cfd_step, cfd_init_state,
   cfd_upload, cfd_set_params, cfd_set_loop, cfd_checkpoint_load, the gain or
   the inflow profile) discards it, and cfd_adjoint_backward then returns
   CFD_ERR_STATE instead of mixing the old trajectory with the new values */
int cfd_adjoint_generation(cfd_ctx *ctx, unsigned long long *generation);

/* ---- output and restart files (host I/O) -------------------------------- */
/* Time_cDcL.dat writer modes */
#define CFD_CDCL_REFERENCE 0  /* Delay_IO_CD_CL as written: every flush writes delay_io copies of the current line */
#define CFD_CDCL_BUFFERED 1   /* Delay_IO_CD_CL as intended: every flush writes the delay_io buffered steps */
#define CFD_CDCL_EVERY_STEP 2 /* noDelay_IO_CD_CL: one line per step */
typedef struct cfd_cdcl_log cfd_cdcl_log;
/* the Tecplot drag / lift history of inc/dfib/virtualForceIntergrator.hpp:44-177
   (simuClass::delayIO and its cd / cl ring buffers), appended to path; zone
   is simuClass::ZONE(), e.g. "ELL (BICG)" */
int cfd_cdcl_open(const char *path, int delay_io, const char *zone, int mode, cfd_cdcl_log **out);
/* one step: simuClass::loop (1-based) and simuClass::time of the step (before
   finishloop), its cD and cL (cfd_drag): header at loop 1, flush when
   loop % delay_io == 0 (every call in CFD_CDCL_EVERY_STEP) */
int cfd_cdcl_record(cfd_cdcl_log *log, int loop, double time, double cd, double cl);
int cfd_cdcl_close(cfd_cdcl_log *log);
/* OutputPlot3D_Xfile (inc/io/Plot3D/xfileWrite.hpp:6-61): one block, cell
   centres Xc, Yc, Zc as float, i fastest */
int cfd_plot3d_write_x(const cfd_grid_desc *grid, const char *path);
/* OutputPLOT3D_Qfile (inc/io/Plot3D/qfile.hpp:7-104): header (1 block,
   nxCal, nyCal, nzCal, mach 0, alpha 0, Re, time as float) then p, the
   cell-centred velocity of the triple vel_field (CFD_T3_U in the reference's
   loop) and eta as float, i fastest; assembled on the device */
int cfd_plot3d_write_q(cfd_ctx *ctx, const char *path, int vel_field, double time);
/* QfileRead (inc/io/Plot3D/qfileRead.hpp:6-100): header[4] = mach, alpha,
   Re, time; p, u, v, w, eta ghosted (G entries each, zero outside the
   interior), face velocities averaged back from the cell-centred values.
   The file's grid must be the given one (CFD_ERR_ARG otherwise). */
int cfd_plot3d_read_q(const char *path, const cfd_grid_desc *grid, int *nblock, float header[4], double *p,
                      double *u, double *v, double *w, double *eta);

/* ---- checkpoint / resume ---------------------------------------------- */
/* An exact FP64 snapshot of what a timestep carries into the next: T0, T1,
   T3, p and eta (ghosted), the source term (3 N), X_result and the solver's
   NaN-reset history (N), and the loop, time, solve count (the restart rule's
   timestep_), pressure-location flag and bicgstabRe2's counters (pResets_,
   numRestarts_, totalNumIter_), so cfd_get_solver_counters continues too.  No reference counterpart: its only
   restart reads the float Plot3D file back, and that block is commented out
   (inc/run/RUN_CPU_PC_METHOD.hpp:98-111).  The operator, cfd_params, the
   forcing gain and the inflow profile are not in the file: a run resumes in a
   context created with the same grid and parameters and assembled the same
   way, where cfd_step continues bit for bit.  The file ends in a checksum.
   cfd_checkpoint_load reads and verifies the whole file before it changes
   the context; it rejects a file of another grid (CFD_ERR_ARG) and a damaged,
   truncated or foreign one (CFD_ERR_RUNTIME). */
int cfd_checkpoint_save(cfd_ctx *ctx, const char *path);
int cfd_checkpoint_load(cfd_ctx *ctx, const char *path);

/* ---- allocation-free solve ---------------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
   memory (code under stream capture, tight serving loops, harnesses that
   forbid hipMalloc): the BiCGSTAB solve of cfd_bicgstab -- a fresh
   bicgstabRe2's first solve (bicgstabRe2.hpp:91-189: no restart rule, NaN
   reset to the start value, px2 zero), the same kernels and reduction order,
   so the result is bit for bit cfd_bicgstab's on a new context -- on a
   caller-owned slot-major ELL operator (col, val: 7 * n_rows entries, the
   layout of cfd_ell_upload) with caller-owned right-hand side b and solution
   x (n_rows entries, x is the warm start and is overwritten).  All four are
   DEVICE pointers on the current HIP device; stream is a hipStream_t (NULL:
   the null stream) the solve is ordered on; the call returns once the solve
   has finished.  It allocates nothing: its work vectors, reduction partials
   and scalars live in a static device array of the library (about 65 MB per
   device, held until the process ends), which limits n_rows to
   CFD_NOALLOC_MAX_ROWS (CFD_ERR_ARG beyond).  Every call on one device
   shares that array, so calls that overlap on one device race. */
#define CFD_NOALLOC_MAX_ROWS 1048576
int cfd_bicgstab_ell_noalloc(const int32_t *col, const double *val, const double *b, double *x, size_t n_rows,
                             double tol, int maxit, void *stream, int *iters, double *relres);

/* ---- measurement ------------------------------------------------------ */
/* average device time (ms) of one launch of the kernel named by `what`, over
   nrep launches bracketed by HIP events on the context stream:
   what = 0 ELL SpMV FP64, 1 ELL SpMV FP32, 2 SPE SpMV FP64, 3 RB-SOR sweep,
   4 BiCGSTAB iteration (FP64), 5 predictor, 6 copy kernel (2*N doubles moved),
   7 CSR SpMV FP64, 8 the iteration's SpMV-dot kernel (ap = A p, partial ap . r0) */
int cfd_time_kernel(cfd_ctx *ctx, int what, int nrep, double *ms_avg);

#ifdef __cplusplus
}
#endif

#endif /* CFD_CAPI_H */
