/*
 * cfd_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the reference solver's per-timestep hot path
 * (jimmy3769tw/CFD_NSsolver_forHPC, inc/...), used as the parity checker for
 * the HIP library and as the timed CPU baseline.  Only tests/, the smoke entry
 * point and the bench's CPU-baseline leg may load it; the product never does.
 *
 * Every function cites the reference file:line it restates.  Floating-point
 * contraction is off for this translation unit; where the reference build
 * (-O3 -march=native, GCC's default -ffp-contract=fast) contracts a
 * multiply-add whose rounding matters for a bit-exact target, the restatement
 * writes fma() explicitly and says so.
 *
 * Parity pins (see tests/): grid, ELL assembly and IB mask against the
 * compiled reference (oracle/_ref) and committed fixtures; predictor, RHS,
 * DFIB update, BC, SpMV, BiCGSTAB and the PC-method step against the compiled
 * reference at its native case and against closed forms.  Red-black SOR has
 * no reference counterpart (the reference SOR is lexicographic): it is pinned
 * by its fixed point (same solution as BiCGSTAB / a dense solve).
 */
#ifndef CFD_ORACLE_H
#define CFD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/cfd_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- grid (inc/grid/structureGrid.hpp) --------------------------------- */
/* init_uniform, one axis (structureGrid.hpp:409-440): X has n+1 entries,
   D and Ds n entries, n = ncal + 4 */
void orc_axis_uniform(int ncal, double L, double *X, double *D, double *Ds);
/* unit_SML + init_gC_dx, one axis (structureGrid.hpp:326-386, 506-559);
   returns 0, or -1 where the reference throws std::invalid_argument */
int orc_axis_sml(int ncal, double L, double lS, int nS, int off_nS, double *X, double *D, double *Ds);
/* tanh clustering around xc (no reference counterpart; the large config's
   stretched grid): widths grow by at most `ratio` per cell away from a band
   of width `band` of uniform spacing h0 centred on xc, then init_gC_dx */
int orc_axis_tanh(int ncal, double L, double xc, double h0, double band, double ratio,
                  double *X, double *D, double *Ds);
/* init_gC_dx for one axis (structureGrid.hpp:506-559) */
void orc_axis_gc_dx(int n, double *X, double *D, double *Ds);

/* ---- individual hot-path functions (ghosted arrays, layout of cfd_capi.h) */
void orc_eta_cylinder_z(const cfd_grid_desc *g, double cx, double cy, double R, int nsub, double *eta);
void orc_predictor(const cfd_grid_desc *g, double nu, double dt,
                   const double *u0, const double *v0, const double *w0,
                   double *u1, double *v1, double *w1);
/* the same with the scheme CFD_CONV_QUICK / _LUD / _UD (convection_and_difussion.hpp:20-165) */
void orc_predictor_ex(const cfd_grid_desc *g, int scheme, double nu, double dt,
                      const double *u0, const double *v0, const double *w0,
                      double *u1, double *v1, double *w1);
/* SmagorinskyModel (smagorinskyModel.hpp:6-79): eddy viscosity vis (N, icelCal) of (u, v, w) */
void orc_smagorinsky(const cfd_grid_desc *g, const double *u, const double *v, const double *w, double *vis);
/* the predictor with nu + vis[icelCal] in the diffusion (vis may be NULL) */
void orc_predictor_les(const cfd_grid_desc *g, int scheme, double nu, double dt, const double *vis,
                       const double *u0, const double *v0, const double *w0,
                       double *u1, double *v1, double *w1);
void orc_source_add(const cfd_grid_desc *g, int quirk, const double *src,
                    double *u1, double *v1, double *w1);
void orc_source_accumulate(const cfd_grid_desc *g, double dt, const double *f, double *src);
void orc_rhs(const cfd_grid_desc *g, double dt, const double *jp,
             const double *u1, const double *v1, const double *w1, double *b);
/* createPressureMatrix; col/val slot-major [7][N]; returns 0 or -1 */
int orc_ell_assemble(const cfd_grid_desc *g, const int p_dirichlet[6], int32_t *col, double *val, double *jp);
void orc_spmv_ell(size_t N, int K, const int32_t *col, const double *val, const double *x, double *y);
/* createPressureMatrixCSR (structureGrid.hpp:630-728): Neumann closure on every
   face whatever the pressure BC; ptr N+1, col/val up to 7N; returns nnz */
long long orc_csr_assemble(const cfd_grid_desc *g, int32_t *ptr, int32_t *col, double *val, double *jp);
/* CSR_matrix::multiply_omp (CSR_sparseMatrix.hpp:278-292) */
void orc_spmv_csr(size_t N, const int32_t *ptr, const int32_t *col, const double *val, const double *x, double *y);
/* SPE_matrix1 setupPressure + set_Jp_pc: coef holds xm[nx],xp[nx],ym[ny],yp[ny],zm[nz],zp[nz] (interior
   axis index) followed by jp[N] */
void orc_spe_setup(const cfd_grid_desc *g, const int p_dirichlet[6], double *coef, double *jp);
void orc_spmv_spe(const cfd_grid_desc *g, const double *coef, const double *jp, const double *x, double *y);
void orc_pressure_scatter(const cfd_grid_desc *g, const double *X, double *p);
void orc_dfib_update(const cfd_grid_desc *g, double dt, const double *eta, const double *p,
                     const double *u1, const double *v1, const double *w1,
                     double *u3, double *v3, double *w3, double *f);
void orc_bc_main(const cfd_grid_desc *g, const double dir[6][3], const double num[6][3],
                 const int p_dirichlet[6], double *u, double *v, double *w, double *p);
void orc_bc_copy(const cfd_grid_desc *g, const double *u0, const double *v0, const double *w0,
                 double *u1, double *v1, double *w1);
void orc_drag(const cfd_grid_desc *g, const double *f, double *cd, double *cl);
void orc_maxdiff(const cfd_grid_desc *g, const double *u0, const double *v0, const double *w0,
                 const double *u3, const double *v3, const double *w3, double out[3]);
/* deterministic dot product (fixed 256-way blocking, serial combine) */
double orc_dot(size_t n, const double *a, const double *b);
/* seeded uniform noise in [-amp, amp): splitmix64(seed ^ (component<<56) ^ icel) */
double orc_noise(uint64_t seed, int component, uint64_t icel, double amp);
/* v[m] = base + orc_noise(seed, tag, m, amp), the seeded inputs of the golden fixtures */
void orc_fill_noise(double *v, size_t n, uint64_t seed, int tag, double base, double amp);

/* ---- solvers ------------------------------------------------------------ */
typedef struct orc_bicg orc_bicg;
/* op: CFD_OP_ELL (col/val) or CFD_OP_SPE (g, coef, jp) */
orc_bicg *orc_bicg_create(size_t N, int op, const int32_t *col, const double *val,
                          const cfd_grid_desc *g, const double *spe_coef, const double *spe_jp);
/* op CFD_OP_CSR: ptr/col/val from orc_csr_assemble */
orc_bicg *orc_bicg_create_csr(size_t N, const int32_t *ptr, const int32_t *col, const double *val);
void orc_bicg_destroy(orc_bicg *s);
void orc_bicg_set_tolerance(orc_bicg *s, double tol);
void orc_bicg_set_maxit(orc_bicg *s, int maxit);
void orc_bicg_set_periodic_z(orc_bicg *s, int pz);
/* bicgstabRe2::solve (inc/solver/bicgstabRe2.hpp:91-189); returns relres, *iters */
double orc_bicg_solve(orc_bicg *s, const double *b, double *x, int *iters);
int orc_bicg_restarts(const orc_bicg *s);
int orc_bicg_resets(const orc_bicg *s);

/* red-black SOR on the ELL operator, natural ordering; colour = (i+j+k) & 1 */
double orc_rbsor(size_t N, const cfd_grid_desc *g, const int32_t *col, const double *val,
                 const double *b, double *x, double omega, double tol, int maxit,
                 int check_every, int *iters);
/* SorPipeLine_seq (inc/pressure/solver/SorPipeLine.seq.hpp:4-122): lexicographic
   SOR on the ghosted p with the volume-weighted coefficients, stops on
   max|dp| < crit; returns the final max|dp| and the reference's iteration count */
double orc_rbsor_csr(size_t N, const cfd_grid_desc *g, const int32_t *ptr, const int32_t *col, const double *val,
                     const double *b, double *x, double omega, double tol, int maxit, int check_every, int *iters);
/* SorPipeLine_seq's volume-weighted update in red-black order (the GPU's
   CFD_SOLVER_SOR_PIPELINE): in place on the ghosted p, max |dp| stop,
   iters = sweeps - 1; returns the last sweep's max |dp| */
double orc_sor_rb_vw(const cfd_grid_desc *g, double dt, double omega, double crit, int itmax,
                     const double *u1, const double *v1, const double *w1, double *p, int *iters);
double orc_sor_lex(const cfd_grid_desc *g, double dt, double omega, double crit, int itmax,
                   const double *u1, const double *v1, const double *w1, double *p, int *iters);

/* FP64 divergence check of a velocity triple, boundary faces from (ub, vb,
   wb): out[0] max |div|, out[1] sqrt(sum div^2) over the interior cells,
   out[2..3] over the fluid cells (no IB forcing on a face, no face on a
   Dirichlet pressure boundary pd) */
void orc_divergence(const cfd_grid_desc *g, const int pd[6], const double *u, const double *v, const double *w,
                    const double *ub, const double *vb, const double *wb, const double *eta, double out[4]);

/* ---- the timestep loop (RUN_CPU_PC_METHOD.hpp:287-627 / RUN_CPU.hpp:290-460) */
typedef struct orc_sim orc_sim;
orc_sim *orc_sim_create(const cfd_grid_desc *g, const cfd_params *prm);
void orc_sim_destroy(orc_sim *s);
/* pointer to a CFD_* array of the simulation; *len receives its length */
double *orc_sim_field(orc_sim *s, int field, size_t *len);
int32_t *orc_sim_ell_col(orc_sim *s);
/* iniU(u_init) + noise, p = 0, X = 0, T1 = T3 = T0, BC main + copy */
int orc_sim_init_state(orc_sim *s, double noise_amp, uint64_t seed);
/* DFIB_CylinderZ and the pressure operator (ELL and SPE, plus CSR when op is CSR) */
int orc_sim_setup(orc_sim *s);
int orc_sim_step(orc_sim *s, cfd_step_stats *st);
/* phases of one pass, exposed for the CPU baseline's timing */
int orc_sim_loop(const orc_sim *s);
void orc_sim_set_loop(orc_sim *s, int loop, double time);

/* ---- adjoint parameters (no reference counterpart) ---------------------- */
/* update_UandF with forcing gain (u_new = (1 - gain*eta_f) T2; gain 1 is the
   reference bit for bit); t2 (3*N, may be NULL) receives T2 */
void orc_dfib_update_ex(const cfd_grid_desc *g, double dt, double gain, const double *eta, const double *p,
                        const double *u1, const double *v1, const double *w1,
                        double *u3, double *v3, double *w3, double *f, double *t2);
/* BC_staggered_main with the face-0 u Dirichlet value replaced by uin[j-2] (uin may be NULL) */
void orc_bc_main_ex(const cfd_grid_desc *g, const double dir[6][3], const double num[6][3],
                    const int p_dirichlet[6], const double *uin, double *u, double *v, double *w, double *p);
/* inflow profile (ny_cal values, NULL: table value) and gain of the simulation */
int orc_sim_set_inflow(orc_sim *s, const double *uin);
void orc_sim_set_gain(orc_sim *s, double gain);

/* ---- discrete adjoint (transposes of the functions above) --------------- */
void orc_bc_main_adjoint(const cfd_grid_desc *g, const double num[6][3], double *g_uin,
                         double *lu, double *lv, double *lw);
void orc_bc_main_adjoint_bl(const cfd_grid_desc *g, const cfd_params *P, double *g_uin, double *lu, double *lv,
                            double *lw);
void orc_dfib_adjoint(const cfd_grid_desc *g, double dt, double gain, const double *eta, const double *t2,
                      const double *lu3, const double *lv3, const double *lw3, const double *lf,
                      double *lu1, double *lv1, double *lw1, double *lX, double *g_gain);
void orc_dfib_adjoint_bl(const cfd_grid_desc *g, const cfd_params *P, double gain, const double *eta,
                         const double *t2, const double *lu3, const double *lv3, const double *lw3,
                         const double *lf, double *lu1, double *lv1, double *lw1, double *lX, double *g_gain);
void orc_rhs_adjoint(const cfd_grid_desc *g, double dt, const double *jp, const double *lb,
                     double *lu1, double *lv1, double *lw1);
void orc_predictor_adjoint(const cfd_grid_desc *g, double nu, double dt,
                           const double *u, const double *v, const double *w,
                           const double *lu1, const double *lv1, const double *lw1,
                           double *lu, double *lv, double *lw);
int orc_ell_transpose(size_t N, const int32_t *col, const double *val, int32_t *colT, double *valT);
/* J = mean corrector-pass cD over steps [avg_from, nsteps) from the current
   state, recording the run for orc_sim_gradient; NaN on bad arguments */
double orc_sim_objective(orc_sim *s, int nsteps, int avg_from);
/* dJ/du_in (ny_cal values) and dJ/dgain of the last objective run; -1 when
   there is none or the configuration is not differentiable here (source
   index quirk, matrix-free operator) */
int orc_sim_gradient(orc_sim *s, double *g_uin, double *g_gain, int *adj_iters);


/* ---- boundary conditions the reference's tables cannot express (no
   reference counterpart): periodic z and a convective outflow on face 1,
   selected by cfd_params.periodic_z / outflow_convective / outflow_speed.
   The _bl functions take the boundary mode from P; the others keep the
   table semantics. */
void orc_predictor_bl(const cfd_grid_desc *g, const cfd_params *P, const double *vis,
                      const double *u, const double *v, const double *w,
                      double *u1, double *v1, double *w1);
int orc_ell_assemble_bl(const cfd_grid_desc *g, const int pd[6], int pz, int32_t *col, double *val, double *jp);
void orc_spe_setup_bl(const cfd_grid_desc *g, const int pd[6], int pz, double *coef, double *jp);
void orc_spmv_spe_bl(const cfd_grid_desc *g, int pz, const double *coef, const double *jp, const double *x,
                     double *y);
void orc_dfib_update_bl(const cfd_grid_desc *g, const cfd_params *P, double gain, const double *eta, const double *p,
                        const double *u1, const double *v1, const double *w1,
                        double *u3, double *v3, double *w3, double *f, double *t2);
void orc_bc_main_bl(const cfd_grid_desc *g, const cfd_params *P, const double *uin, double *u, double *v, double *w,
                    double *p);
void orc_bc_copy_bl(const cfd_grid_desc *g, const cfd_params *P, const double *u0, const double *v0,
                    const double *w0, double *u1, double *v1, double *w1);
void orc_periodic_refresh(const cfd_grid_desc *g, double *u, double *v, double *w);
void orc_periodic_refresh_p(const cfd_grid_desc *g, double *p);
double orc_sor_rb_vw_bl(const cfd_grid_desc *g, int pz, double dt, double omega, double crit, int itmax,
                        const double *u1, const double *v1, const double *w1, double *p, int *iters);
double orc_sor_lex_bl(const cfd_grid_desc *g, int pz, double dt, double omega, double crit, int itmax,
                      const double *u1, const double *v1, const double *w1, double *p, int *iters);
void orc_divergence_bl(const cfd_grid_desc *g, const cfd_params *P, const double *u, const double *v,
                       const double *w, const double *ub, const double *vb, const double *wb, const double *eta,
                       double out[4]);

#ifdef __cplusplus
}
#endif

#endif
