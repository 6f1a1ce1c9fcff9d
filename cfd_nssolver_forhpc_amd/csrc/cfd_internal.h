// cfd_internal.h -- context layout and kernel launchers shared by the HIP
// translation units of libcfd_hip.so.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cfd_capi.h"

namespace cfd {

constexpr int GC = 2;           // ghost layers per side (grid::gC)
constexpr int KSLOTS = 7;       // ELL slots per row (createPressureMatrix resize(N, 7))
constexpr int RED_THREADS = 256;
// threads of the one-block kernels that combine the partials: 16 waves keep
// enough loads in flight to read 32768-65536 partials in a few microseconds
constexpr int FIN_THREADS = 1024;
// upper bound of partial sums per reduction: one row per thread up to 8.4M
// rows (256-thread blocks), which streams fastest on gfx950; larger vectors
// fall back to a grid-stride loop over this many blocks
constexpr int RED_MAX_BLOCKS = 32768;

// Grid extents and 1-D metrics on the device.
struct DevGrid {
  int nx, ny, nz;        // ghosted extents
  int nxc, nyc, nzc;     // interior extents
  int64_t G, N;          // ghosted / interior element counts
  const double *X, *Y, *Z;
  const double *Dx, *Dy, *Dz, *Dxs, *Dys, *Dzs;
  double lx, ly, lz;
};

// Position of slot s of row r in a device ELL array.  Slot-major (sliced
// 0) is the reference's layout, slot s of every row in a slab of N
// (ELL_sparseMatrix.hpp:146-160), as callers hand arrays over.  The library's
// own arrays are sliced: rows in slices of ELL_SLICE, each slice's KSLOTS
// slots contiguous, so a wave's slot loads stay inside one 3.5 KB (FP64
// values) region instead of spreading over seven slabs 67 MB apart (SpMV-dot
// at 512x256x64: 134.0 against 140.4 us).  Row r's slots are then
// base + s * stride.
constexpr int ELL_SLICE = 64;
__host__ __device__ inline int64_t ell_base(int64_t r, int sliced) {
  const uint64_t u = (uint64_t)r;
  return sliced ? (int64_t)((u / ELL_SLICE) * ((uint64_t)KSLOTS * ELL_SLICE) + u % ELL_SLICE) : r;
}
__host__ __device__ inline int64_t ell_stride(int64_t N, int sliced) { return sliced ? ELL_SLICE : N; }
// entries of a sliced array (the last slice padded to ELL_SLICE rows)
inline size_t ell_len(int64_t N) { return (size_t)KSLOTS * (size_t)((N + ELL_SLICE - 1) / ELL_SLICE) * ELL_SLICE; }

__host__ __device__ inline int64_t icel(const DevGrid &g, int i, int j, int k) {
  return ((int64_t)i * g.ny + j) * g.nz + k;
}
__host__ __device__ inline int64_t icelcal(const DevGrid &g, int i, int j, int k) {
  return ((int64_t)(i - GC) * g.nyc + (j - GC)) * g.nzc + (k - GC);
}

// Device-resident scalars of the Krylov solvers: stream-ordered, never
// round-tripped to the host inside an iteration.
struct KrylovScalars {
  double rho, pre_rho, alpha, omega, beta, norm0, rnorm, relres;
  double tol, restart_tol;
  double rmax, delta;  // mixed reliable updates: residual at the last replacement, trigger ratio
  int iters, maxit, done, hit_maxit, restart, timestep, restarts, breakdown;
  int replace, replacements, pad0, pad1;
};

// Reduction workspace: per-block partials plus a completion counter per
// reduction site; the last block to finish combines the partials in block
// order (deterministic) and resets its counter.
struct RedWork {
  double *partials;      // [RED_SITES][3][RED_MAX_BLOCKS]
  unsigned *counters;    // [RED_SITES]
  double *out;           // [RED_SITES][4] combined results
};
enum RedSite { RS_SPMV_DOT = 0, RS_SPMV_DOT2, RS_XR, RS_INIT, RS_NORM, RS_DRAG, RS_MAX, RS_MISC, RED_SITES };

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  cfd_params prm{};
  DevGrid g{};
  // host copies of the metrics
  std::vector<double> hX, hY, hZ, hDx, hDy, hDz, hDxs, hDys, hDzs;
  // fields
  double *f[CFD_NFIELDS] = {};
  // ELL (sliced, ell_base; slot-major where ell_sliced is 0: arrays a
  // caller hands over), FP32 copy, colour-permuted copy for RB-SOR
  int ell_sliced = 1;
  int32_t *col = nullptr;
  double *val = nullptr;
  float *val32 = nullptr;
  int32_t *colP = nullptr;
  double *valP = nullptr;
  float *valP32 = nullptr;
  int32_t *perm = nullptr;  // natural row -> permuted position
  int64_t ncolour[2] = {0, 0};
  bool ell_ready = false, spe_ready = false;
  // SPE coefficients (xm, xp over nxc; ym, yp over nyc; zm, zp over nzc) and jp
  double *spe_coef = nullptr;
  double *spe_jp = nullptr;
  // CSR operator (createPressureMatrixCSR): row pointers (N+1), columns,
  // values and its own Jacobi scaling
  int32_t *csr_ptr = nullptr, *csr_col = nullptr;
  double *csr_val = nullptr, *csr_jp = nullptr;
  int64_t csr_nnz = 0;
  bool csr_ready = false;
  // solver vectors (FP64)
  double *r0 = nullptr, *r1 = nullptr, *p1 = nullptr, *s1 = nullptr, *ap = nullptr, *as = nullptr;
  double *px1 = nullptr, *px2 = nullptr;
  // FP32 inner-solve vectors (mixed precision) and permuted SOR vectors
  float *q32[8] = {};   // FP32 Krylov vectors r0 r1 p1 s1 ap as x r
  double *xhi = nullptr; // FP64 residual of the mixed reliable-update solve
  double *xP = nullptr, *bP = nullptr, *rP = nullptr;
  float *xP32 = nullptr, *bP32 = nullptr;
  // device scalars and reduction workspace
  KrylovScalars *ks = nullptr;   // FP64 outer / plain solve
  KrylovScalars *ks32 = nullptr; // FP32 inner solve
  KrylovScalars *hks = nullptr;  // pinned host mirror
  RedWork red{};
  double *hred = nullptr;        // pinned host mirror of red.out
  double *stepbuf = nullptr;     // per-pass drag sums and max differences of one step
  int solves = 0;                // bicgstabRe2::timestep_ (solve count)
  int nan_resets = 0;            // bicgstabRe2::pResets_ of the FP64 solves
  long long total_iters = 0;     // bicgstabRe2::totalNumIter_ of the FP64 solves
  int solves32 = 0;
  int loop = 1;
  double time = 0.0;
  bool p_from_x = false;         // ghosted p interior lags X (step mode)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // per-phase device time of cfd_step (the reference's clockstruct
  // stopwatches): events recorded at the phase boundaries when enabled
  bool phase_timing = false;
  std::vector<hipEvent_t> phase_ev;
  double phase_ms[CFD_NPHASES] = {};
  int phase_steps = 0;
  bool debug = false;  // CFD_DEBUG=1: solver progress on stderr
  // parameters of the discrete adjoint and its forward record
  double gain = 1.0;             // DFIB forcing gain
  double *uin = nullptr;         // face-0 u Dirichlet profile (ny_cal), device; NULL: table value
  double *t2_rec = nullptr;      // cfd_step records T2 of pass q at t2_rec + 3*N*q when set
  struct AdjState *adj = nullptr;
  // multigrid preconditioner (hierarchy built on first use) and the
  // preconditioned directions of the right-preconditioned BiCGSTAB
  struct MGState *mg = nullptr;
  double *ph64 = nullptr, *sh64 = nullptr;
  float *ph32 = nullptr, *sh32 = nullptr;
  // the step's prediction (u, v, w; ghosted), kept by its first pass
  double *pred[3] = {nullptr, nullptr, nullptr};
  double *viseff = nullptr;  // Smagorinsky eddy viscosity (N), when prm.les
};

// k_mg.hip
int mg_ensure(Ctx &c, bool fp32);
void mg_release(Ctx &c);
int mg_apply64(Ctx &c, const double *r, double *out, const KrylovScalars *ks);
int mg_apply32(Ctx &c, const float *r, float *out, const KrylovScalars *ks);
int mg_apply64_fp32cycle(Ctx &c, const double *r, double *out, const KrylovScalars *ks);
const char *mg_shape(const Ctx &c);
// k_krylov.hip: FP64 BiCGSTAB right-preconditioned by the V-cycle (run in
// FP32 when pc32); the _aux form leaves the solve count untouched
int pbicgstab_solve64(Ctx &c, const double *b, double *x, double tol, int maxit, bool pc32, int *iters,
                      double *relres);
int pbicgstab_solve_aux(Ctx &c, const double *b, double *x, double tol, int maxit, bool pc32, int *iters,
                        double *relres);

// one timestep (cfd_step's body); adjoint.hip's recomputation calls it too
int step_impl(Ctx &c, cfd_step_stats *stats);
void adjoint_release(Ctx &c);
// the last objective run stops being differentiable (state or parameters changed)
void adjoint_invalidate(Ctx &c);
// FP64 BiCGSTAB on the ELL operator that leaves the pressure solver's state
// (solve count, NaN-reset history) untouched
int bicgstab_solve_aux(Ctx &c, const double *b, double *x, double tol, int maxit, int timestep, int *iters,
                       double *relres);

inline double *red_part_host(const Ctx &c, int site) {
  return c.red.partials + (size_t)site * 3 * RED_MAX_BLOCKS;
}

// error reporting
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);
#define CFD_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return ::cfd::fail(CFD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

inline int blocks_for(int64_t n, int threads, int cap) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

// ---- launchers (return hipError_t of the launch) -------------------------
// k_flow.hip
hipError_t launch_init_fields(Ctx &c, double noise_amp, uint64_t seed);
hipError_t launch_eta_cylinder_z(Ctx &c, double inv_n2);
hipError_t launch_predictor(Ctx &c, const double *u0, const double *v0, const double *w0, double *u1,
                            double *v1, double *w1, const double *src, int src_mode, double *const keep[3] = nullptr);
hipError_t launch_pred_add(Ctx &c, double *const keep[3], const double *src, double *u1, double *v1, double *w1);
hipError_t launch_source_add(Ctx &c, double *u1, double *v1, double *w1, const double *src, int quirk);
hipError_t launch_source_quirk_fixup(Ctx &c, double *u1, double *v1, double *w1, const double *src);
hipError_t launch_rhs(Ctx &c, const double *u1, const double *v1, const double *w1, const double *jp, double *b);
hipError_t launch_scatter_p(Ctx &c, const double *X, double *p);
// mode bits for the DFIB update
enum { UPD_WRITE_T3 = 1, UPD_WRITE_T0 = 2, UPD_SRC_SET = 4, UPD_SRC_ADD = 8, UPD_DRAG = 16, UPD_MAXDIFF = 32,
       UPD_P_FROM_X = 64 };
hipError_t launch_dfib_update_into(Ctx &c, int mode, const double *p_or_x, double *drag_out, double *max_out,
                                   double *t2out = nullptr);
hipError_t launch_source_accumulate(Ctx &c);
hipError_t launch_bc(Ctx &c, bool main, bool copy, bool pressure);
// periodic z: the ghost layers of (u, v, w) from their interior images (no-op otherwise)
hipError_t launch_periodic_refresh(Ctx &c, double *u, double *v, double *w);
// periodic z: the pressure's z ghost layers from their interior images (no-op otherwise)
hipError_t launch_periodic_refresh_p(Ctx &c, double *p);
hipError_t launch_drag(Ctx &c);
hipError_t launch_maxdiff(Ctx &c);
// sum over the interior of f[dir] times the cell volume into out[0]
hipError_t launch_force_sum(Ctx &c, int dir, double *out);
// divergence check of a velocity triple whose boundary faces come from bnd:
// out[0] max |div| over every interior cell, out[1] over the fluid cells,
// out[2] / out[3] the sums of div^2
hipError_t launch_divergence(Ctx &c, const double *const vel[3], const double *const bnd[3], double *out);
hipError_t launch_copy(Ctx &c, const double *src, double *dst, int64_t n);
hipError_t launch_fill(Ctx &c, double *dst, double v, int64_t n);
// k_ell.hip
hipError_t launch_ell_assemble(Ctx &c);
hipError_t launch_ell_permute(Ctx &c);
hipError_t launch_spe_assemble(Ctx &c);
hipError_t launch_csr_assemble(Ctx &c);
// ghosted p refreshed from X_result (with its BC ghosts) when the last
// timestep left the pressure there
int sync_pressure(Ctx &c);
// CFD_OK when operator `op` is assembled, else the CFD_ERR_STATE failure
int op_ready(const Ctx &c, int op);
// the FP64 operator of the configured path: prm.op when it is assembled,
// else the ELL operator
inline int fp64_op(const Ctx &c) {
  if (c.prm.op == CFD_OP_SPE && c.spe_ready) return CFD_OP_SPE;
  if (c.prm.op == CFD_OP_CSR && c.csr_ready) return CFD_OP_CSR;
  return CFD_OP_ELL;
}
hipError_t launch_spmv(Ctx &c, int op, int precision, const double *x, double *y);
hipError_t launch_spmv_ell_f32(Ctx &c, const float *x, float *y);
hipError_t launch_cast32(Ctx &c, const double *a, float *b, int64_t n);
int stream_blocks(int64_t n);
// k_krylov.hip
int bicgstab_solve(Ctx &c, const double *b, double *x, double tol, int maxit, int *iters, double *relres);
int defect_correction(Ctx &c, const double *b, double *x, double tol, double tol_inner, int max_outer,
                      int *outer, int *inner_total, double *relres);
int mixed_reliable_solve(Ctx &c, const double *b, double *x, double tol, double delta, int maxit, int *replacements,
                         int *iters, double *relres);
hipError_t launch_bicg_iteration(Ctx &c);
hipError_t launch_bicg_timing_setup(Ctx &c);
hipError_t launch_spmv_dot(Ctx &c);
// k_sor.hip
int rbsor_solve(Ctx &c, const double *b, double *x, double omega, double tol, int maxit, int check_every,
                int *iters, double *relres);
hipError_t launch_rbsor_sweep(Ctx &c, double omega);
// SorPipeLine: the reference's volume-weighted SOR on the ghosted pressure
// (in place, cf6 from T1 into CFD_RHS), red-black order, max |dp| stop;
// iters = sweeps - 1 and error = the last sweep's max |dp|, as the reference
int sor_pipeline_solve(Ctx &c, double omega, double crit, int itmax, int *iters, double *error);
// red-black SOR on the CSR operator (natural row order, colour by (i+j+k) parity)
int rbsor_csr_solve(Ctx &c, const double *b, double *x, double omega, double tol, int maxit, int check_every,
                    int *iters, double *relres);
int rbsor_solve_f32(Ctx &c, const float *b, float *x, double omega, double tol, int maxit, int check_every,
                    int *iters, double *relres);
hipError_t launch_copy_kernel_bench(Ctx &c, const double *src, double *dst, int64_t n);

}  // namespace cfd
