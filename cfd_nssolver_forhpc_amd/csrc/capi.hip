// capi.hip -- the extern "C" boundary of libcfd_hip.so (include/cfd_capi.h).
//
// Host side of the library: context and device-buffer management, argument
// validation (errors become status codes plus a thread-local message, never
// exceptions across the boundary), the host-side grid generators, and the
// sequencing of the fused timestep.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "cfd_internal.h"

namespace cfd {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

hipError_t launch_bicg_timing_setup(Ctx &c);
int rbsor_solve_f32(Ctx &c, const float *b, float *x, double omega, double tol, int maxit, int check_every,
                    int *iters, double *relres);
int stream_blocks(int64_t n);
hipError_t launch_cast32(Ctx &c, const double *a, float *b, int64_t n);

namespace {

template <typename T>
int dmalloc(T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(CFD_ERR_NOMEM, std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) +
                                   " bytes failed: " + hipGetErrorString(e));
  }
  return CFD_OK;
}

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

int check_params(const cfd_params *p) {
  if (!p) return fail(CFD_ERR_ARG, "params is NULL");
  if (!(p->dt > 0.0) || !std::isfinite(p->dt)) return fail(CFD_ERR_ARG, "dt must be positive and finite");
  if (!(p->Re > 0.0) || !std::isfinite(p->Re)) return fail(CFD_ERR_ARG, "Re must be positive and finite");
  if (p->inner_loops < 0 || p->inner_loops > CFD_MAX_PASSES - 1)
    return fail(CFD_ERR_ARG, "inner_loops must be in [0, 7]");
  if (p->solver != CFD_SOLVER_BICGSTAB && p->solver != CFD_SOLVER_RBSOR && p->solver != CFD_SOLVER_SOR_PIPELINE)
    return fail(CFD_ERR_ARG, "unknown solver");
  if (p->solver == CFD_SOLVER_SOR_PIPELINE && p->precision != CFD_FP64)
    return fail(CFD_ERR_ARG, "SorPipeLine runs in FP64");
  if (p->solver == CFD_SOLVER_SOR_PIPELINE && !(p->sor_omega > 0.0 && p->sor_omega < 2.0))
    return fail(CFD_ERR_ARG, "p_sor_omega must be in (0, 2)");
  if (p->op != CFD_OP_ELL && p->op != CFD_OP_SPE && p->op != CFD_OP_CSR) return fail(CFD_ERR_ARG, "unknown operator");
  if (p->precision != CFD_FP64 && p->precision != CFD_MIXED) return fail(CFD_ERR_ARG, "unknown precision");
  if (p->solver == CFD_SOLVER_RBSOR && p->op != CFD_OP_ELL && p->op != CFD_OP_CSR)
    return fail(CFD_ERR_ARG, "red-black SOR runs on the ELL or the CSR operator");
  if (p->solver == CFD_SOLVER_RBSOR && p->op == CFD_OP_CSR && p->precision != CFD_FP64)
    return fail(CFD_ERR_ARG, "red-black SOR on the CSR operator runs in FP64");
  if (p->maxit < 0) return fail(CFD_ERR_ARG, "maxit must be >= 0");
  if (p->precision == CFD_MIXED && p->mixed_mode != CFD_MIXED_RELIABLE && p->mixed_mode != CFD_MIXED_RESTARTED &&
      p->mixed_mode != CFD_MIXED_VALUES32 && p->mixed_mode != CFD_MIXED_PC32)
    return fail(CFD_ERR_ARG, "unknown mixed-precision scheme");
  if (p->precision == CFD_MIXED && p->mixed_mode == CFD_MIXED_PC32 && p->precond != CFD_PRECOND_MG)
    return fail(CFD_ERR_ARG, "CFD_MIXED_PC32 needs precond = CFD_PRECOND_MG");
  if (p->precision == CFD_MIXED && (p->mixed_mode == CFD_MIXED_RELIABLE || p->mixed_mode == CFD_MIXED_VALUES32) &&
      !(p->mixed_delta > 0.0 && p->mixed_delta < 1.0))
    return fail(CFD_ERR_ARG, "mixed_delta must be in (0, 1)");
  if (p->ib_on && (p->ib_nsub < 1 || !(p->ib_r > 0.0))) return fail(CFD_ERR_ARG, "invalid immersed body");
  if (p->precond != CFD_PRECOND_JACOBI && p->precond != CFD_PRECOND_MG) return fail(CFD_ERR_ARG, "unknown preconditioner");
  if (p->convection != CFD_CONV_QUICK && p->convection != CFD_CONV_LUD && p->convection != CFD_CONV_UD)
    return fail(CFD_ERR_ARG, "unknown convection scheme");
  if (p->les != 0 && p->les != 1) return fail(CFD_ERR_ARG, "les must be 0 or 1");
  if (p->precond == CFD_PRECOND_MG && (p->mg_nu < 1 || p->mg_nu > 8)) return fail(CFD_ERR_ARG, "mg_nu must be in [1, 8]");
  if (p->periodic_z != 0 && p->periodic_z != 1) return fail(CFD_ERR_ARG, "periodic_z must be 0 or 1");
  if (p->outflow_convective != 0 && p->outflow_convective != 1)
    return fail(CFD_ERR_ARG, "outflow_convective must be 0 or 1");
  if (p->outflow_convective && !(p->outflow_speed > 0.0 && std::isfinite(p->outflow_speed)))
    return fail(CFD_ERR_ARG, "outflow_speed must be positive and finite");
  if (p->outflow_convective && p->p_dirichlet[1] != 1)
    return fail(CFD_ERR_ARG, "a convective outflow needs the Dirichlet pressure p = 0 on face 1");
  if (p->periodic_z && p->op == CFD_OP_CSR)
    return fail(CFD_ERR_ARG, "the CSR builder closes every face (createPressureMatrixCSR): no periodic z");
  if ((p->periodic_z || p->outflow_convective) && p->source_index_quirk)
    return fail(CFD_ERR_ARG, "periodic / convective boundaries add the source term at its own cell "
                             "(source_index_quirk 0)");
  return CFD_OK;
}

// the checks of cfd_params that need the grid
int check_params_grid(const cfd_params *p, int nz_cal) {
  if (p->periodic_z && (nz_cal < 4 || nz_cal % 2 != 0))
    return fail(CFD_ERR_ARG, "periodic_z needs an even nz_cal of at least 4 (red-black colouring across the seam)");
  return CFD_OK;
}

int field_len(const Ctx &c, int field, size_t *len) {
  if (field < 0 || field >= CFD_NFIELDS) return fail(CFD_ERR_ARG, "unknown field id " + std::to_string(field));
  if (field <= CFD_ETA) *len = (size_t)c.g.G;
  else if (field == CFD_F || field == CFD_SRC) *len = (size_t)(3 * c.g.N);
  else *len = (size_t)c.g.N;
  return CFD_OK;
}

void free_ctx(Ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  adjoint_release(*c);
  mg_release(*c);
  dfree(c->uin);
  dfree(c->viseff);
  for (auto &q : c->pred) dfree(q);
  for (auto &p : c->f) dfree(p);
  dfree(c->col); dfree(c->val); dfree(c->val32); dfree(c->colP); dfree(c->valP); dfree(c->valP32); dfree(c->perm);
  dfree(c->spe_coef); dfree(c->spe_jp);
  dfree(c->csr_ptr); dfree(c->csr_col); dfree(c->csr_val); dfree(c->csr_jp);
  dfree(c->r0); dfree(c->r1); dfree(c->p1); dfree(c->s1); dfree(c->ap); dfree(c->as); dfree(c->px1); dfree(c->px2);
  for (auto &p : c->q32) dfree(p);
  dfree(c->xP); dfree(c->bP); dfree(c->rP); dfree(c->xP32); dfree(c->bP32);
  dfree(c->ks); dfree(c->ks32);
  dfree(c->red.partials); dfree(c->red.counters); dfree(c->red.out);
  dfree(c->stepbuf);
  double *metr[9] = {(double *)c->g.X, (double *)c->g.Y, (double *)c->g.Z, (double *)c->g.Dx, (double *)c->g.Dy,
                     (double *)c->g.Dz, (double *)c->g.Dxs, (double *)c->g.Dys, (double *)c->g.Dzs};
  for (auto p : metr) if (p) (void)hipFree(p);
  if (c->hks) (void)hipHostFree(c->hks);
  if (c->hred) (void)hipHostFree(c->hred);
  for (hipEvent_t e : c->phase_ev) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// FP32 copy of the ELL values (mixed precision)
int ensure_mixed(Ctx &c) {
  if (!c.ell_ready) return fail(CFD_ERR_STATE, "ELL operator not assembled (cfd_ell_assemble)");
  const int64_t N = c.g.N;
  if (!c.val32) {
    int rc = dmalloc(&c.val32, ell_len(N));
    if (rc) return rc;
    CFD_HIP(launch_cast32(c, c.val, c.val32, (int64_t)ell_len(N)));
  }
  for (int q = 0; q < 8; ++q)
    if (!c.q32[q]) {
      int rc = dmalloc(&c.q32[q], (size_t)N);
      if (rc) return rc;
    }
  if (c.valP && !c.valP32) {
    int rc = dmalloc(&c.valP32, ell_len(N));
    if (rc) return rc;
    CFD_HIP(launch_cast32(c, c.valP, c.valP32, (int64_t)ell_len(N)));
  }
  if (c.valP && !c.xP32) {
    int rc = dmalloc(&c.xP32, (size_t)N);
    if (!rc) rc = dmalloc(&c.bP32, (size_t)N);
    if (rc) return rc;
  }
  return CFD_OK;
}

// colour-permuted ELL copy (red-black SOR)
int ensure_perm(Ctx &c) {
  if (!c.ell_ready) return fail(CFD_ERR_STATE, "ELL operator not assembled (cfd_ell_assemble)");
  if (c.colP) return CFD_OK;
  const int64_t N = c.g.N;
  std::vector<int32_t> perm((size_t)N);
  int64_t n0 = 0;
  for (int i = 0; i < c.g.nxc; ++i)
    for (int j = 0; j < c.g.nyc; ++j)
      for (int k = 0; k < c.g.nzc; ++k)
        if (((i + j + k) & 1) == 0) ++n0;
  int64_t q0 = 0, q1 = n0;
  int64_t r = 0;
  for (int i = 0; i < c.g.nxc; ++i)
    for (int j = 0; j < c.g.nyc; ++j)
      for (int k = 0; k < c.g.nzc; ++k, ++r) perm[(size_t)r] = (int32_t)(((i + j + k) & 1) == 0 ? q0++ : q1++);
  c.ncolour[0] = n0;
  c.ncolour[1] = N - n0;
  int rc = dmalloc(&c.perm, (size_t)N);
  if (!rc) rc = dmalloc(&c.colP, ell_len(N));
  if (!rc) rc = dmalloc(&c.valP, ell_len(N));
  if (!rc) rc = dmalloc(&c.xP, (size_t)N);
  if (!rc) rc = dmalloc(&c.bP, (size_t)N);
  if (!rc && c.val32) rc = dmalloc(&c.valP32, ell_len(N));
  if (rc) return rc;
  CFD_HIP(hipMemcpyAsync(c.perm, perm.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, c.stream));
  CFD_HIP(launch_ell_permute(c));
  CFD_HIP(hipStreamSynchronize(c.stream));
  if (c.val32) {
    int rc2 = dmalloc(&c.xP32, (size_t)N);
    if (!rc2) rc2 = dmalloc(&c.bP32, (size_t)N);
    if (rc2) return rc2;
  }
  return CFD_OK;
}

// ghosted p from X (interior) plus the pressure ghosts of BC_staggered_main
int sync_p(Ctx &c) {
  if (!c.p_from_x) return CFD_OK;
  CFD_HIP(launch_scatter_p(c, c.f[CFD_X], c.f[CFD_P]));
  CFD_HIP(launch_bc(c, false, false, true));
  c.p_from_x = false;
  return CFD_OK;
}

int solve_configured(Ctx &c, int *iters, double *relres, int *inner) {
  const cfd_params &P = c.prm;
  *inner = 0;
  if (P.solver == CFD_SOLVER_SOR_PIPELINE) {  // T1 and the ghosted p, in place
    if (int rc = sync_p(c)) return rc;
    return sor_pipeline_solve(c, P.sor_omega, P.tol, P.maxit, iters, relres);
  }
  if (P.precond == CFD_PRECOND_MG && P.solver == CFD_SOLVER_BICGSTAB &&
      (P.precision == CFD_FP64 || P.mixed_mode == CFD_MIXED_PC32)) {
    // FP64 BiCGSTAB preconditioned by the V-cycle, the cycle in FP32 for CFD_MIXED_PC32
    return pbicgstab_solve64(c, c.f[CFD_RHS], c.f[CFD_X], P.tol, P.maxit, P.precision == CFD_MIXED, iters, relres);
  }
  if (P.precision == CFD_MIXED) {
    int rc = ensure_mixed(c);
    if (!rc && P.solver == CFD_SOLVER_RBSOR) {
      rc = ensure_perm(c);
      if (!rc) rc = ensure_mixed(c);
    }
    if (rc) return rc;
    int outer = 0;
    if (P.mixed_mode != CFD_MIXED_RESTARTED && P.solver == CFD_SOLVER_BICGSTAB && P.precond != CFD_PRECOND_MG)
      rc = mixed_reliable_solve(c, c.f[CFD_RHS], c.f[CFD_X], P.tol, P.mixed_delta, P.maxit, &outer, inner, relres);
    else
      rc = defect_correction(c, c.f[CFD_RHS], c.f[CFD_X], P.tol, P.tol_inner, P.maxit, &outer, inner, relres);
    *iters = outer;
    return rc;
  }
  if (P.solver == CFD_SOLVER_RBSOR && P.op == CFD_OP_CSR)
    return rbsor_csr_solve(c, c.f[CFD_RHS], c.f[CFD_X], P.sor_omega, P.tol, P.maxit, P.sor_check, iters, relres);
  if (P.solver == CFD_SOLVER_RBSOR) {
    int rc = ensure_perm(c);
    if (rc) return rc;
    return rbsor_solve(c, c.f[CFD_RHS], c.f[CFD_X], P.sor_omega, P.tol, P.maxit, P.sor_check, iters, relres);
  }
  return bicgstab_solve(c, c.f[CFD_RHS], c.f[CFD_X], P.tol, P.maxit, iters, relres);
}

}  // namespace

int sync_pressure(Ctx &c) { return sync_p(c); }

}  // namespace cfd

using namespace cfd;

#define CTX_CHECK(ctx)                                              \
  do {                                                              \
    if (!(ctx)) return fail(CFD_ERR_ARG, "context is NULL");        \
    hipError_t se_ = hipSetDevice(((Ctx *)(ctx))->device);          \
    if (se_ != hipSuccess) return fail(CFD_ERR_HIP, "hipSetDevice failed"); \
  } while (0)

#define RET_LAUNCH(expr)                                                                              \
  do {                                                                                                \
    hipError_t le_ = (expr);                                                                          \
    if (le_ != hipSuccess) return fail(CFD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(le_)); \
  } while (0)

extern "C" {

int cfd_abi_version(void) { return CFD_ABI_VERSION; }

const char *cfd_last_error(void) { return g_last_error.c_str(); }

int cfd_params_default(cfd_params *out) {
  if (!out) return fail(CFD_ERR_ARG, "out is NULL");
  std::memset(out, 0, sizeof(*out));
  out->Re = 40.0;   // src/CFD_MX.cpp:39
  out->dt = 0.002;  // src/CFD_MX.cpp:37
  out->bc_dir[3][0] = 1.0;  // BC_Cavity_A.hpp:13 (lid on y+)
  out->source_index_quirk = 1;
  out->inner_loops = 3;
  out->solver = CFD_SOLVER_BICGSTAB;
  out->op = CFD_OP_ELL;
  out->precision = CFD_FP64;
  out->tol = 1e-3;        // simuClass::p_criteria (simuClass.hpp:74)
  out->tol_inner = 1e-5;
  out->mixed_mode = CFD_MIXED_RELIABLE;
  out->mixed_delta = 0.1;
  out->maxit = 5000;      // bicgstabRe2::iters_max_ (bicgstabRe2.hpp:57)
  out->sor_omega = 1.8;   // simuClass::p_sor_omega (simuClass.hpp:78)
  out->sor_check = 10;
  out->ib_on = 1;
  out->ib_cx = 6.5;       // setEta.hpp:145-147
  out->ib_cy = 6.0;
  out->ib_r = 0.5;
  out->ib_nsub = 100;
  out->u_init[0] = 1.0;   // RUN_CPU_PC_METHOD.hpp:77
  out->precond = CFD_PRECOND_JACOBI;
  out->mg_nu = 4;  // fastest at 512x256x64: 2 -> 4 halves the iterations for ~1.4x the cycle
  out->convection = CFD_CONV_QUICK;  // controlPanel.hpp:125
  out->les = 0;                      // controlPanel.hpp:135 (TERBULENCE_SMAGORINSKY off)
  out->periodic_z = 0;
  out->outflow_convective = 0;
  out->outflow_speed = 1.0;
  return CFD_OK;
}

// ---- grid generation (host) ---------------------------------------------
static void gc_dx(int n, double *X, double *D, double *Ds) {
  X[1] = 2 * X[2] - X[3];
  X[0] = 2 * X[1] - X[2];
  X[n - 1] = 2 * X[n - 2] - X[n - 3];
  X[n] = 2 * X[n - 1] - X[n - 2];
  for (int i = GC; i < n - GC - 1; ++i) Ds[i] = (X[i + 2] - X[i]) / 2.0;
  for (int i = GC; i < n - GC - 1; ++i) D[i] = X[i + 1] - X[i];
  D[0] = D[1] = D[2];
  Ds[0] = Ds[1] = Ds[2];
  D[n - 1] = D[n - 2] = D[n - 3] = X[n - 2] - X[n - 3];
  Ds[n - 1] = Ds[n - 2] = Ds[n - 3] = Ds[n - 4];
}

// grid::init_uniform for one axis (structureGrid.hpp:409-440)
int cfd_axis_uniform(int ncal, double L, double *X, double *D, double *Ds) {
  if (ncal < 2 || !(L > 0.0) || !X || !D || !Ds) return fail(CFD_ERR_ARG, "cfd_axis_uniform: bad arguments");
  const int n = ncal + 2 * GC;
  const double d = L / (n - 2.0 * GC);
  X[0] = 0.0 - 2.0 * d;
  for (int i = 1; i < n + 1; ++i) X[i] = X[i - 1] + d;
  for (int i = 0; i < n; ++i) {
    D[i] = d;
    Ds[i] = d;
  }
  return CFD_OK;
}

// grid::unit_SML + init_gC_dx for one axis (structureGrid.hpp:326-386, 506-559)
int cfd_axis_sml(int ncal, double L, double lS, int nS, int off_nS, double *X, double *D, double *Ds) {
  if (ncal < 2 || !(L > 0.0) || !X || !D || !Ds) return fail(CFD_ERR_ARG, "cfd_axis_sml: bad arguments");
  const int n = ncal + 2 * GC;
  std::fill(X, X + n + 1, 0.0);
  X[GC] = 0.0;
  if (nS != 0) {
    if ((n - nS) % 2 != 0) return fail(CFD_ERR_ARG, "(n-nSml)%2 != 0");
    const double dS = lS / nS, d = (L - lS) / (ncal - nS);
    int nlEnd0 = (ncal - nS) / 2, nsEnd0 = nlEnd0 + nS;
    const int nlEnd1 = nsEnd0 + nlEnd0;
    nlEnd0 += off_nS;
    nsEnd0 += off_nS;
    if (nlEnd0 <= 0) return fail(CFD_ERR_ARG, "nL[0]");
    if (nsEnd0 > nlEnd1 || nlEnd1 > ncal) return fail(CFD_ERR_ARG, "SML sections exceed the axis");
    for (int i = 0; i < nlEnd0; ++i) X[i + GC + 1] = X[i + GC] + d;
    for (int i = nlEnd0; i < nsEnd0; ++i) X[i + GC + 1] = X[i + GC] + dS;
    for (int i = nsEnd0; i < nlEnd1; ++i) X[i + GC + 1] = X[i + GC] + d;
  } else {
    for (int i = GC; i < n - GC; ++i) X[i + 1] = X[i] + L / ncal;
  }
  gc_dx(n, X, D, Ds);
  return CFD_OK;
}

// Stretched axis of the large configuration (no reference counterpart): tanh
// clustering around xc.  The cell width, as a function of the node index
// offset xi from the centre, blends from h0 inside a band of +-a cells to H
// outside through tanh steps of width w cells,
//   dx/dxi = H - (H - h0)/2 [tanh((xi + a)/w) - tanh((xi - a)/w)],
// so the node positions have the closed form
//   x(xi) = xc + H xi - (H - h0)/2 w [lncosh((xi + a)/w) - lncosh((xi - a)/w)]
// (lncosh(t) = |t| + log1p(exp(-2|t|)) - log 2).  Node i sits at xi0 + i:
// xi0 puts node 0 at 0, H puts node ncal at L (both by bisection; x is
// increasing in xi, and x(xi0 + ncal) in H), and w is the narrowest step
// whose largest ratio of neighbouring widths is at most `ratio`.
namespace {
double lncosh(double t) {
  const double a = std::fabs(t);
  return a + std::log1p(std::exp(-2.0 * a)) - std::log(2.0);
}
struct TanhAxis {
  double h0, a, w, H, xc;
  double f(double xi) const {
    return H * xi - 0.5 * (H - h0) * w * (lncosh((xi + a) / w) - lncosh((xi - a) / w));
  }
};
// xi0 with xc + f(xi0) = 0
double tanh_xi0(const TanhAxis &t, int ncal) {
  double lo = -(double)ncal - 1.0, hi = 0.0;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (t.xc + t.f(mid) > 0.0) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}
// H (and xi0) with node ncal at L; returns the largest neighbouring-width ratio
double tanh_fit(TanhAxis &t, int ncal, double L, double *xi0_out) {
  double lo = t.h0, hi = L;
  double xi0 = 0.0;
  for (int it = 0; it < 200; ++it) {
    t.H = 0.5 * (lo + hi);
    xi0 = tanh_xi0(t, ncal);
    if (t.xc + t.f(xi0 + ncal) > L) hi = t.H;
    else lo = t.H;
  }
  t.H = 0.5 * (lo + hi);
  xi0 = tanh_xi0(t, ncal);
  double rmax = 1.0, prev = t.f(xi0 + 1.0) - t.f(xi0);
  for (int c = 1; c < ncal; ++c) {
    const double h = t.f(xi0 + c + 1.0) - t.f(xi0 + c);
    rmax = std::max(rmax, std::max(h / prev, prev / h));
    prev = h;
  }
  *xi0_out = xi0;
  return rmax;
}
}  // namespace

int cfd_axis_tanh(int ncal, double L, double xc, double h0, double band, double ratio, double *X, double *D,
                  double *Ds) {
  if (ncal < 4 || !(L > 0) || !(h0 > 0) || !(ratio > 1.0) || h0 * ncal >= L || !(xc > 0 && xc < L) ||
      !(band > 0) || !X || !D || !Ds)
    return fail(CFD_ERR_ARG, "cfd_axis_tanh: bad arguments");
  const int n = ncal + 2 * GC;
  TanhAxis t{h0, 0.5 * band / h0, 1.0, h0, xc};
  double xi0 = 0.0;
  // the narrowest tanh step that keeps the width ratio within `ratio`
  double wlo = 0.25, whi = (double)ncal;
  t.w = whi;
  if (tanh_fit(t, ncal, L, &xi0) > ratio) return fail(CFD_ERR_ARG, "cfd_axis_tanh: ratio not reachable");
  for (int it = 0; it < 60; ++it) {
    t.w = 0.5 * (wlo + whi);
    if (tanh_fit(t, ncal, L, &xi0) > ratio) wlo = t.w;
    else whi = t.w;
  }
  t.w = whi;
  tanh_fit(t, ncal, L, &xi0);
  std::fill(X, X + n + 1, 0.0);
  for (int c = 1; c < ncal; ++c) X[GC + c] = xc + t.f(xi0 + c);
  X[GC] = 0.0;
  X[GC + ncal] = L;
  gc_dx(n, X, D, Ds);
  return CFD_OK;
}

int cfd_device_count(int *n) {
  if (!n) return fail(CFD_ERR_ARG, "n is NULL");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *n = (e == hipSuccess) ? c : 0;
  return CFD_OK;
}

// ---- context ----------------------------------------------------------------
int cfd_create(const cfd_grid_desc *grid, const cfd_params *params, int device, cfd_ctx **out) {
  if (!out) return fail(CFD_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (!grid) return fail(CFD_ERR_ARG, "grid is NULL");
  int rc = check_params(params);
  if (rc) return rc;
  if (grid->nx_cal < 2 || grid->ny_cal < 2 || grid->nz_cal < 2)
    return fail(CFD_ERR_ARG, "every axis needs at least 2 interior cells");
  if ((rc = check_params_grid(params, grid->nz_cal))) return rc;
  const int64_t N = (int64_t)grid->nx_cal * grid->ny_cal * grid->nz_cal;
  if (N >= (int64_t)INT32_MAX) return fail(CFD_ERR_ARG, "grid too large for int32 ELL columns");
  const double *m[9] = {grid->X, grid->Y, grid->Z, grid->Dx, grid->Dy, grid->Dz, grid->Dxs, grid->Dys, grid->Dzs};
  for (auto p : m)
    if (!p) return fail(CFD_ERR_ARG, "grid metric array is NULL");
  if (!(grid->lz > 0.0)) return fail(CFD_ERR_ARG, "lz must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(CFD_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(CFD_ERR_ARG, "device index out of range");
  CFD_HIP(hipSetDevice(device));

  Ctx *c = new (std::nothrow) Ctx();
  if (!c) return fail(CFD_ERR_NOMEM, "host allocation failed");
  c->device = device;
  { const char *dbg = std::getenv("CFD_DEBUG"); c->debug = dbg && dbg[0] == '1'; }
  c->prm = *params;
  DevGrid &g = c->g;
  g.nxc = grid->nx_cal; g.nyc = grid->ny_cal; g.nzc = grid->nz_cal;
  g.nx = g.nxc + 2 * GC; g.ny = g.nyc + 2 * GC; g.nz = g.nzc + 2 * GC;
  g.G = (int64_t)g.nx * g.ny * g.nz;
  g.N = N;
  g.lx = grid->lx; g.ly = grid->ly; g.lz = grid->lz;
  const int lens[9] = {g.nx + 1, g.ny + 1, g.nz + 1, g.nx, g.ny, g.nz, g.nx, g.ny, g.nz};
  std::vector<double> *hv[9] = {&c->hX, &c->hY, &c->hZ, &c->hDx, &c->hDy, &c->hDz, &c->hDxs, &c->hDys, &c->hDzs};
  const double **dv[9] = {&g.X, &g.Y, &g.Z, &g.Dx, &g.Dy, &g.Dz, &g.Dxs, &g.Dys, &g.Dzs};
  auto bail = [&](int code) {
    free_ctx(c);
    return code;
  };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(CFD_ERR_HIP, "hipStreamCreate failed"));
  for (int q = 0; q < 9; ++q) {
    hv[q]->assign(m[q], m[q] + lens[q]);
    double *d = nullptr;
    if ((rc = dmalloc(&d, (size_t)lens[q]))) return bail(rc);
    *dv[q] = d;
    if (hipMemcpy(d, m[q], sizeof(double) * lens[q], hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(CFD_ERR_HIP, "metric upload failed"));
  }
  for (int fid = 0; fid < CFD_NFIELDS; ++fid) {
    size_t len = 0;
    field_len(*c, fid, &len);
    if ((rc = dmalloc(&c->f[fid], len))) return bail(rc);
    if (hipMemset(c->f[fid], 0, len * sizeof(double)) != hipSuccess) return bail(fail(CFD_ERR_HIP, "memset"));
  }
  if ((rc = dmalloc(&c->col, ell_len(N))) || (rc = dmalloc(&c->val, ell_len(N)))) return bail(rc);
  double **kv[8] = {&c->r0, &c->r1, &c->p1, &c->s1, &c->ap, &c->as, &c->px1, &c->px2};
  for (auto p : kv) {
    if ((rc = dmalloc(p, (size_t)N))) return bail(rc);
    if (hipMemset(*p, 0, sizeof(double) * N) != hipSuccess) return bail(fail(CFD_ERR_HIP, "memset"));
  }
  if ((rc = dmalloc(&c->spe_coef, (size_t)2 * (g.nxc + g.nyc + g.nzc))) || (rc = dmalloc(&c->spe_jp, (size_t)N)))
    return bail(rc);
  if ((rc = dmalloc(&c->ks, 1)) || (rc = dmalloc(&c->ks32, 1))) return bail(rc);
  (void)hipMemset(c->ks, 0, sizeof(KrylovScalars));
  (void)hipMemset(c->ks32, 0, sizeof(KrylovScalars));
  if ((rc = dmalloc(&c->red.partials, (size_t)RED_SITES * 3 * RED_MAX_BLOCKS)) ||
      (rc = dmalloc(&c->red.counters, (size_t)RED_SITES)) || (rc = dmalloc(&c->red.out, (size_t)RED_SITES * 4)) ||
      (rc = dmalloc(&c->stepbuf, (size_t)64)))
    return bail(rc);
  (void)hipMemset(c->red.counters, 0, sizeof(unsigned) * RED_SITES);
  if (hipHostMalloc((void **)&c->hks, sizeof(KrylovScalars), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&c->hred, 64 * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return bail(fail(CFD_ERR_NOMEM, "pinned host allocation failed"));
  if (hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
    return bail(fail(CFD_ERR_HIP, "hipEventCreate failed"));
  *out = reinterpret_cast<cfd_ctx *>(c);
  return CFD_OK;
}

int cfd_destroy(cfd_ctx *ctx) {
  if (!ctx) return CFD_OK;
  Ctx *c = reinterpret_cast<Ctx *>(ctx);
  (void)hipSetDevice(c->device);
  free_ctx(c);
  return CFD_OK;
}

int cfd_set_params(cfd_ctx *ctx, const cfd_params *params) {
  CTX_CHECK(ctx);
  int rc = check_params(params);
  if (rc) return rc;
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if ((rc = check_params_grid(params, c.g.nzc))) return rc;
  adjoint_invalidate(c);
  // the multigrid hierarchy follows the pressure boundary conditions
  bool mg_stale = params->periodic_z != c.prm.periodic_z;
  for (int f = 0; f < 6; ++f) mg_stale = mg_stale || params->p_dirichlet[f] != c.prm.p_dirichlet[f];
  if (mg_stale && c.mg) {
    CFD_HIP(hipStreamSynchronize(c.stream));
    mg_release(c);
  }
  c.prm = *params;
  return CFD_OK;
}

int cfd_sizes(const cfd_ctx *ctx, size_t *G, size_t *N) {
  if (!ctx) return fail(CFD_ERR_ARG, "context is NULL");
  const Ctx &c = *reinterpret_cast<const Ctx *>(ctx);
  if (G) *G = (size_t)c.g.G;
  if (N) *N = (size_t)c.g.N;
  return CFD_OK;
}

int cfd_upload(cfd_ctx *ctx, int field, const double *host, size_t len) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  size_t n = 0;
  int rc = field_len(c, field, &n);
  if (rc) return rc;
  if (!host || len != n) return fail(CFD_ERR_ARG, "upload length mismatch: expected " + std::to_string(n));
  adjoint_invalidate(c);
  CFD_HIP(hipMemcpyAsync(c.f[field], host, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  CFD_HIP(hipStreamSynchronize(c.stream));
  if (field == CFD_P) c.p_from_x = false;
  return CFD_OK;
}

int cfd_download(cfd_ctx *ctx, int field, double *host, size_t len) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  size_t n = 0;
  int rc = field_len(c, field, &n);
  if (rc) return rc;
  if (!host || len != n) return fail(CFD_ERR_ARG, "download length mismatch: expected " + std::to_string(n));
  if (field == CFD_P && (rc = sync_p(c))) return rc;
  CFD_HIP(hipMemcpyAsync(host, c.f[field], n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

// ---- set-up -----------------------------------------------------------------
int cfd_init_state(cfd_ctx *ctx, double noise_amp, uint64_t seed) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!std::isfinite(noise_amp) || noise_amp < 0.0) return fail(CFD_ERR_ARG, "noise amplitude must be >= 0");
  adjoint_invalidate(c);
  RET_LAUNCH(launch_init_fields(c, noise_amp, seed));
  RET_LAUNCH(launch_fill(c, c.px1, 0.0, c.g.N));
  RET_LAUNCH(launch_fill(c, c.px2, 0.0, c.g.N));
  c.loop = 1;
  c.time = 0.0;
  c.solves = 0;
  c.solves32 = 0;
  c.p_from_x = false;
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

int cfd_set_eta_cylinder_z(cfd_ctx *ctx) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  const double inv_n2 = std::pow((double)c.prm.ib_nsub, -2.0);
  RET_LAUNCH(launch_eta_cylinder_z(c, inv_n2));
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

int cfd_ell_assemble(cfd_ctx *ctx) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  RET_LAUNCH(launch_ell_assemble(c));
  c.ell_ready = true;
  dfree(c.colP); dfree(c.valP); dfree(c.valP32); dfree(c.perm);
  if (c.val32) RET_LAUNCH(launch_cast32(c, c.val, c.val32, (int64_t)ell_len(c.g.N)));
  int rc = CFD_OK;
  if (c.prm.precision == CFD_MIXED) rc = ensure_mixed(c);
  if (!rc && c.prm.solver == CFD_SOLVER_RBSOR) rc = ensure_perm(c);
  if (!rc && c.prm.solver == CFD_SOLVER_RBSOR && c.prm.precision == CFD_MIXED) rc = ensure_mixed(c);
  if (rc) return rc;
  if (c.prm.op == CFD_OP_SPE && c.spe_ready)  // gA.jp follows the configured operator
    RET_LAUNCH(launch_copy(c, c.spe_jp, c.f[CFD_JP], c.g.N));
  if (c.prm.op == CFD_OP_CSR && c.csr_ready) RET_LAUNCH(launch_copy(c, c.csr_jp, c.f[CFD_JP], c.g.N));
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

int cfd_spe_assemble(cfd_ctx *ctx) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  const DevGrid &g = c.g;
  std::vector<double> coef((size_t)2 * (g.nxc + g.nyc + g.nzc));
  double *xm = coef.data(), *xp = xm + g.nxc, *ym = xp + g.nxc, *yp = ym + g.nyc, *zm = yp + g.nyc, *zp = zm + g.nzc;
  for (int i = 0; i < g.nxc; ++i) {  // SPE_sparseMatrix1.hpp:62-75
    xp[i] = 1.0 / c.hDx[i + GC] / c.hDxs[i + GC];
    xm[i] = 1.0 / c.hDx[i + GC] / c.hDxs[i + GC - 1];
  }
  for (int j = 0; j < g.nyc; ++j) {
    yp[j] = 1.0 / c.hDy[j + GC] / c.hDys[j + GC];
    ym[j] = 1.0 / c.hDy[j + GC] / c.hDys[j + GC - 1];
  }
  for (int k = 0; k < g.nzc; ++k) {
    zp[k] = 1.0 / c.hDz[k + GC] / c.hDzs[k + GC];
    zm[k] = 1.0 / c.hDz[k + GC] / c.hDzs[k + GC - 1];
  }
  CFD_HIP(hipMemcpyAsync(c.spe_coef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice, c.stream));
  RET_LAUNCH(launch_spe_assemble(c));
  if (c.prm.op == CFD_OP_SPE) RET_LAUNCH(launch_copy(c, c.spe_jp, c.f[CFD_JP], g.N));
  CFD_HIP(hipStreamSynchronize(c.stream));
  c.spe_ready = true;
  return CFD_OK;
}

int cfd_csr_assemble(cfd_ctx *ctx) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  const DevGrid &g = c.g;
  if (g.N * 7 > (int64_t)INT32_MAX) return fail(CFD_ERR_ARG, "CSR operator: nnz exceeds 32-bit row pointers");
  // row pointers: one entry per row plus one per neighbour that exists
  std::vector<int32_t> ptr((size_t)g.N + 1);
  int64_t e = 0, r = 0;
  ptr[0] = 0;
  for (int i = 0; i < g.nxc; ++i)
    for (int j = 0; j < g.nyc; ++j)
      for (int k = 0; k < g.nzc; ++k) {
        e += 1 + (i != 0) + (j != 0) + (k != 0) + (k != g.nzc - 1) + (j != g.nyc - 1) + (i != g.nxc - 1);
        ptr[(size_t)++r] = (int32_t)e;
      }
  if (c.csr_nnz != e) {
    dfree(c.csr_col); dfree(c.csr_val);
    c.csr_nnz = 0;
  }
  int rc = CFD_OK;
  if (!c.csr_ptr && (rc = dmalloc(&c.csr_ptr, (size_t)g.N + 1))) return rc;
  if (!c.csr_jp && (rc = dmalloc(&c.csr_jp, (size_t)g.N))) return rc;
  if (!c.csr_col && ((rc = dmalloc(&c.csr_col, (size_t)e)) || (rc = dmalloc(&c.csr_val, (size_t)e)))) return rc;
  c.csr_nnz = e;
  CFD_HIP(hipMemcpyAsync(c.csr_ptr, ptr.data(), sizeof(int32_t) * ptr.size(), hipMemcpyHostToDevice, c.stream));
  RET_LAUNCH(launch_csr_assemble(c));
  if (c.prm.op == CFD_OP_CSR) RET_LAUNCH(launch_copy(c, c.csr_jp, c.f[CFD_JP], g.N));
  CFD_HIP(hipStreamSynchronize(c.stream));  // ptr's host copy must outlive the upload
  c.csr_ready = true;
  return CFD_OK;
}

int cfd_csr_nnz(cfd_ctx *ctx, size_t *nnz) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!nnz) return fail(CFD_ERR_ARG, "nnz is NULL");
  if (!c.csr_ready) return fail(CFD_ERR_STATE, "CSR operator not assembled");
  *nnz = (size_t)c.csr_nnz;
  return CFD_OK;
}

int cfd_csr_download(cfd_ctx *ctx, int32_t *ptr, int32_t *col, double *val, size_t n_rows, size_t nnz) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!c.csr_ready) return fail(CFD_ERR_STATE, "CSR operator not assembled");
  if (!ptr || !col || !val || n_rows != (size_t)c.g.N || nnz != (size_t)c.csr_nnz)
    return fail(CFD_ERR_ARG, "CSR shape mismatch");
  CFD_HIP(hipMemcpyAsync(ptr, c.csr_ptr, sizeof(int32_t) * (n_rows + 1), hipMemcpyDeviceToHost, c.stream));
  CFD_HIP(hipMemcpyAsync(col, c.csr_col, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, c.stream));
  CFD_HIP(hipMemcpyAsync(val, c.csr_val, sizeof(double) * nnz, hipMemcpyDeviceToHost, c.stream));
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

int cfd_ell_download(cfd_ctx *ctx, int32_t *col, double *val, size_t n_rows, int k) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!c.ell_ready) return fail(CFD_ERR_STATE, "ELL operator not assembled");
  if (!col || !val || n_rows != (size_t)c.g.N || k != KSLOTS) return fail(CFD_ERR_ARG, "ELL shape mismatch");
  // the device arrays are sliced; the caller gets the reference's slot-major layout
  const int64_t N = c.g.N;
  std::vector<int32_t> hc(ell_len(N));
  std::vector<double> hv(ell_len(N));
  CFD_HIP(hipMemcpyAsync(hc.data(), c.col, sizeof(int32_t) * hc.size(), hipMemcpyDeviceToHost, c.stream));
  CFD_HIP(hipMemcpyAsync(hv.data(), c.val, sizeof(double) * hv.size(), hipMemcpyDeviceToHost, c.stream));
  CFD_HIP(hipStreamSynchronize(c.stream));
  const int64_t st = ell_stride(N, c.ell_sliced);
  for (int64_t r = 0; r < N; ++r) {
    const int64_t b = ell_base(r, c.ell_sliced);
    for (int s = 0; s < KSLOTS; ++s) {
      col[(size_t)s * N + r] = hc[(size_t)(b + s * st)];
      val[(size_t)s * N + r] = hv[(size_t)(b + s * st)];
    }
  }
  return CFD_OK;
}

int cfd_ell_upload(cfd_ctx *ctx, const int32_t *col, const double *val, size_t n_rows, int k) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!col || !val || n_rows != (size_t)c.g.N || k != KSLOTS) return fail(CFD_ERR_ARG, "ELL shape mismatch");
  for (size_t q = 0; q < (size_t)KSLOTS * n_rows; ++q)
    if (col[q] < 0 || (size_t)col[q] >= n_rows) return fail(CFD_ERR_ARG, "ELL column index out of range");
  // slot-major from the caller into the library's sliced arrays (the last
  // slice's padding rows: column 0, value 0, never read)
  const int64_t N = c.g.N;
  std::vector<int32_t> hc(ell_len(N), 0);
  std::vector<double> hv(ell_len(N), 0.0);
  const int64_t st = ell_stride(N, c.ell_sliced);
  for (int64_t r = 0; r < N; ++r) {
    const int64_t b = ell_base(r, c.ell_sliced);
    for (int s = 0; s < KSLOTS; ++s) {
      hc[(size_t)(b + s * st)] = col[(size_t)s * N + r];
      hv[(size_t)(b + s * st)] = val[(size_t)s * N + r];
    }
  }
  CFD_HIP(hipMemcpyAsync(c.col, hc.data(), sizeof(int32_t) * hc.size(), hipMemcpyHostToDevice, c.stream));
  CFD_HIP(hipMemcpyAsync(c.val, hv.data(), sizeof(double) * hv.size(), hipMemcpyHostToDevice, c.stream));
  CFD_HIP(hipStreamSynchronize(c.stream));  // the staging vectors go out of scope
  c.ell_ready = true;
  dfree(c.colP); dfree(c.valP); dfree(c.valP32); dfree(c.perm);
  if (c.val32) RET_LAUNCH(launch_cast32(c, c.val, c.val32, (int64_t)ell_len(c.g.N)));
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

// ---- operators / solves -----------------------------------------------------
int cfd_spmv(cfd_ctx *ctx, int op, int precision, int x_field, int y_field) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  const int vec_ok[] = {CFD_RHS, CFD_X, CFD_JP, CFD_VEC_A, CFD_VEC_B};
  auto is_vec = [&](int f) { return std::find(std::begin(vec_ok), std::end(vec_ok), f) != std::end(vec_ok); };
  if (!is_vec(x_field) || !is_vec(y_field) || x_field == y_field)
    return fail(CFD_ERR_ARG, "x and y must be distinct N-sized vector fields");
  if (op != CFD_OP_ELL && op != CFD_OP_SPE && op != CFD_OP_CSR) return fail(CFD_ERR_ARG, "unknown operator");
  if (int rc = op_ready(c, op)) return rc;
  if (precision == CFD_MIXED) {
    if (op != CFD_OP_ELL) return fail(CFD_ERR_ARG, "FP32 SpMV is defined for the ELL operator");
    int rc = ensure_mixed(c);
    if (rc) return rc;
  }
  RET_LAUNCH(launch_spmv(c, op, precision, c.f[x_field], c.f[y_field]));
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

int cfd_bicgstab(cfd_ctx *ctx, double tol, int maxit, int *iters, double *relres) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!iters || !relres || maxit < 0 || !(tol >= 0.0)) return fail(CFD_ERR_ARG, "bad bicgstab arguments");
  return bicgstab_solve(c, c.f[CFD_RHS], c.f[CFD_X], tol, maxit, iters, relres);
}

int cfd_rbsor(cfd_ctx *ctx, double omega, double tol, int maxit, int check_every, int *iters, double *relres) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!iters || !relres || maxit < 0 || !(omega > 0.0 && omega < 2.0) || !(tol >= 0.0))
    return fail(CFD_ERR_ARG, "bad rbsor arguments");
  if (c.prm.op == CFD_OP_CSR)
    return rbsor_csr_solve(c, c.f[CFD_RHS], c.f[CFD_X], omega, tol, maxit, check_every, iters, relres);
  int rc = ensure_perm(c);
  if (rc) return rc;
  return rbsor_solve(c, c.f[CFD_RHS], c.f[CFD_X], omega, tol, maxit, check_every, iters, relres);
}

int cfd_sor_pipeline(cfd_ctx *ctx, double omega, double crit, int itmax, int *iters, double *error) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!iters || !error || itmax < 1 || !(omega > 0.0 && omega < 2.0) || !(crit >= 0.0))
    return fail(CFD_ERR_ARG, "bad SorPipeLine arguments");
  int rc = sync_p(c);
  if (rc) return rc;
  RET_LAUNCH(launch_periodic_refresh(c, c.f[CFD_T1_U], c.f[CFD_T1_V], c.f[CFD_T1_W]));
  return sor_pipeline_solve(c, omega, crit, itmax, iters, error);
}

int cfd_defect_correction(cfd_ctx *ctx, double tol, double tol_inner, int max_outer, int *outer, int *inner_total,
                          double *relres) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!outer || !inner_total || !relres || max_outer < 0 || !(tol >= 0.0) || !(tol_inner > 0.0))
    return fail(CFD_ERR_ARG, "bad defect-correction arguments");
  int rc = ensure_mixed(c);
  if (!rc && c.prm.solver == CFD_SOLVER_RBSOR) {
    rc = ensure_perm(c);
    if (!rc) rc = ensure_mixed(c);
  }
  if (rc) return rc;
  return defect_correction(c, c.f[CFD_RHS], c.f[CFD_X], tol, tol_inner, max_outer, outer, inner_total, relres);
}

int cfd_solve_pressure(cfd_ctx *ctx, int *iters, double *relres) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!iters || !relres) return fail(CFD_ERR_ARG, "NULL output");
  int inner = 0;
  return solve_configured(c, iters, relres, &inner);
}

// ---- timestep phases (reference semantics) -----------------------------------
int cfd_predictor(cfd_ctx *ctx) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  RET_LAUNCH(launch_predictor(c, c.f[CFD_T0_U], c.f[CFD_T0_V], c.f[CFD_T0_W], c.f[CFD_T1_U], c.f[CFD_T1_V],
                              c.f[CFD_T1_W], nullptr, 0));
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

int cfd_source_add(cfd_ctx *ctx) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  RET_LAUNCH(launch_source_add(c, c.f[CFD_T1_U], c.f[CFD_T1_V], c.f[CFD_T1_W], c.f[CFD_SRC],
                               c.prm.source_index_quirk));
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

int cfd_build_rhs(cfd_ctx *ctx) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  RET_LAUNCH(launch_periodic_refresh(c, c.f[CFD_T1_U], c.f[CFD_T1_V], c.f[CFD_T1_W]));
  RET_LAUNCH(launch_rhs(c, c.f[CFD_T1_U], c.f[CFD_T1_V], c.f[CFD_T1_W], c.f[CFD_JP], c.f[CFD_RHS]));
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

int cfd_pressure_scatter(cfd_ctx *ctx) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  int rc = sync_p(c);
  if (rc) return rc;
  RET_LAUNCH(launch_scatter_p(c, c.f[CFD_X], c.f[CFD_P]));
  RET_LAUNCH(launch_periodic_refresh_p(c, c.f[CFD_P]));  // periodic z: the seam's ghosts follow the interior
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

int cfd_dfib_update(cfd_ctx *ctx) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  int rc = sync_p(c);
  if (rc) return rc;
  RET_LAUNCH(launch_dfib_update_into(c, UPD_WRITE_T3, c.f[CFD_P], nullptr, nullptr));
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

int cfd_source_accumulate(cfd_ctx *ctx) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  RET_LAUNCH(launch_source_accumulate(c));
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

int cfd_bc_main(cfd_ctx *ctx) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  int rc = sync_p(c);
  if (rc) return rc;
  RET_LAUNCH(launch_bc(c, true, false, true));
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

int cfd_bc_copy(cfd_ctx *ctx) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  RET_LAUNCH(launch_bc(c, false, true, false));
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

int cfd_drag(cfd_ctx *ctx, double *cd, double *cl) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!cd || !cl) return fail(CFD_ERR_ARG, "NULL output");
  RET_LAUNCH(launch_drag(c));
  CFD_HIP(hipMemcpyAsync(c.hred, c.red.out + RS_DRAG * 4, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  CFD_HIP(hipStreamSynchronize(c.stream));
  *cd = -2.0 * c.hred[0] / c.g.lz;
  *cl = -2.0 * c.hred[1] / c.g.lz;
  return CFD_OK;
}

int cfd_force_sum(cfd_ctx *ctx, int which_direction, double *sum) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!sum) return fail(CFD_ERR_ARG, "NULL output");
  if (which_direction < 0 || which_direction > 2) return fail(CFD_ERR_ARG, "whichDirection must be 0, 1 or 2");
  RET_LAUNCH(launch_force_sum(c, which_direction, c.red.out + RS_DRAG * 4 + 2));
  CFD_HIP(hipMemcpyAsync(c.hred, c.red.out + RS_DRAG * 4 + 2, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  CFD_HIP(hipStreamSynchronize(c.stream));
  *sum = c.hred[0];
  return CFD_OK;
}

int cfd_maxdiff(cfd_ctx *ctx, double out[3]) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!out) return fail(CFD_ERR_ARG, "NULL output");
  RET_LAUNCH(launch_maxdiff(c));
  CFD_HIP(hipMemcpyAsync(c.hred, c.red.out + RS_MAX * 4, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  CFD_HIP(hipStreamSynchronize(c.stream));
  for (int q = 0; q < 3; ++q) out[q] = c.hred[q];
  return CFD_OK;
}

int cfd_divergence(cfd_ctx *ctx, int vel_field, double out[4]) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!out) return fail(CFD_ERR_ARG, "NULL output");
  if (vel_field != CFD_T0_U && vel_field != CFD_T1_U && vel_field != CFD_T3_U)
    return fail(CFD_ERR_ARG, "vel_field must be CFD_T0_U, CFD_T1_U or CFD_T3_U");
  const double *vel[3] = {c.f[vel_field], c.f[vel_field + 1], c.f[vel_field + 2]};
  RET_LAUNCH(launch_divergence(c, vel, vel, c.red.out + RS_MISC * 4));
  CFD_HIP(hipMemcpyAsync(c.hred, c.red.out + RS_MISC * 4, 4 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  CFD_HIP(hipStreamSynchronize(c.stream));
  out[0] = c.hred[0];
  out[1] = std::sqrt(c.hred[2]);
  out[2] = c.hred[1];
  out[3] = std::sqrt(c.hred[3]);
  return CFD_OK;
}

int cfd_copy_t3_to_t0(cfd_ctx *ctx) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  for (int q = 0; q < 3; ++q) RET_LAUNCH(launch_copy(c, c.f[CFD_T3_U + q], c.f[CFD_T0_U + q], c.g.G));
  CFD_HIP(hipStreamSynchronize(c.stream));
  return CFD_OK;
}

// ---- whole timestep ------------------------------------------------------------
// runSeqOmp_predictionMethod main loop (RUN_CPU_PC_METHOD.hpp:287-627):
//   zero source_termF                    -> folded into pass 0 (source = dt f)
//   passes 0..inner-1: predictor + source add (one kernel), RHS, solve,
//     DFIB update + source accumulate + drag partials (one kernel; the
//     pressure is read from X_result, so the scatter is deferred until the
//     ghosted p is needed)
//   corrector: predictor + source add, RHS, solve, DFIB update writing the
//     new velocity into T0 and T3 with the max-difference check fused
//   BC_staggered_main on T0 fused with BC_staggered_copy into T1
// The predictor is re-evaluated every pass, as in the reference; T0 does not
// change inside a step, so every evaluation is bit-identical.
int cfd_step(cfd_ctx *ctx, cfd_step_stats *stats) {
  CTX_CHECK(ctx);
  adjoint_invalidate(*reinterpret_cast<Ctx *>(ctx));
  return step_impl(*reinterpret_cast<Ctx *>(ctx), stats);
}

}  // extern "C"

namespace cfd {

int op_ready(const Ctx &c, int op) {
  if (op == CFD_OP_ELL && !c.ell_ready) return fail(CFD_ERR_STATE, "ELL operator not assembled");
  if (op == CFD_OP_SPE && !c.spe_ready) return fail(CFD_ERR_STATE, "SPE operator not assembled");
  if (op == CFD_OP_CSR && !c.csr_ready) return fail(CFD_ERR_STATE, "CSR operator not assembled");
  return CFD_OK;
}

int step_impl(Ctx &c, cfd_step_stats *stats) {
  const cfd_params &P = c.prm;
  const bool solve = c.loop > 1;
  if (solve && P.solver != CFD_SOLVER_SOR_PIPELINE) {
    if (int rc = op_ready(c, P.op)) return rc;
  }
  cfd_step_stats st;
  std::memset(&st, 0, sizeof(st));
  st.loop = c.loop;
  st.n_passes = P.inner_loops + 1;
  double *u0 = c.f[CFD_T0_U], *v0 = c.f[CFD_T0_V], *w0 = c.f[CFD_T0_W];
  double *u1 = c.f[CFD_T1_U], *v1 = c.f[CFD_T1_V], *w1 = c.f[CFD_T1_W];
  // phase boundaries (cfd_phase_timing): event k closes phase ev_phase[k]
  int nev = 0;
  int ev_phase[2 + 3 * CFD_MAX_PASSES];
  auto mark = [&](int phase) -> hipError_t {
    if (!c.phase_timing) return hipSuccess;
    ev_phase[nev] = phase;
    return hipEventRecord(c.phase_ev[nev++], c.stream);
  };
  CFD_HIP(mark(-1));
  for (int pass = 0; pass <= P.inner_loops; ++pass) {
    const bool corrector = pass == P.inner_loops;
    const bool with_src = pass > 0 && P.inner_loops > 0;
    const int src_mode = with_src ? (P.source_index_quirk ? 2 : 1) : 0;
    if (src_mode == 2 || P.inner_loops == 0) {
      // the reference's index mapping adds the source outside the predictor's
      // positions too: the predictor is re-evaluated every pass
      RET_LAUNCH(launch_predictor(c, u0, v0, w0, u1, v1, w1, c.f[CFD_SRC], src_mode));
      if (src_mode == 2) RET_LAUNCH(launch_source_quirk_fixup(c, u1, v1, w1, c.f[CFD_SRC]));
    } else if (pass == 0) {
      // T0 is the same in every pass: predict once, keep P for the other passes
      for (int q = 0; q < 3; ++q)
        if (!c.pred[q]) {
          int rc = dmalloc(&c.pred[q], (size_t)c.g.G);
          if (rc) return rc;
        }
      RET_LAUNCH(launch_predictor(c, u0, v0, w0, u1, v1, w1, c.f[CFD_SRC], 0, c.pred));
    } else {
      RET_LAUNCH(launch_pred_add(c, c.pred, c.f[CFD_SRC], u1, v1, w1));
    }
    // periodic z: T1's ghost layers take the new interior's images before
    // the right-hand side reads them
    RET_LAUNCH(launch_periodic_refresh(c, u1, v1, w1));
    CFD_HIP(mark(CFD_PHASE_CONVECTION));
    st.iters[pass] = -1;
    if (solve) {
      const bool sor_p = P.solver == CFD_SOLVER_SOR_PIPELINE;  // in place on the ghosted p, its own RHS
      if (!sor_p) RET_LAUNCH(launch_rhs(c, u1, v1, w1, c.f[CFD_JP], c.f[CFD_RHS]));
      int it = 0, inner = 0;
      double rel = 0.0;
      int rc = solve_configured(c, &it, &rel, &inner);
      if (rc) return rc;
      st.iters[pass] = it;
      st.relres[pass] = rel;
      st.inner_iters[pass] = inner;
      c.p_from_x = !sor_p;
      // SorPipeLine leaves the ghosted p: periodic z refreshes the seam's
      // ghosts the update reads (the X-reading update wraps by itself)
      if (sor_p) RET_LAUNCH(launch_periodic_refresh_p(c, c.f[CFD_P]));
    }
    CFD_HIP(mark(CFD_PHASE_PRESSURE));
    int mode = UPD_DRAG;
    if (c.p_from_x) mode |= UPD_P_FROM_X;
    if (corrector) mode |= UPD_WRITE_T3 | UPD_WRITE_T0 | UPD_MAXDIFF;
    else mode |= (pass == 0 ? UPD_SRC_SET : UPD_SRC_ADD);
    const double *pp = c.p_from_x ? c.f[CFD_X] : c.f[CFD_P];
    double *t2 = c.t2_rec ? c.t2_rec + (size_t)3 * c.g.N * pass : nullptr;
    RET_LAUNCH(
        launch_dfib_update_into(c, mode, pp, c.stepbuf + 2 * pass, corrector ? c.stepbuf + 32 : nullptr, t2));
    if (corrector) {
      // FP64 post-projection divergence check: the corrected velocity (now
      // in T0) on the boundary faces the projection saw (T1's)
      const double *vel[3] = {u0, v0, w0}, *bnd[3] = {u1, v1, w1};
      RET_LAUNCH(launch_divergence(c, vel, bnd, c.stepbuf + 36));
    }
    CFD_HIP(mark(CFD_PHASE_UPDATE));
  }
  RET_LAUNCH(launch_bc(c, true, true, !c.p_from_x));
  CFD_HIP(mark(CFD_PHASE_BC));
  CFD_HIP(hipMemcpyAsync(c.hred, c.stepbuf, 40 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  CFD_HIP(hipStreamSynchronize(c.stream));
  if (c.phase_timing) {
    for (int k = 1; k < nev; ++k) {
      float ms = 0.0f;
      CFD_HIP(hipEventElapsedTime(&ms, c.phase_ev[k - 1], c.phase_ev[k]));
      c.phase_ms[ev_phase[k]] += ms;
    }
    c.phase_steps += 1;
  }
  for (int pass = 0; pass <= P.inner_loops; ++pass) {
    st.cd[pass] = -2.0 * c.hred[2 * pass] / c.g.lz;
    st.cl[pass] = -2.0 * c.hred[2 * pass + 1] / c.g.lz;
  }
  for (int q = 0; q < 3; ++q) st.maxdiff[q] = c.hred[32 + q];
  st.div_max = c.hred[36];
  st.div_fluid_max = c.hred[37];
  st.div_l2 = std::sqrt(c.hred[38]);
  st.div_fluid_l2 = std::sqrt(c.hred[39]);
  c.time += P.dt;  // simuClass::finishloop (simuClass.hpp:152-169)
  c.loop += 1;
  if (stats) *stats = st;
  return CFD_OK;
}

}  // namespace cfd

extern "C" {

int cfd_loop(const cfd_ctx *ctx, int *loop, double *time) {
  if (!ctx) return fail(CFD_ERR_ARG, "context is NULL");
  const Ctx &c = *reinterpret_cast<const Ctx *>(ctx);
  if (loop) *loop = c.loop;
  if (time) *time = c.time;
  return CFD_OK;
}

int cfd_set_loop(cfd_ctx *ctx, int loop, double time) {
  if (!ctx) return fail(CFD_ERR_ARG, "context is NULL");
  if (loop < 1) return fail(CFD_ERR_ARG, "loop starts at 1");
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  adjoint_invalidate(c);
  c.loop = loop;
  c.time = time;
  return CFD_OK;
}

int cfd_phase_timing(cfd_ctx *ctx, int on) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  // one event per boundary: the step's start, three per pass, the BC's end
  const size_t need = 2 + 3 * CFD_MAX_PASSES;
  while (on && c.phase_ev.size() < need) {
    hipEvent_t e = nullptr;
    CFD_HIP(hipEventCreate(&e));
    c.phase_ev.push_back(e);
  }
  c.phase_timing = on != 0;
  for (double &m : c.phase_ms) m = 0.0;
  c.phase_steps = 0;
  return CFD_OK;
}

int cfd_phase_times(cfd_ctx *ctx, double ms[CFD_NPHASES], int *steps) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!ms) return fail(CFD_ERR_ARG, "ms is NULL");
  for (int q = 0; q < CFD_NPHASES; ++q) ms[q] = c.phase_ms[q];
  if (steps) *steps = c.phase_steps;
  return CFD_OK;
}

int cfd_get_solver_counters(cfd_ctx *ctx, cfd_solver_counters *out) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!out) return fail(CFD_ERR_ARG, "out is NULL");
  KrylovScalars h;
  CFD_HIP(hipMemcpyAsync(&h, c.ks, sizeof(h), hipMemcpyDeviceToHost, c.stream));
  CFD_HIP(hipStreamSynchronize(c.stream));
  out->solves = c.solves;
  out->restarts = h.restarts;
  out->resets = c.nan_resets;
  out->pad = 0;
  out->iters = c.total_iters;
  return CFD_OK;
}

// ---- measurement -----------------------------------------------------------------
int cfd_time_kernel(cfd_ctx *ctx, int what, int nrep, double *ms_avg) {
  CTX_CHECK(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!ms_avg || nrep < 1) return fail(CFD_ERR_ARG, "bad timing arguments");
  int rc = CFD_OK;
  if ((what == 0 || what == 1 || what == 3 || what == 4 || what == 8) && !c.ell_ready && c.prm.op == CFD_OP_ELL)
    return fail(CFD_ERR_STATE, "ELL operator not assembled");
  if (what == 1 && (rc = ensure_mixed(c))) return rc;
  if (what == 2 && !c.spe_ready) return fail(CFD_ERR_STATE, "SPE operator not assembled");
  if (what == 7 && !c.csr_ready) return fail(CFD_ERR_STATE, "CSR operator not assembled");
  if (what == 3 && (rc = ensure_perm(c))) return rc;
  if (what < 0 || what > 8) return fail(CFD_ERR_ARG, "unknown kernel id");
  auto body = [&]() -> hipError_t {
    switch (what) {
      case 0: return launch_spmv(c, CFD_OP_ELL, CFD_FP64, c.f[CFD_X], c.f[CFD_VEC_A]);
      case 1: return launch_spmv_ell_f32(c, c.q32[0], c.q32[1]);
      case 2: return launch_spmv(c, CFD_OP_SPE, CFD_FP64, c.f[CFD_X], c.f[CFD_VEC_A]);
      case 3: return launch_rbsor_sweep(c, c.prm.sor_omega);
      case 4: return launch_bicg_iteration(c);
      case 7: return launch_spmv(c, CFD_OP_CSR, CFD_FP64, c.f[CFD_X], c.f[CFD_VEC_A]);
      case 8: return launch_spmv_dot(c);
      case 5:
        return launch_predictor(c, c.f[CFD_T0_U], c.f[CFD_T0_V], c.f[CFD_T0_W], c.f[CFD_T1_U], c.f[CFD_T1_V],
                                c.f[CFD_T1_W], nullptr, 0);
      default: return launch_copy_kernel_bench(c, c.f[CFD_VEC_A], c.f[CFD_VEC_B], c.g.N);
    }
  };
  if (what == 4 || what == 8) RET_LAUNCH(launch_bicg_timing_setup(c));  // r0 = b - A*0 on scratch vectors, no stop test
  RET_LAUNCH(body());  // warm-up
  CFD_HIP(hipEventRecord(c.ev0, c.stream));
  for (int r = 0; r < nrep; ++r) RET_LAUNCH(body());
  CFD_HIP(hipEventRecord(c.ev1, c.stream));
  CFD_HIP(hipEventSynchronize(c.ev1));
  float ms = 0.0f;
  CFD_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  *ms_avg = (double)ms / nrep;
  return CFD_OK;
}

}  // extern "C"
