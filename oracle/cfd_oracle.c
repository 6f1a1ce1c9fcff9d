/*
 * cfd_oracle.c -- TEST INFRASTRUCTURE ONLY: FP64 CPU restatement of the
 * reference hot path (jimmy3769tw/CFD_NSsolver_forHPC).  Parity checker for
 * the HIP library and timed CPU baseline; never linked into the product.
 *
 * Built with -ffp-contract=off: every rounding below is the one written in
 * the source.  Where the reference build contracts a multiply-add into an FMA
 * and the result is a bit-exact target (IB mask sub-sampling, ELL SpMV), the
 * fma() is written out and commented.
 *
 * Parity status per function (pins live in tests/):
 *   orc_axis_*              pinned: closed form + compiled reference (native SML grid)
 *   orc_eta_cylinder_z      pinned: compiled reference hash + area identity
 *   orc_ell_assemble        pinned: compiled reference hash + row-sum invariant
 *   orc_spmv_ell / _spe     pinned: dense brute force + compiled reference
 *   orc_predictor           pinned: compiled reference + closed forms
 *   orc_rhs, orc_dfib_update, orc_bc_*, orc_drag, orc_maxdiff, orc_source_*:
 *                           pinned: compiled reference step dumps + identities
 *   orc_bicg_*              pinned: compiled reference trajectory + dense solve
 *   orc_rbsor               pinned: fixed point equals the dense solve
 *   orc_sor_lex             pinned: fixed point equals the dense solve of the
 *                           symmetric form
 *   orc_sim_step            pinned: compiled reference PC-method loop
 *   orc_axis_tanh           no reference counterpart (large config only):
 *                           pinned by its invariants (sum = L, ratio bound)
 */
#include "cfd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GC 2

typedef struct {
  int nx, ny, nz;    /* ghosted sizes */
  int nxc, nyc, nzc; /* interior sizes */
  size_t nzny, G, N;
} dims_t;

static dims_t dims_of(const cfd_grid_desc *g) {
  dims_t d;
  d.nxc = g->nx_cal; d.nyc = g->ny_cal; d.nzc = g->nz_cal;
  d.nx = d.nxc + 2 * GC; d.ny = d.nyc + 2 * GC; d.nz = d.nzc + 2 * GC;
  d.nzny = (size_t)d.ny * d.nz;
  d.G = (size_t)d.nx * d.nzny;
  d.N = (size_t)d.nxc * d.nyc * d.nzc;
  return d;
}

/* grid::icel (structureGrid.hpp:159-161) and grid::icelCal (:182-185) */
#define IC(d, i, j, k) (((size_t)(i) * (d).ny + (size_t)(j)) * (d).nz + (size_t)(k))
#define ICAL(d, i, j, k) \
  (((size_t)((i) - GC) * (d).nyc + (size_t)((j) - GC)) * (d).nzc + (size_t)((k) - GC))

/* ========================================================================
 * grid
 * ===================================================================== */

/* grid::init_uniform, one axis (structureGrid.hpp:409-440) */
void orc_axis_uniform(int ncal, double L, double *X, double *D, double *Ds) {
  const int n = ncal + 2 * GC;
  const double d = L / (n - 2.0 * GC);
  X[0] = 0.0 - 2.0 * d;
  for (int i = 1; i < n + 1; ++i) X[i] = X[i - 1] + d;
  for (int i = 0; i < n; ++i) { D[i] = d; Ds[i] = d; }
}

/* grid::init_gC_dx, one axis (structureGrid.hpp:506-559) */
void orc_axis_gc_dx(int n, double *X, double *D, double *Ds) {
  X[1] = 2 * X[2] - X[3];
  X[0] = 2 * X[1] - X[2];
  X[n - 1] = 2 * X[n - 2] - X[n - 3];
  X[n] = 2 * X[n - 1] - X[n - 2];
  for (int i = GC; i < n - GC - 1; ++i) Ds[i] = (X[i + 2] - X[i]) / 2.0;
  for (int i = GC; i < n - GC - 1; ++i) D[i] = (X[i + 1] - X[i]);
  D[0] = D[1] = D[2];
  Ds[0] = Ds[1] = Ds[2];
  D[n - 1] = D[n - 2] = D[n - 3] = X[n - 2] - X[n - 3];
  Ds[n - 1] = Ds[n - 2] = Ds[n - 3] = Ds[n - 4];
}

/* grid::unit_SML (structureGrid.hpp:326-386) for one axis */
int orc_axis_sml(int ncal, double L, double lS, int nS, int off_nS, double *X, double *D, double *Ds) {
  const int n = ncal + 2 * GC;
  memset(X, 0, sizeof(double) * (size_t)(n + 1));
  X[GC] = 0.0;
  if (nS != 0) {
    const double dS = lS / nS;
    const double d = (L - lS) / (ncal - nS);
    if ((n - nS) % 2 != 0) return -1;
    int nlEnd0 = (ncal - nS) / 2;
    int nsEnd0 = nlEnd0 + nS;
    const int nlEnd1 = nsEnd0 + nlEnd0;
    nlEnd0 += off_nS;
    nsEnd0 += off_nS;
    if (nlEnd0 <= 0) return -1;
    for (int i = 0; i < nlEnd0; ++i) X[i + GC + 1] = X[i + GC] + d;
    for (int i = nlEnd0; i < nsEnd0; ++i) X[i + GC + 1] = X[i + GC] + dS;
    for (int i = nsEnd0; i < nlEnd1; ++i) X[i + GC + 1] = X[i + GC] + d;
  } else {
    for (int i = GC; i < n - GC; ++i) X[i + 1] = X[i] + L / ncal;
  }
  orc_axis_gc_dx(n, X, D, Ds);
  return 0;
}

/* Stretched axis for the large configuration (no reference counterpart):
   tanh clustering around xc, restated from the library.  Cell width over the
   node index offset xi: dx/dxi = H - (H - h0)/2 [tanh((xi + a)/w) -
   tanh((xi - a)/w)], a = band / (2 h0); node positions
   x(xi) = xc + H xi - (H - h0)/2 w [lncosh((xi + a)/w) - lncosh((xi - a)/w)];
   node i at xi0 + i with xi0 (node 0 at 0) and H (node ncal at L) by
   bisection, w the narrowest step with neighbouring-width ratio <= ratio. */
static double lncosh_(double t) {
  const double a = fabs(t);
  return a + log1p(exp(-2.0 * a)) - log(2.0);
}
typedef struct { double h0, a, w, H, xc; } tanh_axis;
static double tanh_f(const tanh_axis *t, double xi) {
  return t->H * xi - 0.5 * (t->H - t->h0) * t->w * (lncosh_((xi + t->a) / t->w) - lncosh_((xi - t->a) / t->w));
}
static double tanh_xi0(const tanh_axis *t, int ncal) {
  double lo = -(double)ncal - 1.0, hi = 0.0;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (t->xc + tanh_f(t, mid) > 0.0) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}
static double tanh_fit(tanh_axis *t, int ncal, double L, double *xi0_out) {
  double lo = t->h0, hi = L, xi0 = 0.0;
  for (int it = 0; it < 200; ++it) {
    t->H = 0.5 * (lo + hi);
    xi0 = tanh_xi0(t, ncal);
    if (t->xc + tanh_f(t, xi0 + ncal) > L) hi = t->H;
    else lo = t->H;
  }
  t->H = 0.5 * (lo + hi);
  xi0 = tanh_xi0(t, ncal);
  double rmax = 1.0, prev = tanh_f(t, xi0 + 1.0) - tanh_f(t, xi0);
  for (int c = 1; c < ncal; ++c) {
    const double h = tanh_f(t, xi0 + c + 1.0) - tanh_f(t, xi0 + c);
    const double r = h / prev > prev / h ? h / prev : prev / h;
    if (r > rmax) rmax = r;
    prev = h;
  }
  *xi0_out = xi0;
  return rmax;
}

int orc_axis_tanh(int ncal, double L, double xc, double h0, double band, double ratio,
                  double *X, double *D, double *Ds) {
  if (ncal < 4 || !(L > 0) || !(h0 > 0) || !(ratio > 1.0) || h0 * ncal >= L || !(xc > 0 && xc < L) || !(band > 0))
    return -1;
  const int n = ncal + 2 * GC;
  tanh_axis t = {h0, 0.5 * band / h0, 1.0, h0, xc};
  double xi0 = 0.0, wlo = 0.25, whi = (double)ncal;
  t.w = whi;
  if (tanh_fit(&t, ncal, L, &xi0) > ratio) return -1;
  for (int it = 0; it < 60; ++it) {
    t.w = 0.5 * (wlo + whi);
    if (tanh_fit(&t, ncal, L, &xi0) > ratio) wlo = t.w;
    else whi = t.w;
  }
  t.w = whi;
  tanh_fit(&t, ncal, L, &xi0);
  for (int q = 0; q <= n; ++q) X[q] = 0.0;
  for (int c = 1; c < ncal; ++c) X[GC + c] = xc + tanh_f(&t, xi0 + c);
  X[GC] = 0.0;
  X[GC + ncal] = L;
  orc_axis_gc_dx(n, X, D, Ds);
  return 0;
}

/* ========================================================================
 * reductions and noise
 * ===================================================================== */

#define ORC_NBLK 256

/* Fixed 256-way blocked sum with a serial combine: the same bits for every
   OpenMP thread count (the reference's omp reduction order is thread-dependent,
   inc/solver/math.hpp:37-51). */
double orc_dot(size_t n, const double *a, const double *b) {
  double part[ORC_NBLK];
#pragma omp parallel for schedule(static)
  for (int blk = 0; blk < ORC_NBLK; ++blk) {
    size_t lo = n * (size_t)blk / ORC_NBLK, hi = n * (size_t)(blk + 1) / ORC_NBLK;
    double s = 0.0;
    for (size_t i = lo; i < hi; ++i) s += a[i] * b[i];
    part[blk] = s;
  }
  double r = 0.0;
  for (int blk = 0; blk < ORC_NBLK; ++blk) r += part[blk];
  return r;
}

static uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double orc_noise(uint64_t seed, int component, uint64_t icel, double amp) {
  uint64_t z = splitmix64(seed ^ ((uint64_t)component << 56) ^ icel);
  double u01 = (double)(z >> 11) * 0x1.0p-53;
  return amp * (2.0 * u01 - 1.0);
}

void orc_fill_noise(double *v, size_t n, uint64_t seed, int tag, double base, double amp) {
#pragma omp parallel for schedule(static)
  for (long long m = 0; m < (long long)n; ++m) v[m] = base + orc_noise(seed, tag, (uint64_t)m, amp);
}

/* ========================================================================
 * DFIB_CylinderZ (inc/dfib/setEta.hpp:136-204)
 * ===================================================================== */
void orc_eta_cylinder_z(const cfd_grid_desc *g, double cx, double cy, double R, int nsub, double *eta) {
  const dims_t d = dims_of(g);
  const double inv_n2 = pow((double)nsub, -2.0);
  memset(eta, 0, sizeof(double) * d.G);
#pragma omp parallel for schedule(dynamic, 4)
  for (int i = GC; i < d.nx - GC; ++i) {
    double *sdx = (double *)malloc(sizeof(double) * (size_t)(nsub + 1));
    double *sdy = (double *)malloc(sizeof(double) * (size_t)(nsub + 1));
    for (int j = GC; j < d.ny - GC; ++j) {
      const size_t icel = (size_t)i * d.nz * d.ny + (size_t)j * d.nz;
      /* the reference build contracts a*a + b*b into fma(a, a, b*b) */
      const double Diagonal = sqrt(fma(g->Dx[i], g->Dx[i], g->Dy[j] * g->Dy[j])) * 0.5;
      const double ax = ((g->X[i] + g->X[i + 1]) * 0.5) - cx;
      const double ay = ((g->Y[j] + g->Y[j + 1]) * 0.5) - cy;
      const double Distance = sqrt(fma(ax, ax, ay * ay));
      if (fabs(Distance - R) < Diagonal) {
        size_t xi = 0;
        const double dyg = g->Dy[j] / (double)nsub;
        const double dxg = g->Dx[i] / (double)nsub;
        for (int q = 0; q < nsub + 1; ++q) {
          sdy[q] = fma((double)q, dyg, g->Y[j]); /* contracted Y[j] + jj*dyg */
          sdx[q] = fma((double)q, dxg, g->X[i]);
        }
        for (int ii = 0; ii < nsub; ++ii) {
          const double bx = (sdx[ii] + sdx[ii + 1]) * 0.5 - cx;
          for (int jj = 0; jj < nsub; ++jj) {
            const double by = (sdy[jj] + sdy[jj + 1]) * 0.5 - cy;
            const double Sdist = sqrt(fma(bx, bx, by * by));
            if (Sdist <= R) ++xi;
          }
        }
        eta[icel] = (double)xi * inv_n2;
      } else if (Distance > R) {
        eta[icel] = 0.;
      } else {
        eta[icel] = 1.;
      }
    }
    free(sdx);
    free(sdy);
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < d.nx - 1; ++i)
    for (int j = 0; j < d.ny - 1; ++j)
      for (int k = 1; k < d.nz - 1; ++k) eta[IC(d, i, j, k)] = eta[IC(d, i, j, 0)];
}

/* ========================================================================
 * QUICK convection + central diffusion, explicit Euler
 * (inc/source/convection_and_difussion.hpp:20-84, 168-446)
 * ===================================================================== */

/* QuickUnitA (:20-49): face values in the component's own direction */
static inline void quickA(double Up, double Un, size_t pp, size_t p, size_t ic, size_t n, size_t nn,
                          const double *Di, const double *Ds, int I, const double *phi,
                          double *phiP, double *phiN) {
  if (Up > 0.0)
    *phiP = 0.5 * (phi[ic] + phi[p]) - 0.125 * Di[I + 1] * Di[I + 1] / Ds[I] *
                                           ((phi[p] - phi[ic]) / Di[I + 1] - (phi[ic] - phi[n]) / Di[I]);
  else
    *phiP = 0.5 * (phi[ic] + phi[p]) - 0.125 * Di[I + 1] * Di[I + 1] / Ds[I + 1] *
                                           ((phi[pp] - phi[p]) / Di[I + 2] - (phi[p] - phi[ic]) / Di[I + 1]);
  if (Un > 0.0)
    *phiN = 0.5 * (phi[n] + phi[ic]) - 0.125 * Di[I] * Di[I] / Ds[I - 1] *
                                           ((phi[ic] - phi[n]) / Di[I] - (phi[n] - phi[nn]) / Di[I - 1]);
  else
    *phiN = 0.5 * (phi[n] + phi[ic]) - 0.125 * Di[I] * Di[I] / Ds[I] *
                                           ((phi[p] - phi[ic]) / Di[I + 1] - (phi[ic] - phi[n]) / Di[I]);
}

/* QuickUnitB (:53-84): face values across the component's direction */
static inline void quickB(double Up, double Un, size_t pp, size_t p, size_t ic, size_t n, size_t nn,
                          const double *Di, const double *Ds, int I, const double *phi,
                          double *phiP, double *phiN) {
  if (Up > 0.0)
    *phiP = 0.5 * (phi[ic] + phi[p]) - 0.125 * Ds[I] * Ds[I] / Di[I] *
                                           ((phi[p] - phi[ic]) / Ds[I] - (phi[ic] - phi[n]) / Ds[I - 1]);
  else
    *phiP = 0.5 * (phi[ic] + phi[p]) - 0.125 * Ds[I] * Ds[I] / Di[I + 1] *
                                           ((phi[pp] - phi[p]) / Ds[I + 1] - (phi[p] - phi[ic]) / Ds[I]);
  if (Un > 0.0)
    *phiN = 0.5 * (phi[n] + phi[ic]) - 0.125 * Ds[I - 1] * Ds[I - 1] / Di[I - 1] *
                                           ((phi[ic] - phi[n]) / Ds[I - 1] - (phi[n] - phi[nn]) / Ds[I - 2]);
  else
    *phiN = 0.5 * (phi[n] + phi[ic]) - 0.125 * Ds[I - 1] * Ds[I - 1] / Di[I] *
                                           ((phi[p] - phi[ic]) / Ds[I] - (phi[ic] - phi[n]) / Ds[I - 1]);
}

/* LudUnitA / LudUnitB (:90-144): one body serves both */
static inline void lud(double Up, double Un, size_t pp, size_t p, size_t ic, size_t n, size_t nn,
                       const double *Di, const double *Ds, int I, const double *phi, double *phiP, double *phiN) {
  if (Up > 0.0) *phiP = phi[ic] + 0.5 * Ds[I] * (phi[ic] - phi[n]) / Di[I];
  else *phiP = phi[p] + 0.5 * Ds[I + 1] * (phi[p] - phi[pp]) / Di[I + 1];
  if (Un > 0.0) *phiN = phi[n] + 0.5 * Ds[I - 1] * (phi[n] - phi[nn]) / Di[I - 1];
  else *phiN = phi[ic] + 0.5 * Ds[I] * (phi[ic] - phi[p]) / Di[I + 1];
}

/* UdUnit (:148-165) */
static inline void ud(double Up, double Un, size_t p, size_t ic, size_t n, const double *phi, double *phiP,
                      double *phiN) {
  *phiP = Up > 0.0 ? phi[ic] : phi[p];
  *phiN = Un > 0.0 ? phi[n] : phi[ic];
}

/* the face-value unit of a scheme; own: the component's own direction */
static inline void face(int scheme, int own, double Up, double Un, size_t pp, size_t p, size_t ic, size_t n,
                        size_t nn, const double *Di, const double *Ds, int I, const double *phi, double *phiP,
                        double *phiN) {
  if (scheme == CFD_CONV_LUD) lud(Up, Un, pp, p, ic, n, nn, Di, Ds, I, phi, phiP, phiN);
  else if (scheme == CFD_CONV_UD) ud(Up, Un, p, ic, n, phi, phiP, phiN);
  else if (own) quickA(Up, Un, pp, p, ic, n, nn, Di, Ds, I, phi, phiP, phiN);
  else quickB(Up, Un, pp, p, ic, n, nn, Di, Ds, I, phi, phiP, phiN);
}

void orc_predictor(const cfd_grid_desc *g, double nu, double dt,
                   const double *u, const double *v, const double *w,
                   double *u1, double *v1, double *w1) {
  orc_predictor_ex(g, CFD_CONV_QUICK, nu, dt, u, v, w, u1, v1, w1);
}

void orc_predictor_ex(const cfd_grid_desc *g, int scheme, double nu, double dt,
                      const double *u, const double *v, const double *w,
                      double *u1, double *v1, double *w1) {
  orc_predictor_les(g, scheme, nu, dt, NULL, u, v, w, u1, v1, w1);
}

/* SmagorinskyModel + GradientGauss (inc/source/smagorinskyModel.hpp:6-79):
   gradients phi[-] - phi[+] over (Dxs,Dy,Dz) for u, (Dx,Dys,Dz) for v,
   (Dx,Dy,Dzs) for w; |2 ux^2 + vy^2 + wz^2 + (uy+vx)^2 + (uz+wx)^2 +
   (vz+wy)^2|; nu_t = (0.18 (Dx Dy Dz)^(1/3))^2 sqrt(.) at icelCal */
void orc_smagorinsky(const cfd_grid_desc *g, const double *u, const double *v, const double *w, double *vis) {
  const dims_t d = dims_of(g);
  const size_t sx = d.nzny, sy = (size_t)d.nz;
  const double *Dx = g->Dx, *Dy = g->Dy, *Dz = g->Dz, *Dxs = g->Dxs, *Dys = g->Dys, *Dzs = g->Dzs;
#pragma omp parallel for schedule(static)
  for (int i = GC; i < d.nx - GC; ++i)
    for (int j = GC; j < d.ny - GC; ++j)
      for (int k = GC; k < d.nz - GC; ++k) {
        const size_t ic = IC(d, i, j, k);
        const double ux = 0.5 / Dxs[i] * (u[ic - sx] - u[ic + sx]), uy = 0.5 / Dy[j] * (u[ic - sy] - u[ic + sy]),
                     uz = 0.5 / Dz[k] * (u[ic - 1] - u[ic + 1]);
        const double vx = 0.5 / Dx[i] * (v[ic - sx] - v[ic + sx]), vy = 0.5 / Dys[j] * (v[ic - sy] - v[ic + sy]),
                     vz = 0.5 / Dz[k] * (v[ic - 1] - v[ic + 1]);
        const double wx = 0.5 / Dx[i] * (w[ic - sx] - w[ic + sx]), wy = 0.5 / Dy[j] * (w[ic - sy] - w[ic + sy]),
                     wz = 0.5 / Dzs[k] * (w[ic - 1] - w[ic + 1]);
        const double sq = fabs(2.0 * (ux * ux) + vy * vy + wz * wz + (uy + vx) * (uy + vx) + (uz + wx) * (uz + wx) +
                               (vz + wy) * (vz + wy));
        const double cd = 0.18 * pow(Dx[i] * Dy[j] * Dz[k], 1.0 / 3.0);
        vis[ICAL(d, i, j, k)] = (cd * cd) * sqrt(sq);
      }
}

static void predictor_impl(const cfd_grid_desc *g, int scheme, double nu, double dt, const double *vis, int pz,
                           const double *u, const double *v, const double *w,
                           double *u1, double *v1, double *w1);

void orc_predictor_les(const cfd_grid_desc *g, int scheme, double nu, double dt, const double *vis,
                       const double *u, const double *v, const double *w,
                       double *u1, double *v1, double *w1) {
  predictor_impl(g, scheme, nu, dt, vis, 0, u, v, w, u1, v1, w1);
}

/* the predictor with the boundary mode of P (periodic z: the last z face of
   w is a real face and is predicted too) */
void orc_predictor_bl(const cfd_grid_desc *g, const cfd_params *P, const double *vis,
                      const double *u, const double *v, const double *w,
                      double *u1, double *v1, double *w1) {
  predictor_impl(g, P->convection, 1.0 / P->Re, P->dt, vis, P->periodic_z, u, v, w, u1, v1, w1);
}

static void predictor_impl(const cfd_grid_desc *g, int scheme, double nu, double dt, const double *vis, int pz,
                           const double *u, const double *v, const double *w,
                           double *u1, double *v1, double *w1) {
  const dims_t d = dims_of(g);
  const size_t sx = d.nzny, sy = (size_t)d.nz, sz = 1;
  const double *Dx = g->Dx, *Dy = g->Dy, *Dz = g->Dz, *Dxs = g->Dxs, *Dys = g->Dys, *Dzs = g->Dzs;
  const int ib = GC, ie = d.nx - GC, jb = GC, je = d.ny - GC, kb = GC, ke = d.nz - GC;

  /* u: i in [2, nx-3) -- the last face is excluded (:193-197) */
#pragma omp parallel for schedule(static)
  for (int i = ib; i < ie - 1; ++i)
    for (int j = jb; j < je; ++j)
      for (int k = kb; k < ke; ++k) {
        const size_t ic = IC(d, i, j, k);
        const size_t xm = ic - sx, xp = ic + sx, ym = ic - sy, yp = ic + sy, zm = ic - sz, zp = ic + sz;
        const size_t xmm = ic - 2 * sx, xpp = ic + 2 * sx, ymm = ic - 2 * sy, ypp = ic + 2 * sy;
        const size_t zmm = ic - 2 * sz, zpp = ic + 2 * sz;
        const double IT = (Dxs[i] * 0.5) / Dx[i]; /* x_INTRPL (structureGrid.hpp:442) */
        const double Up = 0.5 * (u[xp] + u[ic]);
        const double Un = 0.5 * (u[xm] + u[ic]);
        const double Vp = v[ic] + (v[xp] - v[ic]) * IT;
        const double Vn = v[ym] + (v[ym + xp - ic] - v[ym]) * IT;
        const double Wp = w[ic] + (w[xp] - w[ic]) * IT;
        const double Wn = w[zm] + (w[zm + xp - ic] - w[zm]) * IT;
        double uPx, uNx, uPy, uNy, uPz, uNz;
        face(scheme, 1, Up, Un, xpp, xp, ic, xm, xmm, Dx, Dxs, i, u, &uPx, &uNx);
        face(scheme, 0, Vp, Vn, ypp, yp, ic, ym, ymm, Dy, Dys, j, u, &uPy, &uNy);
        face(scheme, 0, Wp, Wn, zpp, zp, ic, zm, zmm, Dz, Dzs, k, u, &uPz, &uNz);
        const double convection = -dt * ((uPx * uPx - uNx * uNx) / Dxs[i] + (uPy * Vp - uNy * Vn) / Dy[j] +
                                         (uPz * Wp - uNz * Wn) / Dz[k]);
        const double difussion =
            (vis ? nu + vis[ICAL(d, i, j, k)] : nu) * dt *
            (((u[xp] - u[ic]) / Dx[i + 1] - (u[ic] - u[xm]) / Dx[i]) / Dxs[i] +
             ((u[yp] - u[ic]) / Dys[j] - (u[ic] - u[ym]) / Dys[j - 1]) / Dy[j] +
             ((u[zp] - u[ic]) / Dzs[k] - (u[ic] - u[zm]) / Dzs[k - 1]) / Dz[k]);
        u1[ic] = u[ic] + convection + difussion;
      }

  /* v: j in [2, ny-3) (:274-351) */
#pragma omp parallel for schedule(static)
  for (int i = ib; i < ie; ++i)
    for (int j = jb; j < je - 1; ++j)
      for (int k = kb; k < ke; ++k) {
        const size_t ic = IC(d, i, j, k);
        const size_t xm = ic - sx, xp = ic + sx, ym = ic - sy, yp = ic + sy, zm = ic - sz, zp = ic + sz;
        const size_t xmm = ic - 2 * sx, xpp = ic + 2 * sx, ymm = ic - 2 * sy, ypp = ic + 2 * sy;
        const size_t zmm = ic - 2 * sz, zpp = ic + 2 * sz;
        const double IT = (Dys[j] * 0.5) / Dy[j];
        const double Vp = 0.5 * (v[yp] + v[ic]);
        const double Vn = 0.5 * (v[ym] + v[ic]);
        const double Up = u[ic] + (u[yp] - u[ic]) * IT;
        const double Un = u[xm] + (u[xm + yp - ic] - u[xm]) * IT;
        const double Wp = w[ic] + (w[yp] - w[ic]) * IT;
        const double Wn = w[zm] + (w[zm + yp - ic] - w[zm]) * IT;
        double vPx, vNx, vPy, vNy, vPz, vNz;
        face(scheme, 0, Up, Un, xpp, xp, ic, xm, xmm, Dx, Dxs, i, v, &vPx, &vNx);
        face(scheme, 1, Vp, Vn, ypp, yp, ic, ym, ymm, Dy, Dys, j, v, &vPy, &vNy);
        face(scheme, 0, Wp, Wn, zpp, zp, ic, zm, zmm, Dz, Dzs, k, v, &vPz, &vNz);
        const double convection = -dt * ((Up * vPx - Un * vNx) / Dx[i] + (vPy * vPy - vNy * vNy) / Dys[j] +
                                         (Wp * vPz - Wn * vNz) / Dz[k]);
        const double difussion =
            (vis ? nu + vis[ICAL(d, i, j, k)] : nu) * dt *
            (((v[xp] - v[ic]) / Dxs[i] - (v[ic] - v[xm]) / Dxs[i - 1]) / Dx[i] +
             ((v[yp] - v[ic]) / Dy[j + 1] - (v[ic] - v[ym]) / Dy[j]) / Dys[j] +
             ((v[zp] - v[ic]) / Dzs[k] - (v[ic] - v[zm]) / Dzs[k - 1]) / Dz[k]);
        v1[ic] = v[ic] + convection + difussion;
      }

  /* w: k in [2, nz-3) (:357-434); periodic z: [2, nz-2) */
#pragma omp parallel for schedule(static)
  for (int i = ib; i < ie; ++i)
    for (int j = jb; j < je; ++j)
      for (int k = kb; k < ke - 1 + (pz ? 1 : 0); ++k) {
        const size_t ic = IC(d, i, j, k);
        const size_t xm = ic - sx, xp = ic + sx, ym = ic - sy, yp = ic + sy, zm = ic - sz, zp = ic + sz;
        const size_t xmm = ic - 2 * sx, xpp = ic + 2 * sx, ymm = ic - 2 * sy, ypp = ic + 2 * sy;
        const size_t zmm = ic - 2 * sz, zpp = ic + 2 * sz;
        const double IT = (Dzs[k] * 0.5) / Dz[k];
        const double Wp = 0.5 * (w[zp] + w[ic]);
        const double Wn = 0.5 * (w[zm] + w[ic]);
        const double Up = u[ic] + (u[zp] - u[ic]) * IT;
        const double Un = u[xm] + (u[xm + zp - ic] - u[xm]) * IT;
        const double Vp = v[ic] + (v[zp] - v[ic]) * IT;
        const double Vn = v[ym] + (v[ym + zp - ic] - v[ym]) * IT;
        double wPx, wNx, wPy, wNy, wPz, wNz;
        face(scheme, 0, Up, Un, xpp, xp, ic, xm, xmm, Dx, Dxs, i, w, &wPx, &wNx);
        face(scheme, 0, Vp, Vn, ypp, yp, ic, ym, ymm, Dy, Dys, j, w, &wPy, &wNy);
        face(scheme, 1, Wp, Wn, zpp, zp, ic, zm, zmm, Dz, Dzs, k, w, &wPz, &wNz);
        const double convection = -dt * ((Up * wPx - Un * wNx) / Dx[i] + (Vp * wPy - Vn * wNy) / Dy[j] +
                                         (wPz * wPz - wNz * wNz) / Dzs[k]);
        const double difussion =
            (vis ? nu + vis[ICAL(d, i, j, k)] : nu) * dt *
            (((w[xp] - w[ic]) / Dxs[i] - (w[ic] - w[xm]) / Dxs[i - 1]) / Dx[i] +
             ((w[yp] - w[ic]) / Dys[j] - (w[ic] - w[ym]) / Dys[j - 1]) / Dy[j] +
             ((w[zp] - w[ic]) / Dz[k + 1] - (w[ic] - w[zm]) / Dz[k]) / Dzs[k]);
        w1[ic] = w[ic] + convection + difussion;
      }
}

/* T1 += source_termF (RUN_CPU_PC_METHOD.hpp:309-316, 460-467).  The reference
   indexes the ghosted T1 with the interior index icelCal (quirk = 1); quirk = 0
   adds at the cell's own ghosted index. */
void orc_source_add(const cfd_grid_desc *g, int quirk, const double *src,
                    double *u1, double *v1, double *w1) {
  const dims_t d = dims_of(g);
  double *T[3] = {u1, v1, w1};
  for (int c = 0; c < 3; ++c) {
    double *t = T[c];
    const double *s = src + d.N * (size_t)c;
#pragma omp parallel for schedule(static)
    for (int i = GC; i < d.nx - GC; ++i)
      for (int j = GC; j < d.ny - GC; ++j)
        for (int k = GC; k < d.nz - GC; ++k) {
          const size_t r = ICAL(d, i, j, k);
          t[quirk ? r : IC(d, i, j, k)] += s[r];
        }
  }
}

/* source_termF += dt * f (RUN_CPU_PC_METHOD.hpp:393-401) */
void orc_source_accumulate(const cfd_grid_desc *g, double dt, const double *f, double *src) {
  const dims_t d = dims_of(g);
  const size_t n3 = 3 * d.N;
#pragma omp parallel for schedule(static)
  for (long long m = 0; m < (long long)n3; ++m) src[m] += dt * f[m];
}

/* createBMatrix_omp, JACOBI_PC branch (inc/matrix/b_matrix.hpp:51-80) */
void orc_rhs(const cfd_grid_desc *g, double dt, const double *jp,
             const double *u1, const double *v1, const double *w1, double *b) {
  const dims_t d = dims_of(g);
  const double ddt = 1.0 / dt;
#pragma omp parallel for schedule(static)
  for (int i = GC; i < d.nx - GC; ++i)
    for (int j = GC; j < d.ny - GC; ++j)
      for (int k = GC; k < d.nz - GC; ++k) {
        const size_t icel = IC(d, i, j, k);
        const size_t row = ICAL(d, i, j, k);
        b[row] = -jp[row] * ddt *
                 ((u1[icel] - u1[IC(d, i - 1, j, k)]) / g->Dx[i] + (v1[icel] - v1[IC(d, i, j - 1, k)]) / g->Dy[j] +
                  (w1[icel] - w1[IC(d, i, j, k - 1)]) / g->Dz[k]);
      }
}

/* ========================================================================
 * Pressure operator
 * ===================================================================== */

/* createPressureMatrix (inc/matrix/PressureMatrix.hpp:5-117) into the ELL
   layout of mat::ELL_matrix<double>::set (ELL_sparseMatrix.hpp:146-160,
   246-320): slot 0 diagonal, then xm, xp, ym, yp, zm, zp without absent
   neighbours, padding (N-1, 0.0).  Pressure faces marked Dirichlet keep the
   neighbour coefficient in the diagonal (p = 0 in the ghost). */
int orc_ell_assemble(const cfd_grid_desc *g, const int pd[6], int32_t *col, double *val, double *jp) {
  return orc_ell_assemble_bl(g, pd, 0, col, val, jp);
}

/* periodic z (pz): no closure term on the z faces; the neighbour across the
   seam is set after the regular entries, where the reference's periodic
   branch sets it (PressureMatrix.hpp:101-111), with the wrapped cell's own
   index */
int orc_ell_assemble_bl(const cfd_grid_desc *g, const int pd[6], int pz, int32_t *col, double *val, double *jp) {
  const dims_t d = dims_of(g);
  const size_t N = d.N;
  if (N == 0 || N > 0x7fffffffULL) return -1;
  const int xBeg = GC, yBeg = GC, zBeg = GC, xEnd = d.nx - GC - 1, yEnd = d.ny - GC - 1, zEnd = d.nz - GC - 1;
  const double *Dx = g->Dx, *Dy = g->Dy, *Dz = g->Dz, *Dxs = g->Dxs, *Dys = g->Dys, *Dzs = g->Dzs;
#pragma omp parallel for schedule(static)
  for (int i = xBeg; i <= xEnd; ++i)
    for (int j = yBeg; j <= yEnd; ++j)
      for (int k = zBeg; k <= zEnd; ++k) {
        const double xp = 1.0 / Dx[i] / Dxs[i], yp = 1.0 / Dy[j] / Dys[j], zp = 1.0 / Dz[k] / Dzs[k];
        const double xm = 1.0 / Dx[i] / Dxs[i - 1], ym = 1.0 / Dy[j] / Dys[j - 1], zm = 1.0 / Dz[k] / Dzs[k - 1];
        double aa = 0.0, bb = 0.0, cc = 0.0;
        if (i == xBeg && !pd[0]) aa = -xm;
        if (i == xEnd && !pd[1]) aa = -xp;
        if (j == yBeg && !pd[2]) bb = -ym;
        if (j == yEnd && !pd[3]) bb = -yp;
        if (k == zBeg && !pd[4] && !pz) cc = -zm;
        if (k == zEnd && !pd[5] && !pz) cc = -zp;
        const size_t row = ICAL(d, i, j, k);
        const double vdiag = xm + xp + ym + yp + zm + zp + aa + bb + cc;
        const double jr = 1. / vdiag;
        jp[row] = jr;
        int s = 0;
        col[(size_t)s * N + row] = (int32_t)row; val[(size_t)s * N + row] = vdiag * jr; ++s;
        if (i != xBeg) { col[(size_t)s * N + row] = (int32_t)ICAL(d, i - 1, j, k); val[(size_t)s * N + row] = -xm * jr; ++s; }
        if (i != xEnd) { col[(size_t)s * N + row] = (int32_t)ICAL(d, i + 1, j, k); val[(size_t)s * N + row] = -xp * jr; ++s; }
        if (j != yBeg) { col[(size_t)s * N + row] = (int32_t)ICAL(d, i, j - 1, k); val[(size_t)s * N + row] = -ym * jr; ++s; }
        if (j != yEnd) { col[(size_t)s * N + row] = (int32_t)ICAL(d, i, j + 1, k); val[(size_t)s * N + row] = -yp * jr; ++s; }
        if (k != zBeg) { col[(size_t)s * N + row] = (int32_t)ICAL(d, i, j, k - 1); val[(size_t)s * N + row] = -zm * jr; ++s; }
        if (k != zEnd) { col[(size_t)s * N + row] = (int32_t)ICAL(d, i, j, k + 1); val[(size_t)s * N + row] = -zp * jr; ++s; }
        if (pz && k == zBeg) { col[(size_t)s * N + row] = (int32_t)ICAL(d, i, j, zEnd); val[(size_t)s * N + row] = -zm * jr; ++s; }
        if (pz && k == zEnd) { col[(size_t)s * N + row] = (int32_t)ICAL(d, i, j, zBeg); val[(size_t)s * N + row] = -zp * jr; ++s; }
        for (; s < 7; ++s) { col[(size_t)s * N + row] = (int32_t)(N - 1); val[(size_t)s * N + row] = 0.0; }
      }
  return 0;
}

/* ELL_matrix::multiply_omp (ELL_sparseMatrix.hpp:477-492).  The reference
   build contracts r[i] += VAL * x[IDX] into an FMA chain in slot order. */
void orc_spmv_ell(size_t N, int K, const int32_t *col, const double *val, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (long long r = 0; r < (long long)N; ++r) {
    double acc = 0.0;
    for (int s = 0; s < K; ++s) acc = fma(val[(size_t)s * N + r], x[col[(size_t)s * N + r]], acc);
    y[r] = acc;
  }
}

/* SPE_matrix1::setupPressure + set_Jp_pc (SPE_sparseMatrix1.hpp:40-121) */
/* grid::createPressureMatrixCSR (inc/grid/structureGrid.hpp:630-728): the
   Jacobi-scaled rows with their entries in column order xm, ym, zm, diag,
   zp, yp, xp (absent neighbours skipped), the diagonal summed in the order
   xm + ym + zm + xp + yp + zp + aa + bb + cc (not the ELL order), and jp =
   1/diag.  pd marks Dirichlet pressure faces, which keep their coefficient
   in the diagonal as in the ELL assembly (the reference's CSR path has none).
   ptr has N+1 entries; returns the number of entries. */
long long orc_csr_assemble(const cfd_grid_desc *g, int32_t *ptr, int32_t *col, double *val, double *jp) {
  const dims_t d = dims_of(g);
  const long long nzny = (long long)d.nyc * d.nzc;
  const double *Dx = g->Dx, *Dy = g->Dy, *Dz = g->Dz, *Dxs = g->Dxs, *Dys = g->Dys, *Dzs = g->Dzs;
  long long e = 0, row = 0;
  ptr[0] = 0;
  for (int i = 0; i < d.nxc; ++i)
    for (int j = 0; j < d.nyc; ++j)
      for (int k = 0; k < d.nzc; ++k, ++row) {
        const int I = i + GC, J = j + GC, K = k + GC;
        const double xp = 1.0 / Dx[I] / Dxs[I], yp = 1.0 / Dy[J] / Dys[J], zp = 1.0 / Dz[K] / Dzs[K];
        const double xm = 1.0 / Dx[I] / Dxs[I - 1], ym = 1.0 / Dy[J] / Dys[J - 1], zm = 1.0 / Dz[K] / Dzs[K - 1];
        /* the Neumann closure on every face (:658-666 checks no pressure BC) */
        double aa = 0.0, bb = 0.0, cc = 0.0;
        if (i == 0) aa = -xm;
        if (i == d.nxc - 1) aa = -xp;
        if (j == 0) bb = -ym;
        if (j == d.nyc - 1) bb = -yp;
        if (k == 0) cc = -zm;
        if (k == d.nzc - 1) cc = -zp;
        const double ijk = xm + ym + zm + xp + yp + zp + aa + bb + cc;
        const double J_ = 1.0 / ijk;
        jp[row] = J_;
        if (i != 0) { col[e] = (int32_t)(row - nzny); val[e++] = -xm * J_; }
        if (j != 0) { col[e] = (int32_t)(row - d.nzc); val[e++] = -ym * J_; }
        if (k != 0) { col[e] = (int32_t)(row - 1); val[e++] = -zm * J_; }
        col[e] = (int32_t)row; val[e++] = ijk * J_;
        if (k != d.nzc - 1) { col[e] = (int32_t)(row + 1); val[e++] = -zp * J_; }
        if (j != d.nyc - 1) { col[e] = (int32_t)(row + d.nzc); val[e++] = -yp * J_; }
        if (i != d.nxc - 1) { col[e] = (int32_t)(row + nzny); val[e++] = -xp * J_; }
        ptr[row + 1] = (int32_t)e;
      }
  return e;
}

/* CSR_matrix::multiply_omp (inc/matrix/CSR_sparseMatrix.hpp:278-292): r[i] +=
   VAL * x[IDX] in entry order, which the reference build contracts to fma */
void orc_spmv_csr(size_t N, const int32_t *ptr, const int32_t *col, const double *val, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (long long r = 0; r < (long long)N; ++r) {
    double acc = 0.0;
    for (int32_t e = ptr[r]; e < ptr[r + 1]; ++e) acc = fma(val[e], x[col[e]], acc);
    y[r] = acc;
  }
}

void orc_spe_setup(const cfd_grid_desc *g, const int pd[6], double *coef, double *jp) {
  orc_spe_setup_bl(g, pd, 0, coef, jp);
}

void orc_spe_setup_bl(const cfd_grid_desc *g, const int pd[6], int pz, double *coef, double *jp) {
  const dims_t d = dims_of(g);
  double *xm = coef, *xp = xm + d.nxc, *ym = xp + d.nxc, *yp = ym + d.nyc, *zm = yp + d.nyc, *zp = zm + d.nzc;
  for (int i = 0; i < d.nxc; ++i) {
    xp[i] = 1.0 / g->Dx[i + GC] / g->Dxs[i + GC];
    xm[i] = 1.0 / g->Dx[i + GC] / g->Dxs[i + GC - 1];
  }
  for (int j = 0; j < d.nyc; ++j) {
    yp[j] = 1.0 / g->Dy[j + GC] / g->Dys[j + GC];
    ym[j] = 1.0 / g->Dy[j + GC] / g->Dys[j + GC - 1];
  }
  for (int k = 0; k < d.nzc; ++k) {
    zp[k] = 1.0 / g->Dz[k + GC] / g->Dzs[k + GC];
    zm[k] = 1.0 / g->Dz[k + GC] / g->Dzs[k + GC - 1];
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < d.nxc; ++i)
    for (int j = 0; j < d.nyc; ++j)
      for (int k = 0; k < d.nzc; ++k) {
        const size_t ic = (size_t)k + (size_t)j * d.nzc + (size_t)i * d.nyc * d.nzc;
        double ijk = xm[i] + ym[j] + zm[k] + xp[i] + yp[j] + zp[k];
        if (i == 0 && !pd[0]) ijk -= xm[i];
        if (i == d.nxc - 1 && !pd[1]) ijk -= xp[i];
        if (j == 0 && !pd[2]) ijk -= ym[j];
        if (j == d.nyc - 1 && !pd[3]) ijk -= yp[j];
        if (k == 0 && !pd[4] && !pz) ijk -= zm[k];
        if (k == d.nzc - 1 && !pd[5] && !pz) ijk -= zp[k];
        jp[ic] = 1.0 / ijk;
      }
}

/* SPE_matrix1::spmv_ompJP (SPE_sparseMatrix1.hpp:126-144); each
   sum += (-c*jp) * x contracts to an FMA in the reference build */
void orc_spmv_spe(const cfd_grid_desc *g, const double *coef, const double *jp, const double *x, double *y) {
  orc_spmv_spe_bl(g, 0, coef, jp, x, y);
}

/* periodic z (pz): the z neighbours of the first and last planes wrap */
void orc_spmv_spe_bl(const cfd_grid_desc *g, int pz, const double *coef, const double *jp, const double *x,
                     double *y) {
  const dims_t d = dims_of(g);
  const double *xm = coef, *xp = xm + d.nxc, *ym = xp + d.nxc, *yp = ym + d.nyc, *zm = yp + d.nyc, *zp = zm + d.nzc;
  const size_t nz_ = (size_t)d.nzc, nynz_ = (size_t)d.nyc * d.nzc;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < d.nxc; ++i)
    for (int j = 0; j < d.nyc; ++j)
      for (int k = 0; k < d.nzc; ++k) {
        const size_t ic = (size_t)k + (size_t)j * nz_ + (size_t)i * nynz_;
        const double J = jp[ic];
        double sum = 0.0;
        if (i != 0) sum = fma(-xm[i] * J, x[ic - nynz_], sum);
        if (j != 0) sum = fma(-ym[j] * J, x[ic - nz_], sum);
        if (k != 0) sum = fma(-zm[k] * J, x[ic - 1], sum);
        else if (pz) sum = fma(-zm[k] * J, x[ic + (nz_ - 1)], sum);
        sum += x[ic];
        if (k != d.nzc - 1) sum = fma(-zp[k] * J, x[ic + 1], sum);
        else if (pz) sum = fma(-zp[k] * J, x[ic - (nz_ - 1)], sum);
        if (j != d.nyc - 1) sum = fma(-yp[j] * J, x[ic + nz_], sum);
        if (i != d.nxc - 1) sum = fma(-xp[i] * J, x[ic + nynz_], sum);
        y[ic] = sum;
      }
}

/* Pressure_transform_X_result (inc/pressure/Transform.hpp:8-24) */
void orc_pressure_scatter(const cfd_grid_desc *g, const double *X, double *p) {
  const dims_t d = dims_of(g);
#pragma omp parallel for schedule(static)
  for (int i = GC; i < d.nx - GC; ++i)
    for (int j = GC; j < d.ny - GC; ++j)
      for (int k = GC; k < d.nz - GC; ++k) p[IC(d, i, j, k)] = X[ICAL(d, i, j, k)];
}

/* update_UandF_omp (inc/dfib/updateUandF.hpp:95-192), solid velocity 0.
   gain scales the forcing, u_new = (1 - gain*eta_f) T2 (gain = 1 is the
   reference, bit for bit); t2, when not NULL, receives the intermediate
   velocity T2 = u* - dt dp/dx (3*N, interior order) for the adjoint. */
static void dfib_update_impl(const cfd_grid_desc *g, double dt, double gain, int conv, int pz, const double *eta,
                             const double *p, const double *u1, const double *v1, const double *w1,
                             double *u3, double *v3, double *w3, double *f, double *t2);

void orc_dfib_update_ex(const cfd_grid_desc *g, double dt, double gain, const double *eta, const double *p,
                        const double *u1, const double *v1, const double *w1,
                        double *u3, double *v3, double *w3, double *f, double *t2) {
  dfib_update_impl(g, dt, gain, 0, 0, eta, p, u1, v1, w1, u3, v3, w3, f, t2);
}

/* the update with the boundary mode of P: a convective outflow projects the
   outflow face of u too (with the ghost pressure, 0 there) and gives it no
   immersed-boundary forcing; periodic z projects the last z face of w with
   the ghost pressure of the plane across the seam */
void orc_dfib_update_bl(const cfd_grid_desc *g, const cfd_params *P, double gain, const double *eta, const double *p,
                        const double *u1, const double *v1, const double *w1,
                        double *u3, double *v3, double *w3, double *f, double *t2) {
  dfib_update_impl(g, P->dt, gain, P->outflow_convective, P->periodic_z, eta, p, u1, v1, w1, u3, v3, w3, f, t2);
}

static void dfib_update_impl(const cfd_grid_desc *g, double dt, double gain, int conv, int pz, const double *eta,
                             const double *p, const double *u1, const double *v1, const double *w1,
                             double *u3, double *v3, double *w3, double *f, double *t2) {
  const dims_t d = dims_of(g);
  const size_t N = d.N;
  const double uS = 0.0, vS = 0.0, wS = 0.0;
#pragma omp parallel for schedule(static)
  for (int i = GC; i < d.nx - GC - 1 + (conv ? 1 : 0); ++i)
    for (int j = GC; j < d.ny - GC; ++j)
      for (int k = GC; k < d.nz - GC; ++k) {
        const size_t ic = IC(d, i, j, k), icP = IC(d, i + 1, j, k);
        const double T2u = u1[ic] - dt * (p[icP] - p[ic]) / g->Dxs[i];
        const double e = (conv && i == d.nx - 3) ? 0.0
                                                 : eta[ic] + (eta[icP] - eta[ic]) * ((g->Dxs[i] * 0.5) / g->Dx[i]);
        u3[ic] = eta[ic] * uS + (1.0 - gain * e) * T2u;
        f[ICAL(d, i, j, k)] = (u3[ic] - T2u) / dt;
        if (t2) t2[ICAL(d, i, j, k)] = T2u;
      }
#pragma omp parallel for schedule(static)
  for (int i = GC; i < d.nx - GC; ++i)
    for (int j = GC; j < d.ny - GC - 1; ++j)
      for (int k = GC; k < d.nz - GC; ++k) {
        const size_t ic = IC(d, i, j, k), icP = IC(d, i, j + 1, k);
        const double T2v = v1[ic] - dt * (p[icP] - p[ic]) / g->Dys[j];
        const double e = eta[ic] + (eta[icP] - eta[ic]) * ((g->Dys[j] * 0.5) / g->Dy[j]);
        v3[ic] = eta[ic] * vS + (1.0 - gain * e) * T2v;
        f[ICAL(d, i, j, k) + N] = (v3[ic] - T2v) / dt;
        if (t2) t2[ICAL(d, i, j, k) + N] = T2v;
      }
#pragma omp parallel for schedule(static)
  for (int i = GC; i < d.nx - GC; ++i)
    for (int j = GC; j < d.ny - GC; ++j)
      for (int k = GC; k < d.nz - GC - 1 + (pz ? 1 : 0); ++k) {
        const size_t ic = IC(d, i, j, k), icP = IC(d, i, j, k + 1);
        const double T2w = w1[ic] - dt * (p[icP] - p[ic]) / g->Dzs[k];
        const double e = eta[ic] + (eta[icP] - eta[ic]) * ((g->Dzs[k] * 0.5) / g->Dz[k]);
        w3[ic] = eta[ic] * wS + (1.0 - gain * e) * T2w;
        f[ICAL(d, i, j, k) + 2 * N] = (w3[ic] - T2w) / dt;
        if (t2) t2[ICAL(d, i, j, k) + 2 * N] = T2w;
      }
}

void orc_dfib_update(const cfd_grid_desc *g, double dt, const double *eta, const double *p,
                     const double *u1, const double *v1, const double *w1,
                     double *u3, double *v3, double *w3, double *f) {
  orc_dfib_update_ex(g, dt, 1.0, eta, p, u1, v1, w1, u3, v3, w3, f, NULL);
}

/* ========================================================================
 * Boundary conditions
 * ===================================================================== */

/* BC_staggered_main (Boundary_Condition.hpp:37-175), faces in the order the
   reference writes them; normal components: ghost = Dir + Num*interior,
   tangential: ghost = 2 Dir + (2 Num - 1) interior; faces 1/3/5 also set the
   last interior face.  Pressure ghosts copy the first interior cell, or are 0
   on faces marked Dirichlet. */
/* uin (NULL: the table value) replaces Dir[0][0] by a profile over y,
   uin[j - 2] for interior row j: the inflow the adjoint differentiates. */
static void bc_main_impl(const cfd_grid_desc *g, const double Dir[6][3], const double Num[6][3], const int pd[6],
                         const double *uin, int conv, double conv_udt, int pz, double *u, double *v, double *w,
                         double *p);

void orc_bc_main_ex(const cfd_grid_desc *g, const double Dir[6][3], const double Num[6][3],
                    const int pd[6], const double *uin, double *u, double *v, double *w, double *p) {
  bc_main_impl(g, Dir, Num, pd, uin, 0, 0.0, 0, u, v, w, p);
}

/* BC_staggered_main with the boundary mode of P (no reference counterpart
   for these two): a convective outflow on face 1 -- one implicit upwind step
   phi_b = (phi_b + s phi_in) / (1 + s), s = U dt / h, for the outflow face
   of u (h = Dx of the last cell, phi_in the last interior face) and the
   ghosts of v and w (h = Dxs of the last cell, phi_in the last cell); both
   ghost layers take the new value -- and periodic z: the two ghost layers on
   each z side are the images one period away, over whole planes, after the
   x and y faces, for u, v, w and p */
void orc_bc_main_bl(const cfd_grid_desc *g, const cfd_params *P, const double *uin, double *u, double *v, double *w,
                    double *p) {
  bc_main_impl(g, P->bc_dir, P->bc_num, P->p_dirichlet, uin, P->outflow_convective, P->outflow_speed * P->dt,
               P->periodic_z, u, v, w, p);
}

/* the z ghost layers of a field from their periodic images, whole planes */
static void periodic_z_fill(const dims_t d, double *a) {
  const int nz = d.nz;
  for (int i = 0; i < d.nx; ++i)
    for (int j = 0; j < d.ny; ++j) {
      const size_t b = IC(d, i, j, 0);
      a[b + 0] = a[b + nz - 4];
      a[b + 1] = a[b + nz - 3];
      a[b + nz - 2] = a[b + 2];
      a[b + nz - 1] = a[b + 3];
    }
}

void orc_periodic_refresh(const cfd_grid_desc *g, double *u, double *v, double *w) {
  const dims_t d = dims_of(g);
  periodic_z_fill(d, u);
  periodic_z_fill(d, v);
  periodic_z_fill(d, w);
}

void orc_periodic_refresh_p(const cfd_grid_desc *g, double *p) { periodic_z_fill(dims_of(g), p); }

static void bc_main_impl(const cfd_grid_desc *g, const double Dir[6][3], const double Num[6][3], const int pd[6],
                         const double *uin, int conv, double conv_udt, int pz, double *u, double *v, double *w,
                         double *p) {
  const dims_t d = dims_of(g);
  const int nx = d.nx, ny = d.ny, nz = d.nz;
#define NRM(b, c, x) (Dir[b][c] + Num[b][c] * (x))
#define TAN(b, c, x) (Dir[b][c] * 2. + (2.0 * Num[b][c] - 1.0) * (x))
  for (int j = GC; j < ny - GC; ++j)
    for (int k = GC; k < nz - GC; ++k) { /* BC_0 */
      const double din = uin ? uin[j - GC] : Dir[0][0];
      u[IC(d, 0, j, k)] = u[IC(d, 1, j, k)] = din + Num[0][0] * u[IC(d, 2, j, k)];
      v[IC(d, 0, j, k)] = v[IC(d, 1, j, k)] = TAN(0, 1, v[IC(d, 2, j, k)]);
      w[IC(d, 0, j, k)] = w[IC(d, 1, j, k)] = TAN(0, 2, w[IC(d, 2, j, k)]);
    }
  for (int j = GC; j < ny - GC; ++j)
    for (int k = GC; k < nz - GC; ++k) { /* BC_1 */
      if (conv) {
        const double su = conv_udt / g->Dx[nx - 3], st = conv_udt / g->Dxs[nx - 3];
        const double xu = (u[IC(d, nx - 3, j, k)] + su * u[IC(d, nx - 4, j, k)]) / (1.0 + su);
        u[IC(d, nx - 1, j, k)] = u[IC(d, nx - 2, j, k)] = u[IC(d, nx - 3, j, k)] = xu;
        const double xv = (v[IC(d, nx - 2, j, k)] + st * v[IC(d, nx - 3, j, k)]) / (1.0 + st);
        v[IC(d, nx - 1, j, k)] = v[IC(d, nx - 2, j, k)] = xv;
        const double xw = (w[IC(d, nx - 2, j, k)] + st * w[IC(d, nx - 3, j, k)]) / (1.0 + st);
        w[IC(d, nx - 1, j, k)] = w[IC(d, nx - 2, j, k)] = xw;
      } else {
        u[IC(d, nx - 1, j, k)] = u[IC(d, nx - 2, j, k)] = u[IC(d, nx - 3, j, k)] = NRM(1, 0, u[IC(d, nx - 4, j, k)]);
        v[IC(d, nx - 1, j, k)] = v[IC(d, nx - 2, j, k)] = TAN(1, 1, v[IC(d, nx - 3, j, k)]);
        w[IC(d, nx - 1, j, k)] = w[IC(d, nx - 2, j, k)] = TAN(1, 2, w[IC(d, nx - 3, j, k)]);
      }
    }
  for (int i = GC; i < nx - GC; ++i)
    for (int k = GC; k < nz - GC; ++k) { /* BC_2 */
      u[IC(d, i, 0, k)] = u[IC(d, i, 1, k)] = TAN(2, 0, u[IC(d, i, 2, k)]);
      v[IC(d, i, 0, k)] = v[IC(d, i, 1, k)] = NRM(2, 1, v[IC(d, i, 2, k)]);
      w[IC(d, i, 0, k)] = w[IC(d, i, 1, k)] = TAN(2, 2, w[IC(d, i, 2, k)]);
    }
  for (int i = GC; i < nx - GC; ++i)
    for (int k = GC; k < nz - GC; ++k) { /* BC_3 */
      v[IC(d, i, ny - 1, k)] = v[IC(d, i, ny - 2, k)] = v[IC(d, i, ny - 3, k)] = NRM(3, 1, v[IC(d, i, ny - 4, k)]);
      u[IC(d, i, ny - 1, k)] = u[IC(d, i, ny - 2, k)] = TAN(3, 0, u[IC(d, i, ny - 3, k)]);
      w[IC(d, i, ny - 1, k)] = w[IC(d, i, ny - 2, k)] = TAN(3, 2, w[IC(d, i, ny - 3, k)]);
    }
  if (pz) {
    periodic_z_fill(d, u);
    periodic_z_fill(d, v);
    periodic_z_fill(d, w);
  } else {
    for (int i = GC; i < nx - GC; ++i)
      for (int j = GC; j < ny - GC; ++j) { /* BC_4 */
        u[IC(d, i, j, 0)] = u[IC(d, i, j, 1)] = TAN(4, 0, u[IC(d, i, j, 2)]);
        v[IC(d, i, j, 0)] = v[IC(d, i, j, 1)] = TAN(4, 1, v[IC(d, i, j, 2)]);
        w[IC(d, i, j, 0)] = w[IC(d, i, j, 1)] = NRM(4, 2, w[IC(d, i, j, 2)]);
      }
    for (int i = GC; i < nx - GC; ++i)
      for (int j = GC; j < ny - GC; ++j) { /* BC_5 */
        w[IC(d, i, j, nz - 1)] = w[IC(d, i, j, nz - 2)] = w[IC(d, i, j, nz - 3)] = NRM(5, 2, w[IC(d, i, j, nz - 4)]);
        u[IC(d, i, j, nz - 1)] = u[IC(d, i, j, nz - 2)] = TAN(5, 0, u[IC(d, i, j, nz - 3)]);
        v[IC(d, i, j, nz - 1)] = v[IC(d, i, j, nz - 2)] = TAN(5, 1, v[IC(d, i, j, nz - 3)]);
      }
  }
#undef NRM
#undef TAN
  if (!p) return;
  for (int j = GC; j < ny - GC; ++j)
    for (int k = GC; k < nz - GC; ++k) {
      p[IC(d, 0, j, k)] = p[IC(d, 1, j, k)] = pd[0] ? 0.0 : p[IC(d, 2, j, k)];
      p[IC(d, nx - 1, j, k)] = p[IC(d, nx - 2, j, k)] = pd[1] ? 0.0 : p[IC(d, nx - 3, j, k)];
    }
  for (int i = GC; i < nx - GC; ++i)
    for (int k = GC; k < nz - GC; ++k) {
      p[IC(d, i, 0, k)] = p[IC(d, i, 1, k)] = pd[2] ? 0.0 : p[IC(d, i, 2, k)];
      p[IC(d, i, ny - 1, k)] = p[IC(d, i, ny - 2, k)] = pd[3] ? 0.0 : p[IC(d, i, ny - 3, k)];
    }
  if (pz) {
    periodic_z_fill(d, p);
    return;
  }
  for (int i = GC; i < nx - GC; ++i)
    for (int j = GC; j < ny - GC; ++j) {
      p[IC(d, i, j, 0)] = p[IC(d, i, j, 1)] = pd[4] ? 0.0 : p[IC(d, i, j, 2)];
      p[IC(d, i, j, nz - 1)] = p[IC(d, i, j, nz - 2)] = pd[5] ? 0.0 : p[IC(d, i, j, nz - 3)];
    }
}

void orc_bc_main(const cfd_grid_desc *g, const double Dir[6][3], const double Num[6][3],
                 const int pd[6], double *u, double *v, double *w, double *p) {
  orc_bc_main_ex(g, Dir, Num, pd, NULL, u, v, w, p);
}

/* BC_staggered_copy (updateBC.hpp:6-109): boundary faces old -> current */
static void bc_copy_impl(const cfd_grid_desc *g, int pz, const double *u0, const double *v0, const double *w0,
                         double *u1, double *v1, double *w1);

void orc_bc_copy(const cfd_grid_desc *g, const double *u0, const double *v0, const double *w0,
                 double *u1, double *v1, double *w1) {
  bc_copy_impl(g, 0, u0, v0, w0, u1, v1, w1);
}

/* periodic z: the z ghost layers are copied over whole planes, and the last
   z face of w, a real face, is not */
void orc_bc_copy_bl(const cfd_grid_desc *g, const cfd_params *P, const double *u0, const double *v0,
                    const double *w0, double *u1, double *v1, double *w1) {
  bc_copy_impl(g, P->periodic_z, u0, v0, w0, u1, v1, w1);
}

static void bc_copy_impl(const cfd_grid_desc *g, int pz, const double *u0, const double *v0, const double *w0,
                         double *u1, double *v1, double *w1) {
  const dims_t d = dims_of(g);
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  for (int j = GC; j < ny - GC; ++j)
    for (int k = GC; k < nz - GC; ++k) {
      const int gi[4] = {0, 1, nx - 2, nx - 1};
      for (int q = 0; q < 4; ++q) {
        const size_t m = IC(d, gi[q], j, k);
        u1[m] = u0[m]; v1[m] = v0[m]; w1[m] = w0[m];
      }
      const size_t c = IC(d, nx - 3, j, k);
      u1[c] = u0[c];
    }
  for (int i = GC; i < nx - GC; ++i)
    for (int k = GC; k < nz - GC; ++k) {
      const int gj[4] = {0, 1, ny - 2, ny - 1};
      for (int q = 0; q < 4; ++q) {
        const size_t m = IC(d, i, gj[q], k);
        u1[m] = u0[m]; v1[m] = v0[m]; w1[m] = w0[m];
      }
      const size_t c = IC(d, i, ny - 3, k);
      v1[c] = v0[c];
    }
  if (pz) {
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        const int gk[4] = {0, 1, nz - 2, nz - 1};
        for (int q = 0; q < 4; ++q) {
          const size_t m = IC(d, i, j, gk[q]);
          u1[m] = u0[m]; v1[m] = v0[m]; w1[m] = w0[m];
        }
      }
    return;
  }
  for (int i = GC; i < nx - GC; ++i)
    for (int j = GC; j < ny - GC; ++j) {
      const int gk[4] = {0, 1, nz - 2, nz - 1};
      for (int q = 0; q < 4; ++q) {
        const size_t m = IC(d, i, j, gk[q]);
        u1[m] = u0[m]; v1[m] = v0[m]; w1[m] = w0[m];
      }
      const size_t c = IC(d, i, j, nz - 3);
      w1[c] = w0[c];
    }
}

/* virtualF_Int (virtualForceIntergrator.hpp:5-37) for x and y, with
   delta = Dx*Dy*Dz (Generator.hpp:47-53); cD = -2 sum / lz (:123-127) */
void orc_drag(const cfd_grid_desc *g, const double *f, double *cd, double *cl) {
  const dims_t d = dims_of(g);
  const size_t N = d.N;
  double sums[2];
  for (int c = 0; c < 2; ++c) {
    double part[ORC_NBLK];
#pragma omp parallel for schedule(static)
    for (int blk = 0; blk < ORC_NBLK; ++blk) {
      size_t lo = N * (size_t)blk / ORC_NBLK, hi = N * (size_t)(blk + 1) / ORC_NBLK;
      double s = 0.0;
      for (size_t r = lo; r < hi; ++r) {
        const int i = (int)(r / ((size_t)d.nyc * d.nzc)) + GC;
        const int j = (int)((r / (size_t)d.nzc) % (size_t)d.nyc) + GC;
        const int k = (int)(r % (size_t)d.nzc) + GC;
        s += f[(size_t)c * N + r] * (g->Dx[i] * g->Dy[j] * g->Dz[k]);
      }
      part[blk] = s;
    }
    double s = 0.0;
    for (int blk = 0; blk < ORC_NBLK; ++blk) s += part[blk];
    sums[c] = s;
  }
  *cd = -2.0 * sums[0] / g->lz;
  *cl = -2.0 * sums[1] / g->lz;
}

/* CheckSteadyStatebyMaxVal_omp (CheckSteadystate.hpp:165-219) */
void orc_maxdiff(const cfd_grid_desc *g, const double *u0, const double *v0, const double *w0,
                 const double *u3, const double *v3, const double *w3, double out[3]) {
  const dims_t d = dims_of(g);
  const double *A[3] = {u0, v0, w0}, *B[3] = {u3, v3, w3};
  for (int c = 0; c < 3; ++c) {
    double m = 0.0;
    for (int i = GC; i < d.nx - GC; ++i)
      for (int j = GC; j < d.ny - GC; ++j)
        for (int k = GC; k < d.nz - GC; ++k) {
          const size_t ic = IC(d, i, j, k);
          const double dif = A[c][ic] - B[c][ic];
          if (dif > m) m = dif;
        }
    out[c] = m;
  }
}

/* ========================================================================
 * solver::bicgstabRe2 (inc/solver/bicgstabRe2.hpp:13-189)
 * ===================================================================== */
struct orc_bicg {
  size_t N;
  int op;
  const int32_t *ptr; /* CSR row pointers */
  const int32_t *col;
  const double *val;
  const cfd_grid_desc *g;
  const double *coef, *jp;
  int pz; /* SPE: periodic z */
  double *r0, *r1, *p1, *s1, *ap, *as, *px1, *px2;
  int iters_max;
  double tol, restart_factor, restart_tol;
  int timestep, total_iters, resets, restarts;
};

static void bicg_matvec(orc_bicg *s, const double *x, double *y) {
  if (s->op == CFD_OP_SPE) orc_spmv_spe_bl(s->g, s->pz, s->coef, s->jp, x, y);
  else if (s->op == CFD_OP_CSR) orc_spmv_csr(s->N, s->ptr, s->col, s->val, x, y);
  else orc_spmv_ell(s->N, 7, s->col, s->val, x, y);
}

orc_bicg *orc_bicg_create(size_t N, int op, const int32_t *col, const double *val,
                          const cfd_grid_desc *g, const double *spe_coef, const double *spe_jp) {
  orc_bicg *s = (orc_bicg *)calloc(1, sizeof(orc_bicg));
  if (!s) return NULL;
  s->N = N; s->op = op; s->col = col; s->val = val; s->g = g; s->coef = spe_coef; s->jp = spe_jp;
  double **vecs[8] = {&s->r0, &s->r1, &s->p1, &s->s1, &s->ap, &s->as, &s->px1, &s->px2};
  for (int q = 0; q < 8; ++q) {
    *vecs[q] = (double *)calloc(N ? N : 1, sizeof(double));
    if (!*vecs[q]) { orc_bicg_destroy(s); return NULL; }
  }
  s->iters_max = 5000;
  s->tol = 1e-3;
  s->restart_factor = 1e-7;
  s->restart_tol = s->restart_factor * s->tol;
  return s;
}

orc_bicg *orc_bicg_create_csr(size_t N, const int32_t *ptr, const int32_t *col, const double *val) {
  orc_bicg *s = orc_bicg_create(N, CFD_OP_CSR, col, val, NULL, NULL, NULL);
  if (s) s->ptr = ptr;
  return s;
}

void orc_bicg_destroy(orc_bicg *s) {
  if (!s) return;
  free(s->r0); free(s->r1); free(s->p1); free(s->s1); free(s->ap); free(s->as); free(s->px1); free(s->px2);
  free(s);
}

void orc_bicg_set_tolerance(orc_bicg *s, double tol) { s->tol = tol; s->restart_tol = s->restart_factor * tol; }
void orc_bicg_set_maxit(orc_bicg *s, int maxit) { s->iters_max = maxit; }
void orc_bicg_set_periodic_z(orc_bicg *s, int pz) { s->pz = pz; }
int orc_bicg_restarts(const orc_bicg *s) { return s->restarts; }
int orc_bicg_resets(const orc_bicg *s) { return s->resets; }

static void vcopy(size_t n, const double *a, double *r) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) r[i] = a[i];
}

double orc_bicg_solve(orc_bicg *s, const double *rhs, double *x, int *iters_out) {
  const size_t n = s->N;
  s->timestep++;
  double r1r0, pre_r1r0, a, w, b;
  const double norm0 = sqrt(orc_dot(n, rhs, rhs));
  bicg_matvec(s, x, s->r0);
  vcopy(n, s->px1, s->px2);
  vcopy(n, x, s->px1);
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) s->r0[i] = rhs[i] - s->r0[i];
  vcopy(n, s->r0, s->r1);
  vcopy(n, s->r0, s->p1);
  r1r0 = orc_dot(n, s->r1, s->r0);
  int numIter = 0;
  for (;;) {
    bicg_matvec(s, s->p1, s->ap);
    a = r1r0 / orc_dot(n, s->ap, s->r0);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) s->s1[i] = s->r1[i] - a * s->ap[i];
    bicg_matvec(s, s->s1, s->as);
    w = orc_dot(n, s->as, s->s1) / orc_dot(n, s->as, s->as);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
      x[i] += a * s->p1[i] + w * s->s1[i];
      s->r1[i] = s->s1[i] - w * s->as[i];
    }
    if ((sqrt(orc_dot(n, s->r1, s->r1)) / norm0) < s->tol) break;
    if (numIter >= s->iters_max) {
      long long reset = 0;
#pragma omp parallel for schedule(static) reduction(+ : reset)
      for (long long i = 0; i < (long long)n; ++i)
        if (isnan(x[i]) || isinf(x[i])) { x[i] = s->px1[i]; ++reset; }
      if ((double)reset > (double)n * 0.1) { vcopy(n, s->px2, x); ++s->resets; }
      else if (reset != 0) ++s->resets;
      break;
    }
    pre_r1r0 = r1r0;
    r1r0 = orc_dot(n, s->r1, s->r0);
    b = (a / w) * (r1r0 / pre_r1r0);
    if (s->timestep < 10 || fabs(pre_r1r0) > s->restart_tol) {
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < (long long)n; ++i) s->p1[i] = s->r1[i] + b * (s->p1[i] - w * s->ap[i]);
    } else {
      vcopy(n, s->r1, s->r0);
      vcopy(n, s->r1, s->p1);
      ++s->restarts;
    }
    ++numIter;
    ++s->total_iters;
  }
  *iters_out = numIter;
  return sqrt(orc_dot(n, s->r1, s->r1)) / norm0;
}

/* ========================================================================
 * red-black SOR on the ELL operator (no reference counterpart: the
 * reference SOR is lexicographic and in place, SorPipeLine.seq.hpp)
 * ===================================================================== */
double orc_rbsor(size_t N, const cfd_grid_desc *g, const int32_t *col, const double *val,
                 const double *b, double *x, double omega, double tol, int maxit,
                 int check_every, int *iters) {
  const dims_t d = dims_of(g);
  double *r = (double *)malloc(sizeof(double) * (N ? N : 1));
  const double nb = sqrt(orc_dot(N, b, b));
  double rel = INFINITY;
  int it = 0;
  if (check_every < 1) check_every = 1;
  while (it < maxit) {
    for (int colour = 0; colour < 2; ++colour) {
#pragma omp parallel for schedule(static)
      for (int ic = 0; ic < d.nxc; ++ic)
        for (int jc = 0; jc < d.nyc; ++jc)
          for (int kc = ((ic + jc + colour) & 1); kc < d.nzc; kc += 2) {
            const size_t row = ((size_t)ic * d.nyc + jc) * d.nzc + kc;
            double acc = 0.0;
            for (int s = 0; s < 7; ++s) acc = fma(val[(size_t)s * N + row], x[col[(size_t)s * N + row]], acc);
            x[row] = x[row] + omega * (b[row] - acc) / val[row];
          }
    }
    ++it;
    if (it % check_every == 0 || it == maxit) {
      orc_spmv_ell(N, 7, col, val, x, r);
#pragma omp parallel for schedule(static)
      for (long long q = 0; q < (long long)N; ++q) r[q] = b[q] - r[q];
      rel = sqrt(orc_dot(N, r, r)) / nb;
      if (rel < tol) break;
    }
  }
  free(r);
  *iters = it;
  return rel;
}

/* red-black SOR on the CSR operator (no reference counterpart): the
   colouring, update and residual check of orc_rbsor, with each row's
   products summed in stored order (CSR_matrix::multiply_omp,
   CSR_sparseMatrix.hpp:278-292) and the diagonal the entry whose column is
   the row */
double orc_rbsor_csr(size_t N, const cfd_grid_desc *g, const int32_t *ptr, const int32_t *col, const double *val,
                     const double *b, double *x, double omega, double tol, int maxit, int check_every, int *iters) {
  const dims_t d = dims_of(g);
  double *r = (double *)malloc(sizeof(double) * (N ? N : 1));
  const double nb = sqrt(orc_dot(N, b, b));
  double rel = INFINITY;
  int it = 0;
  if (check_every < 1) check_every = 1;
  while (it < maxit) {
    for (int colour = 0; colour < 2; ++colour) {
#pragma omp parallel for schedule(static)
      for (int ic = 0; ic < d.nxc; ++ic)
        for (int jc = 0; jc < d.nyc; ++jc)
          for (int kc = ((ic + jc + colour) & 1); kc < d.nzc; kc += 2) {
            const size_t row = ((size_t)ic * d.nyc + jc) * d.nzc + kc;
            double acc = 0.0, diag = 0.0;
            for (int32_t e = ptr[row]; e < ptr[row + 1]; ++e) {
              acc = fma(val[e], x[col[e]], acc);
              if ((size_t)col[e] == row) diag = val[e];
            }
            x[row] = x[row] + omega * (b[row] - acc) / diag;
          }
    }
    ++it;
    if (it % check_every == 0 || it == maxit) {
      orc_spmv_csr(N, ptr, col, val, x, r);
#pragma omp parallel for schedule(static)
      for (long long q = 0; q < (long long)N; ++q) r[q] = b[q] - r[q];
      rel = sqrt(orc_dot(N, r, r)) / nb;
      if (rel < tol) break;
    }
  }
  free(r);
  *iters = it;
  return rel;
}

/* SorPipeLine_seq's volume-weighted update (SorPipeLine.seq.hpp:26-112) in
   red-black order: the cells with (i + j + k) even, then odd (interior
   indices), each reading the other colour; cf0..cf7 formed per cell in the
   reference's operation order; the same max |pNEW - p| stop and
   iters = sweeps - 1.  Converges to the lexicographic sweep's fixed point. */
double orc_sor_rb_vw(const cfd_grid_desc *g, double dt, double omega, double crit, int itmax,
                     const double *u1, const double *v1, const double *w1, double *p, int *iters) {
  return orc_sor_rb_vw_bl(g, 0, dt, omega, crit, itmax, u1, v1, w1, p, iters);
}

/* periodic z (pz): the z neighbours of the first and last planes are the
   live pressure across the seam, not ghosts */
double orc_sor_rb_vw_bl(const cfd_grid_desc *g, int pz, double dt, double omega, double crit, int itmax,
                        const double *u1, const double *v1, const double *w1, double *p, int *iters) {
  const dims_t d = dims_of(g);
  const size_t N = d.N;
  double *cf6 = (double *)malloc(sizeof(double) * (N ? N : 1));
  const double rdt = 1.0 / dt;
  const double *Dx = g->Dx, *Dy = g->Dy, *Dz = g->Dz, *Dxs = g->Dxs, *Dys = g->Dys, *Dzs = g->Dzs;
  for (int i = GC; i < d.nx - GC; ++i)
    for (int j = GC; j < d.ny - GC; ++j)
      for (int k = GC; k < d.nz - GC; ++k) {
        const size_t icel = IC(d, i, j, k);
        const double mChange = (u1[icel] - u1[IC(d, i - 1, j, k)]) * Dy[j] * Dz[k] +
                               (v1[icel] - v1[IC(d, i, j - 1, k)]) * Dx[i] * Dz[k] +
                               (w1[icel] - w1[IC(d, i, j, k - 1)]) * Dx[i] * Dy[j];
        cf6[ICAL(d, i, j, k)] = mChange * rdt;
      }
  int it = 0;
  double residual;
  do {
    residual = 0.0;
    for (int colour = 0; colour < 2; ++colour) {
      double rmax = 0.0;
#pragma omp parallel for schedule(static) reduction(max : rmax)
      for (int ic = 0; ic < d.nxc; ++ic)
        for (int jc = 0; jc < d.nyc; ++jc)
          for (int kc = ((ic + jc + colour) & 1); kc < d.nzc; kc += 2) {
            const int i = ic + GC, j = jc + GC, k = kc + GC;
            const size_t icel = IC(d, i, j, k);
            const double cf0 = Dy[j] * Dz[k] / Dxs[i], cf1 = Dy[j] * Dz[k] / Dxs[i - 1];
            const double cf2 = Dx[i] * Dz[k] / Dys[j], cf3 = Dx[i] * Dz[k] / Dys[j - 1];
            const double cf4 = Dx[i] * Dy[j] / Dzs[k], cf5 = Dx[i] * Dy[j] / Dzs[k - 1];
            const double coef = Dy[j] * Dz[k] / Dxs[i] + Dy[j] * Dz[k] / Dxs[i - 1] + Dx[i] * Dz[k] / Dys[j] +
                                Dx[i] * Dz[k] / Dys[j - 1] + Dx[i] * Dy[j] / Dzs[k] + Dx[i] * Dy[j] / Dzs[k - 1];
            const double cf7 = -1.0 / coef;
            const double pc = p[icel];
            const int kp = (pz && k == d.nz - 3) ? GC : k + 1, km = (pz && k == GC) ? d.nz - 3 : k - 1;
            const double pNEW = (-p[IC(d, i + 1, j, k)] * cf0 - p[IC(d, i - 1, j, k)] * cf1 -
                                 p[IC(d, i, j + 1, k)] * cf2 - p[IC(d, i, j - 1, k)] * cf3 -
                                 p[IC(d, i, j, kp)] * cf4 - p[IC(d, i, j, km)] * cf5 +
                                 cf6[((size_t)ic * d.nyc + jc) * d.nzc + kc]) * cf7;
            const double ch = fabs(pNEW - pc);
            if (ch > rmax) rmax = ch;
            p[icel] = pc + omega * (pNEW - pc);
          }
      if (rmax > residual) residual = rmax;
    }
    ++it;
  } while (residual > crit && it < itmax);
  free(cf6);
  *iters = it - 1;
  return residual;
}

/* SorPipeLine_seq (inc/pressure/solver/SorPipeLine.seq.hpp:4-122) with its
   coefficients computed on every call (the reference computes them only when
   simuClass::loop == 1, a loop in which the driver never calls it) */
double orc_sor_lex(const cfd_grid_desc *g, double dt, double omega, double crit, int itmax,
                   const double *u1, const double *v1, const double *w1, double *p, int *iters) {
  return orc_sor_lex_bl(g, 0, dt, omega, crit, itmax, u1, v1, w1, p, iters);
}

double orc_sor_lex_bl(const cfd_grid_desc *g, int pz, double dt, double omega, double crit, int itmax,
                      const double *u1, const double *v1, const double *w1, double *p, int *iters) {
  const dims_t d = dims_of(g);
  const size_t N = d.N;
  double *cf = (double *)malloc(sizeof(double) * 8 * (N ? N : 1));
  const double rdt = 1.0 / dt;
  const double *Dx = g->Dx, *Dy = g->Dy, *Dz = g->Dz, *Dxs = g->Dxs, *Dys = g->Dys, *Dzs = g->Dzs;
  for (int i = GC; i < d.nx - GC; ++i)
    for (int j = GC; j < d.ny - GC; ++j)
      for (int k = GC; k < d.nz - GC; ++k) {
        const size_t ii = ICAL(d, i, j, k) * 8;
        cf[ii + 0] = Dy[j] * Dz[k] / Dxs[i];
        cf[ii + 1] = Dy[j] * Dz[k] / Dxs[i - 1];
        cf[ii + 2] = Dx[i] * Dz[k] / Dys[j];
        cf[ii + 3] = Dx[i] * Dz[k] / Dys[j - 1];
        cf[ii + 4] = Dx[i] * Dy[j] / Dzs[k];
        cf[ii + 5] = Dx[i] * Dy[j] / Dzs[k - 1];
        const double coef = Dy[j] * Dz[k] / Dxs[i] + Dy[j] * Dz[k] / Dxs[i - 1] + Dx[i] * Dz[k] / Dys[j] +
                            Dx[i] * Dz[k] / Dys[j - 1] + Dx[i] * Dy[j] / Dzs[k] + Dx[i] * Dy[j] / Dzs[k - 1];
        cf[ii + 7] = -1.0 / coef; /* grid::coef (structureGrid.hpp:278-287) */
        const size_t icel = IC(d, i, j, k);
        const double mChange = (u1[icel] - u1[IC(d, i - 1, j, k)]) * Dy[j] * Dz[k] +
                               (v1[icel] - v1[IC(d, i, j - 1, k)]) * Dx[i] * Dz[k] +
                               (w1[icel] - w1[IC(d, i, j, k - 1)]) * Dx[i] * Dy[j];
        cf[ii + 6] = mChange * rdt;
      }
  int it = 0;
  double residual;
  do {
    residual = 0.0;
    for (int i = GC; i < d.nx - GC; ++i)
      for (int j = GC; j < d.ny - GC; ++j)
        for (int k = GC; k < d.nz - GC; ++k) {
          const size_t icel = IC(d, i, j, k);
          const size_t ii8 = ICAL(d, i, j, k) * 8;
          const int kp = (pz && k == d.nz - 3) ? GC : k + 1, km = (pz && k == GC) ? d.nz - 3 : k - 1;
          const double pNEW = (-p[IC(d, i + 1, j, k)] * cf[ii8] - p[IC(d, i - 1, j, k)] * cf[ii8 + 1] -
                               p[IC(d, i, j + 1, k)] * cf[ii8 + 2] - p[IC(d, i, j - 1, k)] * cf[ii8 + 3] -
                               p[IC(d, i, j, kp)] * cf[ii8 + 4] - p[IC(d, i, j, km)] * cf[ii8 + 5] +
                               cf[ii8 + 6]) * cf[ii8 + 7];
          const double pChange = fabs(pNEW - p[icel]);
          p[icel] += omega * (pNEW - p[icel]);
          if (pChange > residual) residual = pChange;
        }
    ++it;
  } while (residual > crit && it < itmax);
  free(cf);
  *iters = it - 1;
  return residual;
}

/* ========================================================================
 * Discrete adjoint of the timestep (no reference counterpart).
 *
 * Every function below is the transpose of the forward function named in
 * its comment, written as a scatter over the forward loop: where the forward
 * pass computes y = sum_m a_m x_m, the adjoint adds a_m * ly to lx_m.  A
 * forward assignment that overwrites y also clears ly.  The predictor is
 * nonlinear and its transpose is the transpose of its Jacobian at the
 * forward state (QUICK's upwind choice held fixed), so its pin is the
 * dot-product identity <ly, J dx> = <J^T ly, dx> with J dx from central
 * differences of orc_predictor itself.
 * ===================================================================== */

/* transpose of BC_staggered_main's velocity part (orc_bc_main_ex), faces in
   the reverse of the forward order; the adjoint of the face-0 u Dirichlet
   value is added to g_uin[j - 2] (when g_uin is not NULL) */
static void bc_main_adjoint_impl(const cfd_grid_desc *g, const double Num[6][3], double *g_uin, int conv,
                                 double conv_udt, int pz, double *lu, double *lv, double *lw);

void orc_bc_main_adjoint(const cfd_grid_desc *g, const double Num[6][3], double *g_uin,
                         double *lu, double *lv, double *lw) {
  bc_main_adjoint_impl(g, Num, g_uin, 0, 0.0, 0, lu, lv, lw);
}

/* the same with the boundary mode of P (orc_bc_main_bl): the periodic z
   images fold their ghosts' adjoints back (the z faces were written last),
   and the convective outflow's y_b = (y_old + s y_in) / (1 + s) sends
   S / (1 + s) to the old value and s S / (1 + s) to y_in, S the outputs'
   summed adjoint */
void orc_bc_main_adjoint_bl(const cfd_grid_desc *g, const cfd_params *P, double *g_uin, double *lu, double *lv,
                            double *lw) {
  bc_main_adjoint_impl(g, P->bc_num, g_uin, P->outflow_convective, P->outflow_speed * P->dt, P->periodic_z, lu,
                       lv, lw);
}

/* the adjoint of periodic_z_fill: each ghost plane's into its image's */
static void periodic_z_fold(const dims_t d, double *L) {
  const int nz = d.nz;
  for (int i = 0; i < d.nx; ++i)
    for (int j = 0; j < d.ny; ++j) {
      const size_t b = IC(d, i, j, 0);
      L[b + nz - 4] += L[b + 0]; L[b + 0] = 0.0;
      L[b + nz - 3] += L[b + 1]; L[b + 1] = 0.0;
      L[b + 2] += L[b + nz - 2]; L[b + nz - 2] = 0.0;
      L[b + 3] += L[b + nz - 1]; L[b + nz - 1] = 0.0;
    }
}

static void bc_main_adjoint_impl(const cfd_grid_desc *g, const double Num[6][3], double *g_uin, int conv,
                                 double conv_udt, int pz, double *lu, double *lv, double *lw) {
  const dims_t d = dims_of(g);
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  double *L[3] = {lu, lv, lw};
  /* y[m0] = y[m1] = a*x[m2] + c  ->  lx[m2] += a (ly[m0] + ly[m1]); ly[m0..] = 0 */
#define ADJ2(c, a, m0, m1, m2)                 \
  do {                                         \
    const double s_ = L[c][m0] + L[c][m1];     \
    L[c][m0] = 0.0; L[c][m1] = 0.0;            \
    L[c][m2] += (a) * s_;                      \
  } while (0)
#define ADJ3(c, a, m0, m1, m2, m3)                      \
  do {                                                  \
    const double s_ = L[c][m0] + L[c][m1] + L[c][m2];   \
    L[c][m0] = 0.0; L[c][m1] = 0.0; L[c][m2] = 0.0;     \
    L[c][m3] += (a) * s_;                               \
  } while (0)
#define AN(b, c) (Num[b][c])
#define AT(b, c) (2.0 * Num[b][c] - 1.0)
  if (pz) {
    periodic_z_fold(d, lu);
    periodic_z_fold(d, lv);
    periodic_z_fold(d, lw);
  }
  for (int i = GC; i < nx - GC && !pz; ++i)
    for (int j = GC; j < ny - GC; ++j) { /* BC_5, reverse order of its writes */
      ADJ2(1, AT(5, 1), IC(d, i, j, nz - 1), IC(d, i, j, nz - 2), IC(d, i, j, nz - 3));
      ADJ2(0, AT(5, 0), IC(d, i, j, nz - 1), IC(d, i, j, nz - 2), IC(d, i, j, nz - 3));
      ADJ3(2, AN(5, 2), IC(d, i, j, nz - 1), IC(d, i, j, nz - 2), IC(d, i, j, nz - 3), IC(d, i, j, nz - 4));
    }
  for (int i = GC; i < nx - GC && !pz; ++i)
    for (int j = GC; j < ny - GC; ++j) { /* BC_4 */
      ADJ2(2, AN(4, 2), IC(d, i, j, 0), IC(d, i, j, 1), IC(d, i, j, 2));
      ADJ2(1, AT(4, 1), IC(d, i, j, 0), IC(d, i, j, 1), IC(d, i, j, 2));
      ADJ2(0, AT(4, 0), IC(d, i, j, 0), IC(d, i, j, 1), IC(d, i, j, 2));
    }
  for (int i = GC; i < nx - GC; ++i)
    for (int k = GC; k < nz - GC; ++k) { /* BC_3 */
      ADJ2(2, AT(3, 2), IC(d, i, ny - 1, k), IC(d, i, ny - 2, k), IC(d, i, ny - 3, k));
      ADJ2(0, AT(3, 0), IC(d, i, ny - 1, k), IC(d, i, ny - 2, k), IC(d, i, ny - 3, k));
      ADJ3(1, AN(3, 1), IC(d, i, ny - 1, k), IC(d, i, ny - 2, k), IC(d, i, ny - 3, k), IC(d, i, ny - 4, k));
    }
  for (int i = GC; i < nx - GC; ++i)
    for (int k = GC; k < nz - GC; ++k) { /* BC_2 */
      ADJ2(2, AT(2, 2), IC(d, i, 0, k), IC(d, i, 1, k), IC(d, i, 2, k));
      ADJ2(1, AN(2, 1), IC(d, i, 0, k), IC(d, i, 1, k), IC(d, i, 2, k));
      ADJ2(0, AT(2, 0), IC(d, i, 0, k), IC(d, i, 1, k), IC(d, i, 2, k));
    }
  for (int j = GC; j < ny - GC; ++j)
    for (int k = GC; k < nz - GC; ++k) { /* BC_1 */
      const size_t n1 = IC(d, nx - 1, j, k), n2 = IC(d, nx - 2, j, k), n3 = IC(d, nx - 3, j, k),
                   n4 = IC(d, nx - 4, j, k);
      if (conv) {
        const double su = conv_udt / g->Dx[nx - 3], st = conv_udt / g->Dxs[nx - 3];
        for (int c = 2; c >= 1; --c) { /* w, v: (y[n2] + st y[n3]) / (1 + st) -> n2, n1 */
          const double S = L[c][n1] + L[c][n2];
          L[c][n1] = 0.0;
          L[c][n2] = S / (1.0 + st);
          L[c][n3] += st * S / (1.0 + st);
        }
        const double S = lu[n1] + lu[n2] + lu[n3]; /* u: (u[n3] + su u[n4]) / (1 + su) -> n3, n2, n1 */
        lu[n1] = 0.0; lu[n2] = 0.0;
        lu[n3] = S / (1.0 + su);
        lu[n4] += su * S / (1.0 + su);
      } else {
        ADJ2(2, AT(1, 2), n1, n2, n3);
        ADJ2(1, AT(1, 1), n1, n2, n3);
        ADJ3(0, AN(1, 0), n1, n2, n3, n4);
      }
    }
  for (int j = GC; j < ny - GC; ++j)
    for (int k = GC; k < nz - GC; ++k) { /* BC_0 */
      ADJ2(2, AT(0, 2), IC(d, 0, j, k), IC(d, 1, j, k), IC(d, 2, j, k));
      ADJ2(1, AT(0, 1), IC(d, 0, j, k), IC(d, 1, j, k), IC(d, 2, j, k));
      const size_t m0 = IC(d, 0, j, k), m1 = IC(d, 1, j, k);
      const double s_ = lu[m0] + lu[m1];
      if (g_uin) g_uin[j - GC] += s_;
      lu[m0] = 0.0; lu[m1] = 0.0;
      lu[IC(d, 2, j, k)] += AN(0, 0) * s_;
    }
#undef ADJ2
#undef ADJ3
#undef AN
#undef AT
}

/* transpose of orc_dfib_update_ex at its forward input (T2 recorded by the
   forward pass): from the adjoints of the new velocity (lu3.., ghosted; NULL
   when the pass's velocity is overwritten later) and of f (lf, 3*N) to the
   adjoints of u* (added into lu1.. at the update positions), of the
   interior pressure (added into lX, N) and of the gain (added to *g_gain) */
static void dfib_adjoint_impl(const cfd_grid_desc *g, double dt, double gain, int conv, int pz, const double *eta,
                              const double *t2, const double *lu3, const double *lv3, const double *lw3,
                              const double *lf, double *lu1, double *lv1, double *lw1, double *lX, double *g_gain);

void orc_dfib_adjoint(const cfd_grid_desc *g, double dt, double gain, const double *eta, const double *t2,
                      const double *lu3, const double *lv3, const double *lw3, const double *lf,
                      double *lu1, double *lv1, double *lw1, double *lX, double *g_gain) {
  dfib_adjoint_impl(g, dt, gain, 0, 0, eta, t2, lu3, lv3, lw3, lf, lu1, lv1, lw1, lX, g_gain);
}

/* the same with the boundary mode of P (orc_dfib_update_bl): the outflow
   face of u (no forcing; its + pressure neighbour is the Dirichlet ghost,
   no unknown) and the seam face of w (its + neighbour the first plane) */
void orc_dfib_adjoint_bl(const cfd_grid_desc *g, const cfd_params *P, double gain, const double *eta,
                         const double *t2, const double *lu3, const double *lv3, const double *lw3,
                         const double *lf, double *lu1, double *lv1, double *lw1, double *lX, double *g_gain) {
  dfib_adjoint_impl(g, P->dt, gain, P->outflow_convective, P->periodic_z, eta, t2, lu3, lv3, lw3, lf, lu1, lv1,
                    lw1, lX, g_gain);
}

static void dfib_adjoint_impl(const cfd_grid_desc *g, double dt, double gain, int conv, int pz, const double *eta,
                              const double *t2, const double *lu3, const double *lv3, const double *lw3,
                              const double *lf, double *lu1, double *lv1, double *lw1, double *lX, double *g_gain) {
  const dims_t d = dims_of(g);
  const size_t N = d.N;
  const double *L3[3] = {lu3, lv3, lw3};
  double *L1[3] = {lu1, lv1, lw1};
  const double *Ds[3] = {g->Dxs, g->Dys, g->Dzs}, *Dd[3] = {g->Dx, g->Dy, g->Dz};
  const size_t sg[3] = {d.nzny, (size_t)d.nz, 1};
  const size_t si[3] = {(size_t)d.nyc * d.nzc, (size_t)d.nzc, 1};
  double gg = 0.0;
  for (int q = 0; q < 3; ++q) {
    const int iend = d.nx - GC - (q == 0 && !conv), jend = d.ny - GC - (q == 1), kend = d.nz - GC - (q == 2 && !pz);
    for (int i = GC; i < iend; ++i)
      for (int j = GC; j < jend; ++j)
        for (int k = GC; k < kend; ++k) {
          const size_t ic = IC(d, i, j, k), t = ICAL(d, i, j, k);
          const int I = q == 0 ? i : (q == 1 ? j : k);
          const double ds = Ds[q][I];
          const int outflow = q == 0 && conv && i == d.nx - 3, seam = q == 2 && pz && k == d.nz - 3;
          const double e = outflow ? 0.0 : eta[ic] + (eta[ic + sg[q]] - eta[ic]) * ((ds * 0.5) / Dd[q][I]);
          const double T2 = t2[q * N + t];
          /* u3 = (1 - gain e) T2; f = (u3 - T2)/dt */
          const double lu3t = (L3[q] ? L3[q][ic] : 0.0) + lf[q * N + t] / dt;
          const double lT2 = (1.0 - gain * e) * lu3t - lf[q * N + t] / dt;
          gg += -e * T2 * lu3t;
          L1[q][ic] += lT2;
          /* T2 = u* - dt (p[+] - p) / ds */
          lX[t] += dt / ds * lT2;
          if (seam) lX[t - (size_t)(d.nzc - 1)] -= dt / ds * lT2;
          else if (!outflow) lX[t + si[q]] -= dt / ds * lT2;
        }
  }
  *g_gain += gg;
}

/* transpose of createBMatrix_omp (orc_rhs): lb (N) -> lu1.. (ghosted, added) */
void orc_rhs_adjoint(const cfd_grid_desc *g, double dt, const double *jp, const double *lb,
                     double *lu1, double *lv1, double *lw1) {
  const dims_t d = dims_of(g);
  const double ddt = 1.0 / dt;
  for (int i = GC; i < d.nx - GC; ++i)
    for (int j = GC; j < d.ny - GC; ++j)
      for (int k = GC; k < d.nz - GC; ++k) {
        const size_t ic = IC(d, i, j, k), row = ICAL(d, i, j, k);
        const double s = -jp[row] * ddt * lb[row];
        lu1[ic] += s / g->Dx[i];
        lu1[IC(d, i - 1, j, k)] -= s / g->Dx[i];
        lv1[ic] += s / g->Dy[j];
        lv1[IC(d, i, j - 1, k)] -= s / g->Dy[j];
        lw1[ic] += s / g->Dz[k];
        lw1[IC(d, i, j, k - 1)] -= s / g->Dz[k];
      }
}

/* Coefficients of a QUICK face pair (quickA / quickB above) with respect to
   phi at (pp, p, c, n, nn), for the upwind branches taken at (Up, Un). */
static void quick_coef(int typeA, double Up, double Un, const double *Di, const double *Ds, int I,
                       double cP[5], double cN[5]) {
  for (int m = 0; m < 5; ++m) cP[m] = cN[m] = 0.0;
  if (Up > 0.0) { /* phiP = (c+p)/2 - a ((p-c)/h1 - (c-n)/h0) */
    const double a = typeA ? 0.125 * Di[I + 1] * Di[I + 1] / Ds[I] : 0.125 * Ds[I] * Ds[I] / Di[I];
    const double h1 = typeA ? Di[I + 1] : Ds[I], h0 = typeA ? Di[I] : Ds[I - 1];
    cP[1] = 0.5 - a / h1; cP[2] = 0.5 + a / h1 + a / h0; cP[3] = -a / h0;
  } else {        /* phiP = (c+p)/2 - a ((pp-p)/h2 - (p-c)/h1) */
    const double a = typeA ? 0.125 * Di[I + 1] * Di[I + 1] / Ds[I + 1] : 0.125 * Ds[I] * Ds[I] / Di[I + 1];
    const double h2 = typeA ? Di[I + 2] : Ds[I + 1], h1 = typeA ? Di[I + 1] : Ds[I];
    cP[0] = -a / h2; cP[1] = 0.5 + a / h2 + a / h1; cP[2] = 0.5 - a / h1;
  }
  if (Un > 0.0) { /* phiN = (n+c)/2 - b ((c-n)/h0 - (n-nn)/hm) */
    const double b = typeA ? 0.125 * Di[I] * Di[I] / Ds[I - 1] : 0.125 * Ds[I - 1] * Ds[I - 1] / Di[I - 1];
    const double h0 = typeA ? Di[I] : Ds[I - 1], hm = typeA ? Di[I - 1] : Ds[I - 2];
    cN[2] = 0.5 - b / h0; cN[3] = 0.5 + b / h0 + b / hm; cN[4] = -b / hm;
  } else {        /* phiN = (n+c)/2 - b ((p-c)/h1 - (c-n)/h0) */
    const double b = typeA ? 0.125 * Di[I] * Di[I] / Ds[I] : 0.125 * Ds[I - 1] * Ds[I - 1] / Di[I];
    const double h1 = typeA ? Di[I + 1] : Ds[I], h0 = typeA ? Di[I] : Ds[I - 1];
    cN[1] = -b / h1; cN[2] = 0.5 + b / h1 + b / h0; cN[3] = 0.5 - b / h0;
  }
}

/* face values of quickA / quickB from their coefficients */
static double dot5(const double c[5], const double *phi, const size_t m[5]) {
  double s = 0.0;
  for (int q = 0; q < 5; ++q)
    if (c[q] != 0.0) s += c[q] * phi[m[q]];
  return s;
}

/* transpose of the Jacobian of orc_predictor at (u, v, w): ly1.. hold the
   adjoints of the predicted velocity at the positions the predictor writes;
   the adjoints of its inputs are added into lu, lv, lw (ghosted) */
static void predictor_adjoint_impl(const cfd_grid_desc *g, double nu, double dt, int pz, const double *u,
                                   const double *v, const double *w, const double *lu1, const double *lv1,
                                   const double *lw1, double *lu, double *lv, double *lw);

void orc_predictor_adjoint(const cfd_grid_desc *g, double nu, double dt,
                           const double *u, const double *v, const double *w,
                           const double *lu1, const double *lv1, const double *lw1,
                           double *lu, double *lv, double *lw) {
  predictor_adjoint_impl(g, nu, dt, 0, u, v, w, lu1, lv1, lw1, lu, lv, lw);
}

static void predictor_adjoint_impl(const cfd_grid_desc *g, double nu, double dt, int pz, const double *u,
                                   const double *v, const double *w, const double *lu1, const double *lv1,
                                   const double *lw1, double *lu, double *lv, double *lw) {
  const dims_t d = dims_of(g);
  const size_t sa[3] = {d.nzny, (size_t)d.nz, 1};
  const double *Di[3] = {g->Dx, g->Dy, g->Dz}, *Ds[3] = {g->Dxs, g->Dys, g->Dzs};
  const double *F[3] = {u, v, w};
  const double *LY[3] = {lu1, lv1, lw1};
  double *LX[3] = {lu, lv, lw};
  const double nudt = nu * dt;
  for (int c = 0; c < 3; ++c) {
    const int iend = d.nx - GC - (c == 0), jend = d.ny - GC - (c == 1), kend = d.nz - GC - (c == 2 && !pz);
    for (int i = GC; i < iend; ++i)
      for (int j = GC; j < jend; ++j)
        for (int k = GC; k < kend; ++k) {
          const size_t ic = IC(d, i, j, k);
          const double ly = LY[c][ic];
          if (ly == 0.0) continue;
          const int idx[3] = {i, j, k};
          const double *phi = F[c];
          double *lphi = LX[c];
          lphi[ic] += ly; /* u1 = u + ... */
          for (int a = 0; a < 3; ++a) { /* flux direction a */
            const size_t s = sa[a];
            const int I = idx[a];
            const size_t m[5] = {ic + 2 * s, ic + s, ic, ic - s, ic - 2 * s}; /* pp p c n nn */
            const int own = a == c;
            /* convecting velocities on the +a / -a faces of the component's cell */
            double Vp, Vn;
            int vc = a; /* the velocity component normal to direction a */
            double IT = 0.0;
            size_t vpm0 = 0, vpm1 = 0, vnm0 = 0, vnm1 = 0;
            if (own) {
              Vp = 0.5 * (phi[ic + s] + phi[ic]);
              Vn = 0.5 * (phi[ic - s] + phi[ic]);
            } else {
              /* interpolation of velocity a to the c-face, along direction c */
              const int Ic = idx[c];
              IT = (Ds[c][Ic] * 0.5) / Di[c][Ic];
              const size_t sc = sa[c];
              vpm0 = ic; vpm1 = ic + sc;           /* Vp = F[a][ic] + (F[a][ic+sc] - F[a][ic]) IT */
              vnm0 = ic - s; vnm1 = ic - s + sc;   /* Vn at the -a neighbour */
              Vp = F[vc][vpm0] + (F[vc][vpm1] - F[vc][vpm0]) * IT;
              Vn = F[vc][vnm0] + (F[vc][vnm1] - F[vc][vnm0]) * IT;
            }
            double cP[5], cN[5];
            quick_coef(own, Vp, Vn, Di[a], Ds[a], I, cP, cN);
            const double phP = dot5(cP, phi, m), phN = dot5(cN, phi, m);
            /* conv_a = -dt (phP*Vp' - phN*Vn') / H, own: Vp' = phP, H = Ds; else H = Di */
            const double H = own ? Ds[a][I] : Di[a][I];
            double lP, lN;
            if (own) {
              lP = -dt / H * 2.0 * phP * ly;
              lN = dt / H * 2.0 * phN * ly;
            } else {
              lP = -dt / H * Vp * ly;
              lN = dt / H * Vn * ly;
              const double lVp = -dt / H * phP * ly, lVn = dt / H * phN * ly;
              LX[vc][vpm0] += (1.0 - IT) * lVp;
              LX[vc][vpm1] += IT * lVp;
              LX[vc][vnm0] += (1.0 - IT) * lVn;
              LX[vc][vnm1] += IT * lVn;
            }
            for (int q = 0; q < 5; ++q) lphi[m[q]] += cP[q] * lP + cN[q] * lN;
            /* diffusion along a: ((phi+ - phi)/hp - (phi - phi-)/hm) / H2 */
            double hp, hm, H2;
            if (own) { hp = Di[a][I + 1]; hm = Di[a][I]; H2 = Ds[a][I]; }
            else { hp = Ds[a][I]; hm = Ds[a][I - 1]; H2 = Di[a][I]; }
            const double k1 = nudt * ly / H2;
            lphi[ic + s] += k1 / hp;
            lphi[ic] -= k1 / hp + k1 / hm;
            lphi[ic - s] += k1 / hm;
          }
        }
  }
}

/* transpose of a slot-major ELL matrix into the same layout (padding
   entries, value 0, dropped; rows of the transpose padded with (N-1, 0)) */
int orc_ell_transpose(size_t N, const int32_t *col, const double *val, int32_t *colT, double *valT) {
  int *cnt = (int *)calloc(N ? N : 1, sizeof(int));
  if (!cnt) return -1;
  for (size_t q = 0; q < 7 * N; ++q) { colT[q] = (int32_t)(N - 1); valT[q] = 0.0; }
  for (size_t r = 0; r < N; ++r)
    for (int s = 0; s < 7; ++s) {
      const double a = val[(size_t)s * N + r];
      if (a == 0.0) continue;
      const size_t c = (size_t)col[(size_t)s * N + r];
      if (cnt[c] >= 7) { free(cnt); return -1; }
      colT[(size_t)cnt[c] * N + c] = (int32_t)r;
      valT[(size_t)cnt[c] * N + c] = a;
      cnt[c]++;
    }
  free(cnt);
  return 0;
}

/* ========================================================================
 * runSeqOmp_predictionMethod main loop (RUN_CPU_PC_METHOD.hpp:287-627)
 * ===================================================================== */
struct orc_sim {
  cfd_grid_desc g;
  dims_t d;
  cfd_params prm;
  double *T0[3], *T1[3], *T3[3], *p, *eta, *f, *src, *b, *X, *jp, *va, *vb;
  int32_t *col;
  double *val;
  double *spe_coef, *spe_jp;
  int32_t *csr_ptr, *csr_col;
  double *csr_val, *csr_jp;
  double *vis;      /* Smagorinsky eddy viscosity (N) */
  orc_bicg *bicg;
  int loop;
  double time;
  /* adjoint parameters and the forward record of an objective run */
  double *uin;      /* inflow profile (ny_cal) or NULL */
  double gain;      /* DFIB forcing gain */
  double *t2_rec;   /* where sim_pass writes T2 (3*N), or NULL */
  int adj_steps, adj_avg_from, adj_passes;
  double *adj_x;    /* T0 at the start of every step: steps x 3G */
  double *adj_t2;   /* T2 of every pass: steps x passes x 3N */
  int *adj_solved;  /* loop > 1 for every step */
};

orc_sim *orc_sim_create(const cfd_grid_desc *g, const cfd_params *prm) {
  orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
  if (!s) return NULL;
  s->g = *g;
  s->d = dims_of(g);
  s->prm = *prm;
  const size_t G = s->d.G, N = s->d.N;
  for (int c = 0; c < 3; ++c) {
    s->T0[c] = (double *)calloc(G, sizeof(double));
    s->T1[c] = (double *)calloc(G, sizeof(double));
    s->T3[c] = (double *)calloc(G, sizeof(double));
  }
  s->p = (double *)calloc(G, sizeof(double));
  s->eta = (double *)calloc(G, sizeof(double));
  s->f = (double *)calloc(3 * N, sizeof(double));
  s->src = (double *)calloc(3 * N, sizeof(double));
  s->b = (double *)calloc(N, sizeof(double));
  s->X = (double *)calloc(N, sizeof(double));
  s->jp = (double *)calloc(N, sizeof(double));
  s->va = (double *)calloc(N, sizeof(double));
  s->vb = (double *)calloc(N, sizeof(double));
  s->col = (int32_t *)calloc(7 * N, sizeof(int32_t));
  s->val = (double *)calloc(7 * N, sizeof(double));
  s->spe_coef = (double *)calloc((size_t)2 * (s->d.nxc + s->d.nyc + s->d.nzc), sizeof(double));
  s->spe_jp = (double *)calloc(N, sizeof(double));
  s->csr_ptr = (int32_t *)calloc(N + 1, sizeof(int32_t));
  s->csr_col = (int32_t *)calloc(7 * N, sizeof(int32_t));
  s->csr_val = (double *)calloc(7 * N, sizeof(double));
  s->csr_jp = (double *)calloc(N, sizeof(double));
  s->vis = (double *)calloc(N, sizeof(double));
  s->loop = 1;
  s->time = 0.0;
  s->gain = 1.0;
  return s;
}

static void adj_free(orc_sim *s) {
  free(s->adj_x); free(s->adj_t2); free(s->adj_solved);
  s->adj_x = s->adj_t2 = NULL;
  s->adj_solved = NULL;
  s->adj_steps = 0;
}

void orc_sim_destroy(orc_sim *s) {
  if (!s) return;
  for (int c = 0; c < 3; ++c) { free(s->T0[c]); free(s->T1[c]); free(s->T3[c]); }
  free(s->p); free(s->eta); free(s->f); free(s->src); free(s->b); free(s->X); free(s->jp);
  free(s->va); free(s->vb); free(s->col); free(s->val); free(s->spe_coef); free(s->spe_jp);
  free(s->csr_ptr); free(s->csr_col); free(s->csr_val); free(s->csr_jp); free(s->vis);
  orc_bicg_destroy(s->bicg);
  free(s->uin);
  adj_free(s);
  free(s);
}

int orc_sim_set_inflow(orc_sim *s, const double *uin) {
  free(s->uin);
  s->uin = NULL;
  if (!uin) return 0;
  s->uin = (double *)malloc(sizeof(double) * (size_t)s->d.nyc);
  if (!s->uin) return -1;
  memcpy(s->uin, uin, sizeof(double) * (size_t)s->d.nyc);
  return 0;
}

void orc_sim_set_gain(orc_sim *s, double gain) { s->gain = gain; }

double *orc_sim_field(orc_sim *s, int field, size_t *len) {
  const size_t G = s->d.G, N = s->d.N;
  double *ptr = NULL;
  size_t n = G;
  switch (field) {
    case CFD_T0_U: case CFD_T0_V: case CFD_T0_W: ptr = s->T0[field - CFD_T0_U]; break;
    case CFD_T1_U: case CFD_T1_V: case CFD_T1_W: ptr = s->T1[field - CFD_T1_U]; break;
    case CFD_T3_U: case CFD_T3_V: case CFD_T3_W: ptr = s->T3[field - CFD_T3_U]; break;
    case CFD_P: ptr = s->p; break;
    case CFD_ETA: ptr = s->eta; break;
    case CFD_F: ptr = s->f; n = 3 * N; break;
    case CFD_SRC: ptr = s->src; n = 3 * N; break;
    case CFD_RHS: ptr = s->b; n = N; break;
    case CFD_X: ptr = s->X; n = N; break;
    case CFD_JP: ptr = s->jp; n = N; break;
    case CFD_VEC_A: ptr = s->va; n = N; break;
    case CFD_VEC_B: ptr = s->vb; n = N; break;
    default: n = 0; break;
  }
  if (len) *len = n;
  return ptr;
}

int32_t *orc_sim_ell_col(orc_sim *s) { return s->col; }
int orc_sim_loop(const orc_sim *s) { return s->loop; }
void orc_sim_set_loop(orc_sim *s, int loop, double time) { s->loop = loop; s->time = time; }

int orc_sim_init_state(orc_sim *s, double noise_amp, uint64_t seed) {
  const dims_t d = s->d;
  for (int c = 0; c < 3; ++c) {
    double *t = s->T0[c];
    const double u0 = s->prm.u_init[c];
    for (size_t m = 0; m < d.G; ++m) t[m] = u0;
    if (noise_amp != 0.0) {
      for (int i = GC; i < d.nx - GC; ++i)
        for (int j = GC; j < d.ny - GC; ++j)
          for (int k = GC; k < d.nz - GC; ++k) {
            const size_t ic = IC(d, i, j, k);
            t[ic] += orc_noise(seed, c, ic, noise_amp);
          }
    }
    memcpy(s->T1[c], t, sizeof(double) * d.G);
    memcpy(s->T3[c], t, sizeof(double) * d.G);
  }
  memset(s->p, 0, sizeof(double) * d.G);
  memset(s->X, 0, sizeof(double) * d.N);
  memset(s->f, 0, sizeof(double) * 3 * d.N);
  memset(s->src, 0, sizeof(double) * 3 * d.N);
  orc_bc_main_bl(&s->g, &s->prm, s->uin, s->T0[0], s->T0[1], s->T0[2], s->p);
  orc_bc_copy_bl(&s->g, &s->prm, s->T0[0], s->T0[1], s->T0[2], s->T1[0], s->T1[1], s->T1[2]);
  s->loop = 1;
  s->time = 0.0;
  return 0;
}

int orc_sim_setup(orc_sim *s) {
  const cfd_params *P = &s->prm;
  if (orc_ell_assemble_bl(&s->g, P->p_dirichlet, P->periodic_z, s->col, s->val, s->jp) != 0) return -1;
  orc_spe_setup_bl(&s->g, P->p_dirichlet, P->periodic_z, s->spe_coef, s->spe_jp);
  if (P->op == CFD_OP_SPE) memcpy(s->jp, s->spe_jp, sizeof(double) * s->d.N); /* gA.jp = set_Jp_pc() */
  orc_bicg_destroy(s->bicg);
  if (P->op == CFD_OP_CSR) {
    /* RUN_CPU_PC_METHOD.hpp:153-160: the CSR builder's jp becomes gA.jp */
    orc_csr_assemble(&s->g, s->csr_ptr, s->csr_col, s->csr_val, s->csr_jp);
    memcpy(s->jp, s->csr_jp, sizeof(double) * s->d.N);
    s->bicg = orc_bicg_create_csr(s->d.N, s->csr_ptr, s->csr_col, s->csr_val);
  } else {
    s->bicg = orc_bicg_create(s->d.N, P->op, s->col, s->val, &s->g, s->spe_coef, s->spe_jp);
  }
  if (!s->bicg) return -1;
  orc_bicg_set_periodic_z(s->bicg, P->periodic_z);
  orc_bicg_set_tolerance(s->bicg, P->tol);
  orc_bicg_set_maxit(s->bicg, P->maxit);
  if (P->ib_on) orc_eta_cylinder_z(&s->g, P->ib_cx, P->ib_cy, P->ib_r, P->ib_nsub, s->eta);
  else memset(s->eta, 0, sizeof(double) * s->d.G);
  return 0;
}

static void sim_pass(orc_sim *s, int pass, int accumulate, cfd_step_stats *st) {
  const cfd_params *P = &s->prm;
  const cfd_grid_desc *g = &s->g;
  if (P->les) orc_smagorinsky(g, s->T0[0], s->T0[1], s->T0[2], s->vis);  /* RUN_CPU_PC_METHOD.hpp:298-300 */
  orc_predictor_bl(g, P, P->les ? s->vis : NULL, s->T0[0], s->T0[1], s->T0[2], s->T1[0], s->T1[1], s->T1[2]);
  orc_source_add(g, P->source_index_quirk, s->src, s->T1[0], s->T1[1], s->T1[2]);
  if (P->periodic_z) orc_periodic_refresh(g, s->T1[0], s->T1[1], s->T1[2]); /* before the RHS reads the ghosts */
  int it = -1;
  double rel = 0.0;
  if (s->loop > 1) {
    if (P->solver != CFD_SOLVER_SOR_PIPELINE) orc_rhs(g, P->dt, s->jp, s->T1[0], s->T1[1], s->T1[2], s->b);
    if (P->solver == CFD_SOLVER_SOR_PIPELINE) /* SorPipeLine: in place on the ghosted p, no X */
      rel = orc_sor_rb_vw_bl(g, P->periodic_z, P->dt, P->sor_omega, P->tol, P->maxit, s->T1[0], s->T1[1],
                             s->T1[2], s->p, &it);
    else if (P->solver == CFD_SOLVER_RBSOR && P->op == CFD_OP_CSR)
      rel = orc_rbsor_csr(s->d.N, g, s->csr_ptr, s->csr_col, s->csr_val, s->b, s->X, P->sor_omega, P->tol,
                          P->maxit, P->sor_check, &it);
    else if (P->solver == CFD_SOLVER_RBSOR)
      rel = orc_rbsor(s->d.N, g, s->col, s->val, s->b, s->X, P->sor_omega, P->tol, P->maxit,
                      P->sor_check, &it);
    else
      rel = orc_bicg_solve(s->bicg, s->b, s->X, &it);
    if (P->solver != CFD_SOLVER_SOR_PIPELINE) orc_pressure_scatter(g, s->X, s->p);
    if (P->periodic_z) orc_periodic_refresh_p(g, s->p); /* the update's +z neighbour across the seam */
  }
  orc_dfib_update_bl(g, P, s->gain, s->eta, s->p, s->T1[0], s->T1[1], s->T1[2], s->T3[0], s->T3[1], s->T3[2],
                     s->f, s->t2_rec ? s->t2_rec + (size_t)pass * 3 * s->d.N : NULL);
  if (accumulate) orc_source_accumulate(g, P->dt, s->f, s->src);
  if (st && pass < CFD_MAX_PASSES) {
    st->iters[pass] = it;
    st->relres[pass] = rel;
    orc_drag(g, s->f, &st->cd[pass], &st->cl[pass]);
  }
}

/* FP64 divergence check (no reference counterpart; the divergence is the
   finite-volume form of createBMatrix_omp, inc/matrix/b_matrix.hpp:68-80,
   without its -jp/dt scaling).  Boundary faces (index 1 and n-3 along each
   component's own axis) come from (ub, vb, wb), the others from (u, v, w).
   out[0] max |d| and out[1] sqrt(sum d^2) over the interior cells, out[2] /
   out[3] the same over the fluid cells: eta zero at the cell and its six
   neighbours, and no face on a Dirichlet pressure boundary (pd), whose
   velocity update_UandF leaves uncorrected (inc/dfib/updateUandF.hpp:107-111) */
static void divergence_impl(const cfd_grid_desc *g, const int pd[6], int conv, int pz, const double *u,
                            const double *v, const double *w, const double *ub, const double *vb, const double *wb,
                            const double *eta, double out[4]);

void orc_divergence(const cfd_grid_desc *g, const int pd[6], const double *u, const double *v, const double *w,
                    const double *ub, const double *vb, const double *wb, const double *eta, double out[4]) {
  divergence_impl(g, pd, 0, 0, u, v, w, ub, vb, wb, eta, out);
}

/* with the boundary mode of P: the outflow face (convective outflow) and
   the last z face (periodic z, also the -z face of the first plane) are
   projected faces, read from (u, v, w) */
void orc_divergence_bl(const cfd_grid_desc *g, const cfd_params *P, const double *u, const double *v,
                       const double *w, const double *ub, const double *vb, const double *wb, const double *eta,
                       double out[4]) {
  divergence_impl(g, P->p_dirichlet, P->outflow_convective, P->periodic_z, u, v, w, ub, vb, wb, eta, out);
}

static void divergence_impl(const cfd_grid_desc *g, const int pd[6], int conv, int pz, const double *u,
                            const double *v, const double *w, const double *ub, const double *vb, const double *wb,
                            const double *eta, double out[4]) {
  const dims_t d = dims_of(g);
  const size_t sx = (size_t)d.ny * d.nz, sy = (size_t)d.nz;
  double mx[2] = {0.0, 0.0}, sm[2] = {0.0, 0.0};
  for (int i = GC; i < d.nx - GC; ++i)
    for (int j = GC; j < d.ny - GC; ++j)
      for (int k = GC; k < d.nz - GC; ++k) {
        const size_t ic = IC(d, i, j, k);
        const double uP = (i == d.nx - 3 && !conv ? ub : u)[ic], uM = (i == GC ? ub : u)[ic - sx];
        const double vP = (j == d.ny - 3 ? vb : v)[ic], vM = (j == GC ? vb : v)[ic - sy];
        const double wP = (k == d.nz - 3 && !pz ? wb : w)[ic];
        const double wM = pz ? w[k == GC ? IC(d, i, j, d.nz - 3) : ic - 1] : (k == GC ? wb : w)[ic - 1];
        const double dv = (uP - uM) / g->Dx[i] + (vP - vM) / g->Dy[j] + (wP - wM) / g->Dz[k];
        const double a = fabs(dv);
        if (a > mx[0]) mx[0] = a;
        sm[0] += dv * dv;
        const int dir_face = (pd[0] && i == GC) || (pd[1] && !conv && i == d.nx - 3) || (pd[2] && j == GC) ||
                             (pd[3] && j == d.ny - 3) || (pd[4] && !pz && k == GC) || (pd[5] && !pz && k == d.nz - 3);
        if (!dir_face && eta[ic] == 0.0 && eta[ic - sx] == 0.0 && eta[ic + sx] == 0.0 && eta[ic - sy] == 0.0 &&
            eta[ic + sy] == 0.0 && eta[ic - 1] == 0.0 && eta[ic + 1] == 0.0) {
          if (a > mx[1]) mx[1] = a;
          sm[1] += dv * dv;
        }
      }
  out[0] = mx[0];
  out[1] = sqrt(sm[0]);
  out[2] = mx[1];
  out[3] = sqrt(sm[1]);
}

int orc_sim_step(orc_sim *s, cfd_step_stats *st) {
  const cfd_params *P = &s->prm;
  const cfd_grid_desc *g = &s->g;
  const dims_t d = s->d;
  if (st) { memset(st, 0, sizeof(*st)); st->loop = s->loop; st->n_passes = P->inner_loops + 1; }
  memset(s->src, 0, sizeof(double) * 3 * d.N);
  for (int pass = 0; pass < P->inner_loops; ++pass) sim_pass(s, pass, 1, st);
  sim_pass(s, P->inner_loops, 0, st);
  if (st) {  /* the corrected velocity on the boundary faces the projection saw */
    double dv[4];
    orc_divergence_bl(g, P, s->T3[0], s->T3[1], s->T3[2], s->T1[0], s->T1[1], s->T1[2], s->eta, dv);
    st->div_max = dv[0];
    st->div_l2 = dv[1];
    st->div_fluid_max = dv[2];
    st->div_fluid_l2 = dv[3];
  }
  double md[3];
  orc_maxdiff(g, s->T0[0], s->T0[1], s->T0[2], s->T3[0], s->T3[1], s->T3[2], md);
  if (st) memcpy(st->maxdiff, md, sizeof(md));
  /* a convective outflow steps from the ghost values the last BC left in
     T0: the whole-array T0 <- T3 copy must not replace them by T3's */
  const size_t plane = (size_t)d.ny * d.nz;
  double *keep = NULL;
  if (P->outflow_convective) {
    keep = (double *)malloc(sizeof(double) * 4 * plane);
    for (int c = 1; c < 3; ++c)
      for (int q = 0; q < 2; ++q) memcpy(keep + ((c - 1) * 2 + q) * plane, s->T0[c] + (size_t)(d.nx - 2 + q) * plane,
                                         sizeof(double) * plane);
  }
  for (int c = 0; c < 3; ++c) vcopy(d.G, s->T3[c], s->T0[c]);
  if (keep) {
    for (int c = 1; c < 3; ++c)
      for (int q = 0; q < 2; ++q) memcpy(s->T0[c] + (size_t)(d.nx - 2 + q) * plane, keep + ((c - 1) * 2 + q) * plane,
                                         sizeof(double) * plane);
    free(keep);
  }
  orc_bc_main_bl(g, P, s->uin, s->T0[0], s->T0[1], s->T0[2], s->p);
  orc_bc_copy_bl(g, P, s->T0[0], s->T0[1], s->T0[2], s->T1[0], s->T1[1], s->T1[2]);
  /* simuClass::finishloop (simuClass.hpp:152-169) */
  s->time += P->dt;
  s->loop++;
  return 0;
}

/* ========================================================================
 * Objective and gradient (no reference counterpart): J = mean of the
 * corrector pass's cD over steps [avg_from, nsteps) of a run from the
 * current state; gradient with respect to the inflow profile u_in(y) of the
 * face-0 u Dirichlet value and to the DFIB forcing gain.  The forward run
 * records every step's starting T0 and every pass's T2, so the reverse
 * sweep needs no recomputation; the transposed pressure solve runs
 * BiCGSTAB on the explicit transpose of the assembled ELL operator.
 * ===================================================================== */

double orc_sim_objective(orc_sim *s, int nsteps, int avg_from) {
  const size_t G = s->d.G, N = s->d.N;
  const int passes = s->prm.inner_loops + 1;
  adj_free(s);
  if (nsteps < 1 || avg_from < 0 || avg_from >= nsteps) return NAN;
  s->adj_x = (double *)malloc(sizeof(double) * 3 * G * (size_t)nsteps);
  s->adj_t2 = (double *)calloc(3 * N * (size_t)passes * (size_t)nsteps, sizeof(double));
  s->adj_solved = (int *)calloc((size_t)nsteps, sizeof(int));
  if (!s->adj_x || !s->adj_t2 || !s->adj_solved) { adj_free(s); return NAN; }
  s->adj_steps = nsteps;
  s->adj_avg_from = avg_from;
  s->adj_passes = passes;
  double J = 0.0;
  for (int st = 0; st < nsteps; ++st) {
    for (int c = 0; c < 3; ++c) memcpy(s->adj_x + ((size_t)st * 3 + c) * G, s->T0[c], sizeof(double) * G);
    s->adj_solved[st] = s->loop > 1;
    s->t2_rec = s->adj_t2 + (size_t)st * passes * 3 * N;
    cfd_step_stats stats;
    orc_sim_step(s, &stats);
    s->t2_rec = NULL;
    if (st >= avg_from) J += stats.cd[passes - 1];
  }
  return J / (double)(nsteps - avg_from);
}

/* transpose of BC_staggered_copy (orc_bc_copy): ly at the copied positions
   added into lx */
static void bc_copy_adjoint(const dims_t d, int pz, double *const ly[3], double *const lx[3]) {
  for (int j = GC; j < d.ny - GC; ++j)
    for (int k = GC; k < d.nz - GC; ++k) {
      const int gi[4] = {0, 1, d.nx - 2, d.nx - 1};
      for (int q = 0; q < 4; ++q)
        for (int c = 0; c < 3; ++c) lx[c][IC(d, gi[q], j, k)] += ly[c][IC(d, gi[q], j, k)];
      lx[0][IC(d, d.nx - 3, j, k)] += ly[0][IC(d, d.nx - 3, j, k)];
    }
  for (int i = GC; i < d.nx - GC; ++i)
    for (int k = GC; k < d.nz - GC; ++k) {
      const int gj[4] = {0, 1, d.ny - 2, d.ny - 1};
      for (int q = 0; q < 4; ++q)
        for (int c = 0; c < 3; ++c) lx[c][IC(d, i, gj[q], k)] += ly[c][IC(d, i, gj[q], k)];
      lx[1][IC(d, i, d.ny - 3, k)] += ly[1][IC(d, i, d.ny - 3, k)];
    }
  for (int i = GC; i < d.nx - GC; ++i)
    for (int j = GC; j < d.ny - GC; ++j) {
      const int gk[4] = {0, 1, d.nz - 2, d.nz - 1};
      for (int q = 0; q < 4; ++q)
        for (int c = 0; c < 3; ++c) lx[c][IC(d, i, j, gk[q])] += ly[c][IC(d, i, j, gk[q])];
      if (!pz) lx[2][IC(d, i, j, d.nz - 3)] += ly[2][IC(d, i, j, d.nz - 3)];
    }
}

int orc_sim_gradient(orc_sim *s, double *g_uin, double *g_gain, int *adj_iters) {
  const cfd_params *P = &s->prm;
  const cfd_grid_desc *g = &s->g;
  const dims_t d = s->d;
  const size_t G = d.G, N = d.N;
  if (!s->adj_steps || P->source_index_quirk || P->op != CFD_OP_ELL || P->convection != CFD_CONV_QUICK || P->les)
    return -1;
  const int pz = P->periodic_z, conv = P->outflow_convective;
  const int passes = s->adj_passes, L = passes - 1;
  const double dt = P->dt, nu = 1.0 / P->Re;
  const double wJ = 1.0 / (double)(s->adj_steps - s->adj_avg_from);
  double *buf = (double *)calloc(12 * G + 6 * N + 3 * N + 7 * N, sizeof(double));
  int32_t *colT = (int32_t *)malloc(sizeof(int32_t) * 7 * N);
  if (!buf || !colT) { free(buf); free(colT); return -1; }
  double *lx[3] = {buf, buf + G, buf + 2 * G};
  double *lsum[3] = {buf + 3 * G, buf + 4 * G, buf + 5 * G};
  double *lT1[3] = {buf + 6 * G, buf + 7 * G, buf + 8 * G};
  double *Lam = buf + 12 * G, *lf = Lam + 3 * N, *lX = lf + 3 * N, *lb = lX + N, *zl = lb + N;
  double *valT = zl + N;
  (void)lb;
  if (orc_ell_transpose(N, s->col, s->val, colT, valT) != 0) { free(buf); free(colT); return -1; }
  orc_bicg *bt = orc_bicg_create(N, CFD_OP_ELL, colT, valT, g, NULL, NULL);
  orc_bicg_set_tolerance(bt, P->tol);
  orc_bicg_set_maxit(bt, P->maxit);
  double gg = 0.0;
  int total_it = 0;
  for (int q = 0; q < d.nyc; ++q) g_uin[q] = 0.0;
  for (int st = s->adj_steps - 1; st >= 0; --st) {
    const double *x = s->adj_x + (size_t)st * 3 * G;
    const double *t2 = s->adj_t2 + (size_t)st * passes * 3 * N;
    /* T0' = BC(T3): lx <- adjoint of T3 */
    orc_bc_main_adjoint_bl(g, P, g_uin, lx[0], lx[1], lx[2]);
    memset(lsum[0], 0, sizeof(double) * 3 * G);
    memset(Lam, 0, sizeof(double) * 3 * N);
    for (int q = L; q >= 0; --q) {
      memset(lT1[0], 0, sizeof(double) * 3 * G);
      memset(lX, 0, sizeof(double) * N);
      for (size_t m = 0; m < 3 * N; ++m) lf[m] = dt * Lam[m];
      if (q == L && st >= s->adj_avg_from) /* cD = -2 sum f_x dV / lz */
        for (int i = GC; i < d.nx - GC; ++i)
          for (int j = GC; j < d.ny - GC; ++j)
            for (int k = GC; k < d.nz - GC; ++k)
              lf[ICAL(d, i, j, k)] += wJ * (-2.0 / g->lz) * (g->Dx[i] * g->Dy[j] * g->Dz[k]);
      orc_dfib_adjoint_bl(g, P, s->gain, s->eta, t2 + (size_t)q * 3 * N, q == L ? lx[0] : NULL,
                          q == L ? lx[1] : NULL, q == L ? lx[2] : NULL, lf, lT1[0], lT1[1], lT1[2], lX, &gg);
      if (s->adj_solved[st]) {
        int it = 0;
        orc_bicg_solve(bt, lX, zl, &it); /* A^T zl = lX, warm start from the last adjoint solve */
        total_it += it;
        orc_rhs_adjoint(g, dt, s->jp, zl, lT1[0], lT1[1], lT1[2]);
      }
      /* periodic z: the pass refreshed T1's z ghosts from their images */
      if (pz)
        for (int c = 0; c < 3; ++c) periodic_z_fold(d, lT1[c]);
      /* T1 = P + S_q at every interior cell (source added at its own cell) */
      for (int i = GC; i < d.nx - GC; ++i)
        for (int j = GC; j < d.ny - GC; ++j)
          for (int k = GC; k < d.nz - GC; ++k) {
            const size_t m = IC(d, i, j, k), t = ICAL(d, i, j, k);
            for (int c = 0; c < 3; ++c) Lam[(size_t)c * N + t] += lT1[c][m];
          }
      for (int c = 0; c < 3; ++c)
        for (size_t m = 0; m < G; ++m) lsum[c][m] += lT1[c][m];
    }
    /* T1 = predictor(T0) at the predictor's positions, T0 copied elsewhere;
       the ghosts of v and w on the convective outflow keep their value
       through the step (the T0 <- T3 copy leaves them), so the adjoint their
       end-of-step update left there carries over */
    const size_t plane = (size_t)d.ny * d.nz;
    double *keep = NULL;
    if (conv) {
      keep = (double *)malloc(sizeof(double) * 4 * plane);
      for (int c = 1; c < 3; ++c)
        for (int q = 0; q < 2; ++q)
          memcpy(keep + ((c - 1) * 2 + q) * plane, lx[c] + (size_t)(d.nx - 2 + q) * plane, sizeof(double) * plane);
    }
    memset(lx[0], 0, sizeof(double) * 3 * G);
    if (keep) {
      for (int c = 1; c < 3; ++c)
        for (int q = 0; q < 2; ++q)
          memcpy(lx[c] + (size_t)(d.nx - 2 + q) * plane, keep + ((c - 1) * 2 + q) * plane, sizeof(double) * plane);
      free(keep);
    }
    predictor_adjoint_impl(g, nu, dt, pz, x, x + G, x + 2 * G, lsum[0], lsum[1], lsum[2], lx[0], lx[1], lx[2]);
    bc_copy_adjoint(d, pz, lsum, lx);
  }
  /* the initial state's ghosts: BC_staggered_main of the initial field */
  orc_bc_main_adjoint_bl(g, P, g_uin, lx[0], lx[1], lx[2]);
  *g_gain = gg;
  if (adj_iters) *adj_iters = total_it;
  orc_bicg_destroy(bt);
  free(buf);
  free(colT);
  return 0;
}
