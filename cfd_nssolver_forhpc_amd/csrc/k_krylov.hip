// k_krylov.hip -- restarted BiCGSTAB (solver::bicgstabRe2,
// inc/solver/bicgstabRe2.hpp:91-189) and FP32-inner / FP64-outer defect
// correction on gfx950.
//
// One iteration is five streaming kernels, each fusing the reference's
// BLAS-1 work with the reduction that follows it:
//   K1  ap = A p,               partial (ap . r0)           -> F1: alpha
//   K2  s  = r - alpha ap
//   K3  as = A s,               partials (as . s), (as . as) -> F2: omega
//   K4  x += alpha p + omega s; r = s - omega as;
//                               partials (r . r), (r . r0)   -> F3: stop test, beta, restart rule
//   K5  p  = r + beta (p - omega ap)   (or r0 = p = r on restart)
// F1-F3 are one-block kernels combining the per-block partials in a fixed
// order (deterministic results) and updating the solver's scalars in device
// memory; every kernel returns at once when the device-side `done` flag is
// set, so the host launches iterations in batches and reads the flag once per
// batch instead of once per dot product.  FP64 dots are kept for every
// scalar, also when the vectors are FP32.
#include <algorithm>
#include <cstdio>
#include <mutex>
#include <string>
#include <type_traits>
#include <cmath>

#include "dev_common.h"
#include "dev_spmv.h"

namespace cfd {

namespace {

template <typename T>
__device__ inline T op_apply(const OpArgs &a, int64_t r, const T *__restrict__ x);
template <>
__device__ inline double op_apply<double>(const OpArgs &a, int64_t r, const double *__restrict__ x) {
  return op_row(a, r, x);
}
template <>
__device__ inline float op_apply<float>(const OpArgs &a, int64_t r, const float *__restrict__ x) {
  return op_row32(a, r, x);
}

// (A x)[base + threadIdx.x] for the block of rows starting at base (0 past
// the last row); every thread of the block calls it with the same base
template <typename T>
__device__ inline T op_apply_block(const OpArgs &a, int64_t base, const T *__restrict__ x);
template <>
__device__ inline double op_apply_block<double>(const OpArgs &a, int64_t base, const double *__restrict__ x) {
  return op_row_block(a, base, x);
}
template <>
__device__ inline float op_apply_block<float>(const OpArgs &a, int64_t base, const float *__restrict__ x) {
  const int64_t r = base + threadIdx.x;
  return r < a.N ? op_row32(a, r, x) : 0.0f;
}

__global__ void ks_setup_kernel(KrylovScalars *ks, double tol, double restart_tol, int maxit, int timestep) {
  ks->tol = tol;
  ks->restart_tol = restart_tol;
  ks->maxit = maxit;
  ks->timestep = timestep;
  ks->done = 0;
  ks->hit_maxit = 0;
  ks->iters = 0;
  ks->restart = 0;
}

// r0 = b - A x; r1 = p1 = r0; px2 = px1; px1 = x; partials (b.b), (r0.r0)
template <typename T>
__global__ void __launch_bounds__(256) init_kernel(OpArgs a, const T *__restrict__ b, const T *__restrict__ x,
                                                   T *__restrict__ r0, T *__restrict__ r1, T *__restrict__ p1,
                                                   double *__restrict__ px1, double *__restrict__ px2,
                                                   double *part) {
  double acc[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < a.N; base += stride) {
    const T ax = op_apply_block<T>(a, base, x);
    const int64_t r = base + threadIdx.x;
    if (r >= a.N) continue;
    const T br = b[r];
    const T rr = br - ax;
    r0[r] = rr;
    r1[r] = rr;
    p1[r] = rr;
    if (px1) {  // px2: NULL for a solver without history (its px2 is all zero)
      if (px2) px2[r] = px1[r];
      px1[r] = (double)x[r];
    }
    acc[0] += (double)br * (double)br;
    acc[1] += (double)rr * (double)rr;
  }
  block_partials<2>(acc, part);
}

__global__ void __launch_bounds__(FIN_THREADS) init_finalize(const double *part, int nblk, KrylovScalars *ks) {
  double tot[2];
  combine_partials<2>(part, nblk, tot);
  if (threadIdx.x == 0) {
    ks->norm0 = sqrt(tot[0]);
    ks->rho = tot[1];
    ks->relres = sqrt(tot[1]) / ks->norm0;
  }
}

// BT threads a block: 256 for the CSR operator (op_apply_block stages 256
// rows), SPMV_DOT_THREADS for the others
template <typename T, int BT = 256>
__global__ void __launch_bounds__(BT) spmv_dot_kernel(OpArgs a, const KrylovScalars *__restrict__ ks,
                                                      const T *__restrict__ p1, T *__restrict__ ap,
                                                      const T *__restrict__ r0, double *part) {
  if (ks->done) return;
  double acc[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < a.N; base += stride) {
    const T y = op_apply_block<T>(a, base, p1);
    const int64_t r = base + threadIdx.x;
    if (r >= a.N) continue;
    ntstore(y, ap + r);
    acc[0] += (double)y * (double)r0[r];
  }
  block_partials<1, false, BT>(acc, part);
}

// spmv_dot and spmv_dot2 for the SPE operator over plane columns
// (spe_column_rows): the same rows, a third of the loads; the partials of a
// block sum its columns' planes, so the dot products are ordered differently
// from the row-wise kernels (deterministically)
__global__ void __launch_bounds__(256) spe_spmv_dot_kernel(OpArgs a, SpeTiles T, const KrylovScalars *__restrict__ ks,
                                                           const double *__restrict__ p1, double *__restrict__ ap,
                                                           const double *__restrict__ r0, double *part) {
  if (ks->done) return;
  double acc[1] = {0.0};
  const SpeColumn c = spe_column(a, T);
  if (c.ok) {
    const int64_t nynz = (int64_t)a.nyc * a.nzc, col = (int64_t)c.j * a.nzc + c.k;
    double w[SPE_PLANES];
    spe_plane_values(a, c, r0, w);
    double y[SPE_PLANES], xc[SPE_PLANES];
    const int np = spe_column_rows(a, c, p1, y, xc);
#pragma unroll
    for (int p = 0; p < SPE_PLANES; ++p)
      if (p < np) {
        const int64_t r = (int64_t)(c.i0 + p) * nynz + col;
        ntstore(y[p], ap + r);
        acc[0] += y[p] * w[p];
      }
  }
  block_partials<1>(acc, part);
}

__global__ void __launch_bounds__(256) spe_spmv_dot2_kernel(OpArgs a, SpeTiles T, const KrylovScalars *__restrict__ ks,
                                                            const double *__restrict__ s1, double *__restrict__ as,
                                                            double *part) {
  if (ks->done) return;
  double acc[2] = {0.0, 0.0};
  const SpeColumn c = spe_column(a, T);
  if (c.ok) {
    double y[SPE_PLANES], xc[SPE_PLANES];
    const int np = spe_column_rows(a, c, s1, y, xc);
    const int64_t nynz = (int64_t)a.nyc * a.nzc, col = (int64_t)c.j * a.nzc + c.k;
#pragma unroll
    for (int p = 0; p < SPE_PLANES; ++p)
      if (p < np) {
        ntstore(y[p], as + (int64_t)(c.i0 + p) * nynz + col);
        acc[0] += y[p] * xc[p];
        acc[1] += y[p] * y[p];
      }
  }
  block_partials<2>(acc, part);
}

// The finalize kernels read the scalars they use before the partials, so
// that those reads share the partials' round trip to memory instead of
// preceding it; with `done` set the totals are combined and dropped.
__global__ void __launch_bounds__(FIN_THREADS) alpha_finalize(const double *part, int nblk, KrylovScalars *ks) {
  const int done = ks->done;
  const double rho = ks->rho;
  asm volatile("" ::: "memory");  // issue both loads here, not after the combine
  double tot[1];
  combine_partials<1>(part, nblk, tot);
  if (done) return;
  if (threadIdx.x == 0) ks->alpha = rho / tot[0];
}

template <typename T>
__global__ void __launch_bounds__(256) s_update_kernel(int64_t N, const KrylovScalars *__restrict__ ks,
                                                       const T *__restrict__ r1, const T *__restrict__ ap,
                                                       T *__restrict__ s1) {
  if (ks->done) return;
  const double alpha = ks->alpha;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += stride)
    s1[r] = (T)((double)r1[r] - alpha * (double)ap[r]);
}

template <typename T, int BT = 256>
__global__ void __launch_bounds__(BT) spmv_dot2_kernel(OpArgs a, const KrylovScalars *__restrict__ ks,
                                                       const T *__restrict__ s1, T *__restrict__ as, double *part) {
  if (ks->done) return;
  double acc[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < a.N; base += stride) {
    const T y = op_apply_block<T>(a, base, s1);
    const int64_t r = base + threadIdx.x;
    if (r >= a.N) continue;
    ntstore(y, as + r);
    acc[0] += (double)y * (double)s1[r];
    acc[1] += (double)y * (double)y;
  }
  block_partials<2, false, BT>(acc, part);
}

__global__ void __launch_bounds__(FIN_THREADS) omega_finalize(const double *part, int nblk, KrylovScalars *ks) {
  const int done = ks->done;
  asm volatile("" ::: "memory");
  double tot[2];
  combine_partials<2>(part, nblk, tot);
  if (done) return;
  if (threadIdx.x == 0) ks->omega = tot[0] / tot[1];
}

template <typename T>
__global__ void __launch_bounds__(256) xr_update_kernel(int64_t N, const KrylovScalars *__restrict__ ks,
                                                        T *__restrict__ x, const T *__restrict__ p1,
                                                        const T *__restrict__ s1, const T *__restrict__ as,
                                                        T *__restrict__ r1, const T *__restrict__ r0, double *part) {
  if (ks->done) return;
  const double alpha = ks->alpha, omega = ks->omega;
  double acc[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += stride) {
    const double s = (double)s1[r];
    ntstore((T)((double)x[r] + (alpha * (double)p1[r] + omega * s)), x + r);
    const T rn = (T)(s - omega * (double)as[r]);
    ntstore(rn, r1 + r);
    acc[0] += (double)rn * (double)rn;
    acc[1] += (double)rn * (double)r0[r];
  }
  block_partials<2>(acc, part);
}

// stop test, beta and the restart rule (bicgstabRe2.hpp:137-183)
__global__ void __launch_bounds__(FIN_THREADS) conv_finalize(const double *part, int nblk, KrylovScalars *ks) {
  const KrylovScalars k = *ks;
  asm volatile("" ::: "memory");
  double tot[2];
  combine_partials<2>(part, nblk, tot);
  if (k.done) return;
  if (threadIdx.x == 0) {
    const double rnorm = sqrt(tot[0]), relres = rnorm / k.norm0;
    ks->rnorm = rnorm;
    ks->relres = relres;
    if (relres < k.tol) {
      ks->done = 1;
    } else if (k.iters >= k.maxit) {
      ks->done = 1;
      ks->hit_maxit = 1;
    } else {
      const double pre_rho = k.rho, rho = tot[1];
      ks->pre_rho = pre_rho;
      ks->rho = rho;
      ks->beta = (k.alpha / k.omega) * (rho / pre_rho);
      const int restart = (k.timestep < 10 || fabs(pre_rho) > k.restart_tol) ? 0 : 1;
      ks->restart = restart;
      ks->restarts = k.restarts + restart;
      ks->iters = k.iters + 1;
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) p_update_kernel(int64_t N, const KrylovScalars *__restrict__ ks,
                                                       const T *__restrict__ r1, T *__restrict__ p1,
                                                       const T *__restrict__ ap, T *__restrict__ r0) {
  if (ks->done) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if (ks->restart) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += stride) {
      r0[r] = r1[r];
      p1[r] = r1[r];
    }
    return;
  }
  const double beta = ks->beta, omega = ks->omega;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += stride)
    p1[r] = (T)((double)r1[r] + beta * ((double)p1[r] - omega * (double)ap[r]));
}

// right-preconditioned variant: t = A sh, partials (t . s), (t . t)
template <typename T>
__global__ void __launch_bounds__(256) spmv_dot2_pc_kernel(OpArgs a, const KrylovScalars *__restrict__ ks,
                                                           const T *__restrict__ sh, const T *__restrict__ s1,
                                                           T *__restrict__ tv, double *part) {
  if (ks->done) return;
  double acc[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < a.N; base += stride) {
    const T y = op_apply_block<T>(a, base, sh);
    const int64_t r = base + threadIdx.x;
    if (r >= a.N) continue;
    ntstore(y, tv + r);
    acc[0] += (double)y * (double)s1[r];
    acc[1] += (double)y * (double)y;
  }
  block_partials<2>(acc, part);
}

// the same over SPE plane columns (FP64)
__global__ void __launch_bounds__(256) spe_spmv_dot2_pc_kernel(OpArgs a, SpeTiles T,
                                                               const KrylovScalars *__restrict__ ks,
                                                               const double *__restrict__ sh,
                                                               const double *__restrict__ s1, double *__restrict__ tv,
                                                               double *part) {
  if (ks->done) return;
  double acc[2] = {0.0, 0.0};
  const SpeColumn c = spe_column(a, T);
  if (c.ok) {
    const int64_t nynz = (int64_t)a.nyc * a.nzc, col = (int64_t)c.j * a.nzc + c.k;
    double w[SPE_PLANES];
    spe_plane_values(a, c, s1, w);
    double y[SPE_PLANES], xc[SPE_PLANES];
    const int np = spe_column_rows(a, c, sh, y, xc);
#pragma unroll
    for (int p = 0; p < SPE_PLANES; ++p)
      if (p < np) {
        const int64_t r = (int64_t)(c.i0 + p) * nynz + col;
        ntstore(y[p], tv + r);
        acc[0] += y[p] * w[p];
        acc[1] += y[p] * y[p];
      }
  }
  block_partials<2>(acc, part);
}

// x += alpha ph + omega sh; r = s - omega t; partials (r . r), (r . r0)
template <typename T>
__global__ void __launch_bounds__(256) xr_update_pc_kernel(int64_t N, const KrylovScalars *__restrict__ ks,
                                                           T *__restrict__ x, const T *__restrict__ ph,
                                                           const T *__restrict__ sh, const T *__restrict__ s1,
                                                           const T *__restrict__ tv, T *__restrict__ r1,
                                                           const T *__restrict__ r0, double *part) {
  if (ks->done) return;
  const double alpha = ks->alpha, omega = ks->omega;
  double acc[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += stride) {
    ntstore((T)((double)x[r] + (alpha * (double)ph[r] + omega * (double)sh[r])), x + r);
    const T rn = (T)((double)s1[r] - omega * (double)tv[r]);
    ntstore(rn, r1 + r);
    acc[0] += (double)rn * (double)rn;
    acc[1] += (double)rn * (double)r0[r];
  }
  block_partials<2>(acc, part);
}

// NaN / Inf reset after iters_max (bicgstabRe2.hpp:143-163)
__global__ void __launch_bounds__(256) nan_reset_kernel(int64_t N, double *__restrict__ x,
                                                        const double *__restrict__ px1, double *part) {
  double acc[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += stride) {
    const double v = x[r];
    if (isnan(v) || isinf(v)) {
      x[r] = px1[r];
      acc[0] += 1.0;
    }
  }
  block_partials<1>(acc, part);
}

template <int NV>
__global__ void __launch_bounds__(FIN_THREADS) sum_finalize(const double *part, int nblk, double *out) {
  double tot[NV];
  combine_partials<NV>(part, nblk, tot);
  if (threadIdx.x == 0)
    for (int q = 0; q < NV; ++q) out[q] = tot[q];
}

// defect correction: r = b - A64 x in FP64, r32 = (float) r, partials (r.r), (b.b)
__global__ void __launch_bounds__(256) dc_residual_kernel(OpArgs a, const double *__restrict__ b,
                                                          const double *__restrict__ x, float *__restrict__ r32,
                                                          double *part) {
  double acc[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < a.N; base += stride) {
    const double ax = op_row_block(a, base, x);
    const int64_t r = base + threadIdx.x;
    if (r >= a.N) continue;
    const double br = b[r];
    const double rr = br - ax;
    r32[r] = (float)rr;
    acc[0] += rr * rr;
    acc[1] += br * br;
  }
  block_partials<2>(acc, part);
}

__global__ void __launch_bounds__(256) dc_update_kernel(int64_t N, double *__restrict__ x, float *__restrict__ d) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += stride) {
    x[r] += (double)d[r];
    d[r] = 0.0f;
  }
}

__global__ void __launch_bounds__(256) fill32_kernel(float *__restrict__ a, float v, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

template <typename T>
struct Vecs {
  T *x, *r0, *r1, *p1, *s1, *ap, *as;
  const T *b;
  double *px1, *px2;
};

// threads a block of the iteration's SpMV-dot kernels (ELL and SPE-as-rows
// operators): 512 halves the partials of 256-thread blocks at the same rows
// a thread
constexpr int SPMV_DOT_THREADS = 512;

template <typename T>
hipError_t launch_iters(Ctx &c, const OpArgs &a, const Vecs<T> &v, KrylovScalars *ks, int nit) {
  const int64_t N = c.g.N;
  // The SpMV-dots run SPMV_DOT_THREADS-thread blocks (256 on the CSR
  // operator) and the x / r update four rows a thread, which cuts the
  // partials their one-block finalize kernels combine; the s / p updates
  // take two rows a thread (3% faster streams).
  const bool wide = a.op != CFD_OP_CSR;
  const int nb = blocks_for(N, wide ? SPMV_DOT_THREADS : 256, RED_MAX_BLOCKS);
  const int nbu = blocks_for(N, 2 * 256, RED_MAX_BLOCKS), nbx = blocks_for(N, 4 * 256, RED_MAX_BLOCKS);
  double *p1 = red_part_host(c, RS_SPMV_DOT), *p2 = red_part_host(c, RS_SPMV_DOT2), *p3 = red_part_host(c, RS_XR);
  // the SPE operator's SpMV-dot kernels over plane columns (FP64)
  const SpeTiles tiles = spe_tiles(c.g.nxc, c.g.nyc, c.g.nzc);
  const bool spe = std::is_same<T, double>::value && a.op == CFD_OP_SPE && tiles.blocks <= RED_MAX_BLOCKS;
  const int nbs = spe ? (int)tiles.blocks : nb;
  for (int it = 0; it < nit; ++it) {
    if (spe)
      spe_spmv_dot_kernel<<<nbs, 256, 0, c.stream>>>(a, tiles, ks, reinterpret_cast<const double *>(v.p1),
                                                     reinterpret_cast<double *>(v.ap),
                                                     reinterpret_cast<const double *>(v.r0), p1);
    else if (wide)
      spmv_dot_kernel<T, SPMV_DOT_THREADS><<<nb, SPMV_DOT_THREADS, 0, c.stream>>>(a, ks, v.p1, v.ap, v.r0, p1);
    else
      spmv_dot_kernel<T><<<nb, 256, op_lds_bytes(a), c.stream>>>(a, ks, v.p1, v.ap, v.r0, p1);
    alpha_finalize<<<1, FIN_THREADS, 0, c.stream>>>(p1, nbs, ks);
    s_update_kernel<T><<<nbu, 256, 0, c.stream>>>(N, ks, v.r1, v.ap, v.s1);
    if (spe)
      spe_spmv_dot2_kernel<<<nbs, 256, 0, c.stream>>>(a, tiles, ks, reinterpret_cast<const double *>(v.s1),
                                                      reinterpret_cast<double *>(v.as), p2);
    else if (wide)
      spmv_dot2_kernel<T, SPMV_DOT_THREADS><<<nb, SPMV_DOT_THREADS, 0, c.stream>>>(a, ks, v.s1, v.as, p2);
    else
      spmv_dot2_kernel<T><<<nb, 256, op_lds_bytes(a), c.stream>>>(a, ks, v.s1, v.as, p2);
    omega_finalize<<<1, FIN_THREADS, 0, c.stream>>>(p2, nbs, ks);
    xr_update_kernel<T><<<nbx, 256, 0, c.stream>>>(N, ks, v.x, v.p1, v.s1, v.as, v.r1, v.r0, p3);
    conv_finalize<<<1, FIN_THREADS, 0, c.stream>>>(p3, nbx, ks);
    p_update_kernel<T><<<nbu, 256, 0, c.stream>>>(N, ks, v.r1, v.p1, v.ap, v.r0);
  }
  return hipGetLastError();
}

// Generic driver: set up scalars, r0 = b - A x, then batches of iterations
// with one host read of the scalars per batch.  The batch length follows
// the observed convergence rate so that few no-op launches follow the
// converged iteration.
template <typename T>
int solve_bicg(Ctx &c, const OpArgs &a, const Vecs<T> &v, KrylovScalars *ks, int timestep, double tol, int maxit,
               int *iters, double *relres) {
  const int64_t N = c.g.N;
  const int nb = blocks_for(N, 256, RED_MAX_BLOCKS);
  ks_setup_kernel<<<1, 1, 0, c.stream>>>(ks, tol, 1e-7 * tol, maxit, timestep);
  init_kernel<T><<<nb, 256, op_lds_bytes(a), c.stream>>>(a, v.b, v.x, v.r0, v.r1, v.p1, v.px1, v.px2,
                                                      red_part_host(c, RS_INIT));
  init_finalize<<<1, FIN_THREADS, 0, c.stream>>>(red_part_host(c, RS_INIT), nb, ks);
  CFD_HIP(hipGetLastError());
  int batch = 2;
  double prev_rel = -1.0;
  int prev_it = 0;
  for (;;) {
    CFD_HIP(launch_iters<T>(c, a, v, ks, batch));
    CFD_HIP(hipMemcpyAsync(c.hks, ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, c.stream));
    CFD_HIP(hipStreamSynchronize(c.stream));
    const KrylovScalars h = *c.hks;
    if (h.done) break;
    // predict the iterations still needed from the last batch's rate
    int next = batch * 2;
    if (prev_rel > 0.0 && h.relres > 0.0 && h.relres < prev_rel && h.iters > prev_it) {
      const double rate = std::log(h.relres / prev_rel) / (double)(h.iters - prev_it);
      const double need = std::log(h.tol / h.relres) / rate;
      if (need > 0.0 && std::isfinite(need)) next = (int)std::ceil(need * 1.05) + 1;
    }
    prev_rel = h.relres;
    prev_it = h.iters;
    next = std::max(1, std::min(next, 64));
    next = std::min(next, maxit - h.iters + 1);
    batch = std::max(next, 1);
  }
  const KrylovScalars h = *c.hks;
  if (h.hit_maxit && v.px1) {
    double *part = red_part_host(c, RS_MISC);
    nan_reset_kernel<<<nb, 256, 0, c.stream>>>(N, (double *)v.x, v.px1, part);
    sum_finalize<1><<<1, FIN_THREADS, 0, c.stream>>>(part, nb, c.red.out + RS_MISC * 4);
    CFD_HIP(hipMemcpyAsync(c.hred, c.red.out + RS_MISC * 4, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CFD_HIP(hipStreamSynchronize(c.stream));
    if (c.hred[0] > (double)N * 0.1) {
      if (v.px2) CFD_HIP(hipMemcpyAsync(v.x, v.px2, sizeof(double) * N, hipMemcpyDeviceToDevice, c.stream));
      else CFD_HIP(launch_fill(c, (double *)v.x, 0.0, N));
    }
    if (c.hred[0] > 0.0 && ks == c.ks) c.nan_resets++;
  }
  if (ks == c.ks) c.total_iters += h.iters;
  *iters = h.iters;
  *relres = h.relres;
  return CFD_OK;
}

// BiCGSTAB right-preconditioned by the multigrid V-cycle (an addition; the
// reference's bicgstabRe2 is Jacobi-scaled only):
//   ph = M p; v = A ph; alpha = rho / (v . r0); s = r - alpha v;
//   sh = M s; t = A sh; omega = (t . s) / (t . t);
//   x += alpha ph + omega sh; r = s - omega t; beta; p = r + beta (p - omega v)
// with the fused kernels and device-side scalars of the plain solve.
template <typename T>
struct PVecs {
  T *x, *r0, *r1, *p1, *s1, *v, *t, *ph, *sh;
  const T *b;
};

template <typename T>
int mg_apply_t(Ctx &c, const T *r, T *out, const KrylovScalars *ks);
template <>
int mg_apply_t<double>(Ctx &c, const double *r, double *out, const KrylovScalars *ks) {
  return mg_apply64(c, r, out, ks);
}
template <>
int mg_apply_t<float>(Ctx &c, const float *r, float *out, const KrylovScalars *ks) {
  return mg_apply32(c, r, out, ks);
}

// the preconditioner: the V-cycle in the vectors' precision, or (pc32, FP64
// vectors only) in FP32
template <typename T>
int precondition(Ctx &c, const T *r, T *out, const KrylovScalars *ks, bool pc32) {
  if constexpr (std::is_same<T, double>::value) {
    if (pc32) return mg_apply64_fp32cycle(c, r, out, ks);
  }
  return mg_apply_t<T>(c, r, out, ks);
}

template <typename T>
int launch_pc_iters(Ctx &c, const OpArgs &a, const PVecs<T> &v, KrylovScalars *ks, int nit, bool pc32) {
  const int64_t N = c.g.N;
  const int nb = blocks_for(N, 256, RED_MAX_BLOCKS);
  double *p1 = red_part_host(c, RS_SPMV_DOT), *p2 = red_part_host(c, RS_SPMV_DOT2), *p3 = red_part_host(c, RS_XR);
  const int nbu = blocks_for(N, 2 * 256, RED_MAX_BLOCKS);  // two rows a thread, as in launch_iters
  // the SPE operator's kernels over plane columns (FP64), as in launch_iters
  const SpeTiles tiles = spe_tiles(c.g.nxc, c.g.nyc, c.g.nzc);
  const bool spe = std::is_same<T, double>::value && a.op == CFD_OP_SPE && tiles.blocks <= RED_MAX_BLOCKS;
  const int nbs = spe ? (int)tiles.blocks : nb;
  for (int it = 0; it < nit; ++it) {
    int rc = precondition<T>(c, v.p1, v.ph, ks, pc32);
    if (rc) return rc;
    if (spe)
      spe_spmv_dot_kernel<<<nbs, 256, 0, c.stream>>>(a, tiles, ks, reinterpret_cast<const double *>(v.ph),
                                                     reinterpret_cast<double *>(v.v),
                                                     reinterpret_cast<const double *>(v.r0), p1);
    else
      spmv_dot_kernel<T><<<nb, 256, op_lds_bytes(a), c.stream>>>(a, ks, v.ph, v.v, v.r0, p1);
    alpha_finalize<<<1, FIN_THREADS, 0, c.stream>>>(p1, nbs, ks);
    s_update_kernel<T><<<nbu, 256, 0, c.stream>>>(N, ks, v.r1, v.v, v.s1);
    if ((rc = precondition<T>(c, v.s1, v.sh, ks, pc32))) return rc;
    if (spe)
      spe_spmv_dot2_pc_kernel<<<nbs, 256, 0, c.stream>>>(a, tiles, ks, reinterpret_cast<const double *>(v.sh),
                                                         reinterpret_cast<const double *>(v.s1),
                                                         reinterpret_cast<double *>(v.t), p2);
    else
      spmv_dot2_pc_kernel<T><<<nb, 256, op_lds_bytes(a), c.stream>>>(a, ks, v.sh, v.s1, v.t, p2);
    omega_finalize<<<1, FIN_THREADS, 0, c.stream>>>(p2, nbs, ks);
    xr_update_pc_kernel<T><<<nbu, 256, 0, c.stream>>>(N, ks, v.x, v.ph, v.sh, v.s1, v.t, v.r1, v.r0, p3);
    conv_finalize<<<1, FIN_THREADS, 0, c.stream>>>(p3, nbu, ks);
    p_update_kernel<T><<<nbu, 256, 0, c.stream>>>(N, ks, v.r1, v.p1, v.v, v.r0);
    CFD_HIP(hipGetLastError());
  }
  return CFD_OK;
}

template <typename T>
int solve_pbicg(Ctx &c, const OpArgs &a, const PVecs<T> &v, KrylovScalars *ks, double tol, int maxit, int *iters,
                double *relres, bool pc32 = false) {
  const int64_t N = c.g.N;
  const int nb = blocks_for(N, 256, RED_MAX_BLOCKS);
  ks_setup_kernel<<<1, 1, 0, c.stream>>>(ks, tol, 0.0, maxit, 0);
  init_kernel<T><<<nb, 256, op_lds_bytes(a), c.stream>>>(a, v.b, v.x, v.r0, v.r1, v.p1, nullptr, nullptr,
                                                      red_part_host(c, RS_INIT));
  init_finalize<<<1, FIN_THREADS, 0, c.stream>>>(red_part_host(c, RS_INIT), nb, ks);
  CFD_HIP(hipGetLastError());
  CFD_HIP(hipMemcpyAsync(c.hks, ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, c.stream));
  CFD_HIP(hipStreamSynchronize(c.stream));
  if (!(c.hks->relres >= tol)) {  // already converged (or a zero right-hand side)
    *iters = 0;
    *relres = c.hks->norm0 > 0.0 ? c.hks->relres : 0.0;
    return CFD_OK;
  }
  // Each preconditioned iteration is ~100 launches, so the batch follows the
  // observed convergence rate: launches after the converged iteration are
  // no-ops but still cost their launch latency.
  int batch = 1;
  double prev_rel = c.hks->relres;
  int prev_it = 0;
  for (;;) {
    int rc = launch_pc_iters<T>(c, a, v, ks, batch, pc32);
    if (rc) return rc;
    CFD_HIP(hipMemcpyAsync(c.hks, ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, c.stream));
    CFD_HIP(hipStreamSynchronize(c.stream));
    const KrylovScalars h = *c.hks;
    if (h.done) break;
    int next = 1;
    if (h.relres > 0.0 && h.relres < prev_rel && h.iters > prev_it) {
      const double rate = std::log(h.relres / prev_rel) / (double)(h.iters - prev_it);
      const double need = std::log(h.tol / h.relres) / rate;
      if (need > 0.0 && std::isfinite(need)) next = (int)std::floor(need);
    }
    prev_rel = h.relres;
    prev_it = h.iters;
    batch = std::max(1, std::min(next, 8));
  }
  *iters = c.hks->iters;
  *relres = c.hks->relres;
  return CFD_OK;
}

}  // namespace

int pbicgstab_solve_aux(Ctx &c, const double *b, double *x, double tol, int maxit, bool pc32, int *iters,
                        double *relres) {
  if (!c.ell_ready) return fail(CFD_ERR_STATE, "ELL operator not assembled");
  int rc = mg_ensure(c, pc32);
  if (!rc && pc32) rc = mg_ensure(c, false);  // the FP64 preconditioned directions
  if (rc) return rc;
  PVecs<double> v{x, c.r0, c.r1, c.p1, c.s1, c.ap, c.as, c.ph64, c.sh64, b};
  // the configured operator (ELL, matrix-free SPE or CSR); the V-cycle
  // approximates its inverse from the grid metrics
  return solve_pbicg<double>(c, op_args(c, fp64_op(c)), v, c.ks, tol, maxit, iters, relres, pc32);
}

int pbicgstab_solve64(Ctx &c, const double *b, double *x, double tol, int maxit, bool pc32, int *iters,
                      double *relres) {
  c.solves++;
  return pbicgstab_solve_aux(c, b, x, tol, maxit, pc32, iters, relres);
}

int bicgstab_solve(Ctx &c, const double *b, double *x, double tol, int maxit, int *iters, double *relres) {
  if (int rc = op_ready(c, c.prm.op)) return rc;
  c.solves++;
  Vecs<double> v{x, c.r0, c.r1, c.p1, c.s1, c.ap, c.as, b, c.px1, c.px2};
  return solve_bicg<double>(c, op_args(c, c.prm.op), v, c.ks, c.solves, tol, maxit, iters, relres);
}

int bicgstab_solve_aux(Ctx &c, const double *b, double *x, double tol, int maxit, int timestep, int *iters,
                       double *relres) {
  if (!c.ell_ready) return fail(CFD_ERR_STATE, "ELL operator not assembled");
  Vecs<double> v{x, c.r0, c.r1, c.p1, c.s1, c.ap, c.as, b, nullptr, nullptr};
  return solve_bicg<double>(c, op_args(c, CFD_OP_ELL), v, c.ks, timestep, tol, maxit, iters, relres);
}

// FP64 defect correction around FP32 BiCGSTAB on the FP32 ELL copy:
// r = b - A x (FP64), solve A32 d = r32 to tol_inner, x += d, until
// ||b - A x|| / ||b|| < tol.
int defect_correction(Ctx &c, const double *b, double *x, double tol, double tol_inner, int max_outer, int *outer,
                      int *inner_total, double *relres) {
  if (!c.ell_ready || !c.val32) return fail(CFD_ERR_STATE, "FP32 ELL operator not assembled");
  const int64_t N = c.g.N;
  const int nb = blocks_for(N, 256, RED_MAX_BLOCKS);
  const OpArgs a64 = op_args(c, fp64_op(c));
  const OpArgs a32 = op_args(c, CFD_OP_ELL);
  float *r32 = c.q32[7], *d32 = c.q32[6];
  Vecs<float> v{d32, c.q32[0], c.q32[1], c.q32[2], c.q32[3], c.q32[4], c.q32[5], r32, nullptr, nullptr};
  fill32_kernel<<<blocks_for(N, 256, 1 << 30), 256, 0, c.stream>>>(d32, 0.0f, N);
  CFD_HIP(hipGetLastError());
  int k = 0, inner = 0;
  double rel = 0.0;
  c.solves++;
  for (;;) {
    double *part = red_part_host(c, RS_NORM);
    dc_residual_kernel<<<nb, 256, op_lds_bytes(a64), c.stream>>>(a64, b, x, r32, part);
    sum_finalize<2><<<1, FIN_THREADS, 0, c.stream>>>(part, nb, c.red.out + RS_NORM * 4);
    CFD_HIP(hipGetLastError());
    CFD_HIP(hipMemcpyAsync(c.hred, c.red.out + RS_NORM * 4, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CFD_HIP(hipStreamSynchronize(c.stream));
    rel = std::sqrt(c.hred[0]) / std::sqrt(c.hred[1]);
    if (!(rel >= tol) || k >= max_outer) break;
    int it = 0;
    double rin = 0.0;
    c.solves32++;
    if (c.prm.precond == CFD_PRECOND_MG && c.prm.solver == CFD_SOLVER_BICGSTAB) {
      int rc = mg_ensure(c, true);
      if (rc) return rc;
      PVecs<float> pv{d32, c.q32[0], c.q32[1], c.q32[2], c.q32[3], c.q32[4], c.q32[5], c.ph32, c.sh32, r32};
      rc = solve_pbicg<float>(c, a32, pv, c.ks32, tol_inner, c.prm.maxit, &it, &rin);
      if (rc != CFD_OK) return rc;
    } else if (c.prm.solver == CFD_SOLVER_RBSOR) {
      int rc = rbsor_solve_f32(c, r32, d32, c.prm.sor_omega, tol_inner, c.prm.maxit, c.prm.sor_check, &it, &rin);
      if (rc != CFD_OK) return rc;
    } else {
      int rc = solve_bicg<float>(c, a32, v, c.ks32, c.solves32, tol_inner, c.prm.maxit, &it, &rin);
      if (rc != CFD_OK) return rc;
    }
    inner += it + 1;
    dc_update_kernel<<<nb, 256, 0, c.stream>>>(N, x, d32);
    CFD_HIP(hipGetLastError());
    ++k;
  }
  *outer = k;
  *inner_total = inner;
  *relres = rel;
  return CFD_OK;
}

// Timing of one iteration: the solver runs on scratch vectors (x = VEC_B
// starting from 0) with the stop test disabled; the warm-start history of the
// real solver (px1, px2, X) is not touched.
hipError_t launch_bicg_iteration(Ctx &c) {
  Vecs<double> v{c.f[CFD_VEC_B], c.r0, c.r1, c.p1, c.s1, c.ap, c.as, c.f[CFD_RHS], nullptr, nullptr};
  return launch_iters<double>(c, op_args(c, c.prm.op), v, c.ks32, 1);
}

// Timing of the iteration's first kernel alone (K1: ap = A p with the
// partial dot ap . r0), the iteration's largest share, on the same scratch
// vectors and scalars as launch_bicg_iteration
hipError_t launch_spmv_dot(Ctx &c) {
  const OpArgs a = op_args(c, c.prm.op);
  const SpeTiles tiles = spe_tiles(c.g.nxc, c.g.nyc, c.g.nzc);
  double *part = red_part_host(c, RS_SPMV_DOT);
  if (a.op == CFD_OP_SPE && tiles.blocks <= RED_MAX_BLOCKS)
    spe_spmv_dot_kernel<<<(int)tiles.blocks, 256, 0, c.stream>>>(a, tiles, c.ks32, c.p1, c.ap, c.r0, part);
  else if (a.op != CFD_OP_CSR)  // launch_iters' launch
    spmv_dot_kernel<double, SPMV_DOT_THREADS>
        <<<blocks_for(c.g.N, SPMV_DOT_THREADS, RED_MAX_BLOCKS), SPMV_DOT_THREADS, 0, c.stream>>>(a, c.ks32, c.p1,
                                                                                               c.ap, c.r0, part);
  else
    spmv_dot_kernel<double><<<blocks_for(c.g.N, 256, RED_MAX_BLOCKS), 256, op_lds_bytes(a), c.stream>>>(
        a, c.ks32, c.p1, c.ap, c.r0, part);
  return hipGetLastError();
}

hipError_t launch_bicg_timing_setup(Ctx &c) {
  const int64_t N = c.g.N;
  const int nb = blocks_for(N, 256, RED_MAX_BLOCKS);
  hipError_t e = launch_fill(c, c.f[CFD_VEC_B], 0.0, N);
  if (e != hipSuccess) return e;
  ks_setup_kernel<<<1, 1, 0, c.stream>>>(c.ks32, 0.0, 0.0, 1 << 30, 100);
  const OpArgs a = op_args(c, c.prm.op);
  init_kernel<double><<<nb, 256, op_lds_bytes(a), c.stream>>>(a, c.f[CFD_RHS], c.f[CFD_VEC_B], c.r0, c.r1, c.p1,
                                                              nullptr, nullptr, red_part_host(c, RS_INIT));
  init_finalize<<<1, FIN_THREADS, 0, c.stream>>>(red_part_host(c, RS_INIT), nb, c.ks32);
  return hipGetLastError();
}

}  // namespace cfd

// ---------------------------------------------------------------------------
// Mixed precision by reliable updates: BiCGSTAB iterates in FP32 (FP32 ELL
// values and vectors, FP64 accumulation in the operator and in every dot);
// the FP32 iterate xl is folded into the FP64 solution x and the FP32
// residual is replaced by the FP64 true residual b - A64 x whenever the
// recurrence residual has fallen by `delta` since the last replacement, and
// at the end.  The search direction and shadow residual are kept across a
// replacement, so the Krylov space is not restarted; the stop test is on the
// FP64 true residual, the same criterion as the FP64 solve.
// ---------------------------------------------------------------------------
namespace cfd {
namespace {

__global__ void set_delta_kernel(KrylovScalars *ks, double delta) { ks->delta = delta; }

// r = b - A64 x (FP64); rv = r0 = p = (T) r; xl = 0; partials (b.b), (r.r), (rv.rv)
template <typename T>
__global__ void __launch_bounds__(256) mr_init_kernel(OpArgs a64, const double *__restrict__ b,
                                                      const double *__restrict__ x, T *__restrict__ r32,
                                                      T *__restrict__ r0, T *__restrict__ p, T *__restrict__ xl,
                                                      double *part) {
  double acc[3] = {0.0, 0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < a64.N; base += stride) {
    const double ax = op_row_block(a64, base, x);
    const int64_t r = base + threadIdx.x;
    if (r >= a64.N) continue;
    const double br = b[r];
    const double rr = br - ax;
    const T rf = (T)rr;
    r32[r] = rf;
    r0[r] = rf;
    p[r] = rf;
    if (xl) xl[r] = T(0);
    acc[0] += br * br;
    acc[1] += rr * rr;
    acc[2] += (double)rf * (double)rf;
  }
  block_partials<3>(acc, part);
}

__global__ void __launch_bounds__(FIN_THREADS) mr_init_finalize(const double *part, int nblk, KrylovScalars *ks) {
  double tot[3];
  combine_partials<3>(part, nblk, tot);
  if (threadIdx.x == 0) {
    ks->norm0 = sqrt(tot[0]);
    ks->rmax = sqrt(tot[1]);
    ks->relres = ks->rmax / ks->norm0;
    ks->rho = tot[2];
    ks->replace = 0;
    ks->replacements = 0;
    ks->breakdown = 0;
    if (!(ks->relres >= ks->tol)) ks->done = 1;
  }
}

__global__ void __launch_bounds__(FIN_THREADS) mr_alpha_finalize(const double *part, int nblk, KrylovScalars *ks) {
  if (ks->done) return;
  double tot[1];
  combine_partials<1>(part, nblk, tot);
  if (threadIdx.x == 0) {
    const double a = ks->rho / tot[0];
    if (!isfinite(a)) {  // breakdown: fold what has been gained and re-anchor on the true residual
      ks->breakdown += 1;
      ks->replace = 1;
      ks->alpha = 0.0;
    } else {
      ks->alpha = a;
    }
  }
}

__global__ void __launch_bounds__(FIN_THREADS) mr_omega_finalize(const double *part, int nblk, KrylovScalars *ks) {
  if (ks->done || ks->replace) return;
  double tot[2];
  combine_partials<2>(part, nblk, tot);
  if (threadIdx.x == 0) {
    const double w = tot[0] / tot[1];
    ks->omega = isfinite(w) ? w : 0.0;
    if (!isfinite(w)) {
      ks->breakdown += 1;
      ks->replace = 1;
    }
  }
}

// recurrence residual test: trigger a replacement on a `delta` drop, on the
// stop criterion, or at the iteration cap; otherwise beta and the restart rule
__global__ void __launch_bounds__(FIN_THREADS) mr_conv_finalize(const double *part, int nblk, KrylovScalars *ks) {
  if (ks->done || ks->replace) return;
  double tot[2];
  combine_partials<2>(part, nblk, tot);
  if (threadIdx.x == 0) {
    const double rn = sqrt(tot[0]);
    ks->rnorm = rn;
    if (ks->iters >= ks->maxit) ks->hit_maxit = 1;
    if (rn < ks->delta * ks->rmax || rn / ks->norm0 < ks->tol || ks->hit_maxit || !isfinite(rn)) {
      ks->replace = 1;
    } else {
      ks->pre_rho = ks->rho;
      ks->rho = tot[1];
      ks->beta = (ks->alpha / ks->omega) * (ks->rho / ks->pre_rho);
      ks->restart = (ks->timestep < 10 || fabs(ks->pre_rho) > ks->restart_tol) ? 0 : 1;
      ks->restarts += ks->restart;
      ks->iters += 1;
    }
  }
}

// x += xl; xl = 0 (only when a replacement is due)
__global__ void __launch_bounds__(256) mr_fold_kernel(int64_t N, const KrylovScalars *__restrict__ ks,
                                                      double *__restrict__ x, float *__restrict__ xl) {
  if (ks->done || !ks->replace) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += stride) {
    x[r] += (double)xl[r];
    xl[r] = 0.0f;
  }
}

// rv = (T)(b - A64 x); partials (r.r) in FP64, (rv.r0)
template <typename T>
__global__ void __launch_bounds__(256) mr_replace_kernel(OpArgs a64, const KrylovScalars *__restrict__ ks,
                                                         const double *__restrict__ b, const double *__restrict__ x,
                                                         T *__restrict__ r32, const T *__restrict__ r0,
                                                         double *part) {
  if (ks->done || !ks->replace) return;
  double acc[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < a64.N; base += stride) {
    const double ax = op_row_block(a64, base, x);
    const int64_t r = base + threadIdx.x;
    if (r >= a64.N) continue;
    const double rr = b[r] - ax;
    const T rf = (T)rr;
    r32[r] = rf;
    acc[0] += rr * rr;
    acc[1] += (double)rf * (double)r0[r];
  }
  block_partials<2>(acc, part);
}

__global__ void __launch_bounds__(FIN_THREADS) mr_replace_finalize(const double *part, int nblk, KrylovScalars *ks) {
  if (ks->done || !ks->replace) return;
  double tot[2];
  combine_partials<2>(part, nblk, tot);
  if (threadIdx.x == 0) {
    const double rt = sqrt(tot[0]);
    ks->replacements += 1;
    ks->relres = rt / ks->norm0;
    ks->replace = 0;
    if (ks->relres < ks->tol || ks->hit_maxit || ks->breakdown > 8) {
      ks->done = 1;
      return;
    }
    const bool broke = !(ks->alpha != 0.0 && ks->omega != 0.0);
    ks->rmax = rt;
    ks->pre_rho = ks->rho;
    ks->rho = tot[1];
    ks->beta = broke ? 0.0 : (ks->alpha / ks->omega) * (ks->rho / ks->pre_rho);
    ks->restart = broke ? 1 : ((ks->timestep < 10 || fabs(ks->pre_rho) > ks->restart_tol) ? 0 : 1);
    ks->restarts += ks->restart;
    ks->iters += 1;
  }
}

// p update, shared by both branches: runs unless the solve is done
template <typename T>
__global__ void __launch_bounds__(256) mr_p_update_kernel(int64_t N, const KrylovScalars *__restrict__ ks,
                                                          const T *__restrict__ r1, T *__restrict__ p1,
                                                          const T *__restrict__ ap, T *__restrict__ r0) {
  if (ks->done) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if (ks->restart) {
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += stride) {
      r0[r] = r1[r];
      p1[r] = r1[r];
    }
    return;
  }
  const double beta = ks->beta, omega = ks->omega;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += stride)
    p1[r] = (T)((double)r1[r] + beta * ((double)p1[r] - omega * (double)ap[r]));
}

// the kernels of one iteration skip their work while a replacement is pending
template <typename T>
__global__ void __launch_bounds__(256) mr_s_update_kernel(int64_t N, const KrylovScalars *__restrict__ ks,
                                                          const T *__restrict__ r1, const T *__restrict__ ap,
                                                          T *__restrict__ s1) {
  if (ks->done || ks->replace) return;
  const double alpha = ks->alpha;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += stride)
    s1[r] = (T)((double)r1[r] - alpha * (double)ap[r]);
}

template <typename T>
__global__ void __launch_bounds__(256) mr_spmv_dot2_kernel(OpArgs a, const KrylovScalars *__restrict__ ks,
                                                           const T *__restrict__ s1, T *__restrict__ as,
                                                           double *part) {
  if (ks->done || ks->replace) return;
  double acc[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < a.N; r += stride) {
    const T y = op_apply<T>(a, r, s1);
    ntstore(y, as + r);
    acc[0] += (double)y * (double)s1[r];
    acc[1] += (double)y * (double)y;
  }
  block_partials<2>(acc, part);
}

template <typename T, typename TX>
__global__ void __launch_bounds__(256) mr_xr_update_kernel(int64_t N, const KrylovScalars *__restrict__ ks,
                                                           TX *__restrict__ xl, const T *__restrict__ p1,
                                                           const T *__restrict__ s1, const T *__restrict__ as,
                                                           T *__restrict__ r1, const T *__restrict__ r0,
                                                           double *part) {
  if (ks->done || ks->replace) return;
  const double alpha = ks->alpha, omega = ks->omega;
  double acc[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += stride) {
    const double s = (double)s1[r];
    xl[r] = (TX)((double)xl[r] + (alpha * (double)p1[r] + omega * s));
    const T rn = (T)(s - omega * (double)as[r]);
    r1[r] = rn;
    acc[0] += (double)rn * (double)rn;
    acc[1] += (double)rn * (double)r0[r];
  }
  block_partials<2>(acc, part);
}

// T = float: FP32 vectors, FP32 iterate xl folded into the FP64 x.
// T = double: FP64 vectors on the FP32-valued operator; x updated directly.
template <typename T>
int reliable_solve(Ctx &c, const double *b, double *x, double tol, double delta, int maxit, int *replacements,
                   int *iters, double *relres) {
  if (!c.ell_ready || !c.val32) return fail(CFD_ERR_STATE, "FP32 ELL operator not assembled");
  const int64_t N = c.g.N;
  const int nb = blocks_for(N, 256, RED_MAX_BLOCKS);
  const OpArgs a64 = op_args(c, fp64_op(c));
  const OpArgs ain = op_args(c, std::is_same<T, float>::value ? CFD_OP_ELL : OP_ELL_VAL32);
  T *r0, *r1, *p1, *s1, *ap, *as;
  if constexpr (std::is_same<T, float>::value) {
    r0 = c.q32[0]; r1 = c.q32[1]; p1 = c.q32[2]; s1 = c.q32[3]; ap = c.q32[4]; as = c.q32[5];
  } else {
    r0 = c.r0; r1 = c.r1; p1 = c.p1; s1 = c.s1; ap = c.ap; as = c.as;
  }
  float *xl = std::is_same<T, float>::value ? c.q32[6] : nullptr;
  KrylovScalars *ks = c.ks32;
  double *P1 = red_part_host(c, RS_SPMV_DOT), *P2 = red_part_host(c, RS_SPMV_DOT2), *P3 = red_part_host(c, RS_XR),
         *P4 = red_part_host(c, RS_NORM);
  c.solves32++;
  ks_setup_kernel<<<1, 1, 0, c.stream>>>(ks, tol, 1e-7 * tol, maxit, c.solves32);
  set_delta_kernel<<<1, 1, 0, c.stream>>>(ks, delta);
  mr_init_kernel<T><<<nb, 256, op_lds_bytes(a64), c.stream>>>(a64, b, x, r1, r0, p1, (T *)xl,
                                                               red_part_host(c, RS_INIT));
  mr_init_finalize<<<1, FIN_THREADS, 0, c.stream>>>(red_part_host(c, RS_INIT), nb, ks);
  CFD_HIP(hipGetLastError());
  int batch = 4;
  double prev_rel = -1.0;
  int prev_it = 0;
  for (;;) {
    for (int it = 0; it < batch; ++it) {
      spmv_dot_kernel<T><<<nb, 256, op_lds_bytes(ain), c.stream>>>(ain, ks, p1, ap, r0, P1);
      mr_alpha_finalize<<<1, FIN_THREADS, 0, c.stream>>>(P1, nb, ks);
      mr_s_update_kernel<T><<<nb, 256, 0, c.stream>>>(N, ks, r1, ap, s1);
      mr_spmv_dot2_kernel<T><<<nb, 256, 0, c.stream>>>(ain, ks, s1, as, P2);
      mr_omega_finalize<<<1, FIN_THREADS, 0, c.stream>>>(P2, nb, ks);
      if constexpr (std::is_same<T, float>::value)
        mr_xr_update_kernel<T, float><<<nb, 256, 0, c.stream>>>(N, ks, xl, p1, s1, as, r1, r0, P3);
      else
        mr_xr_update_kernel<T, double><<<nb, 256, 0, c.stream>>>(N, ks, x, p1, s1, as, r1, r0, P3);
      mr_conv_finalize<<<1, FIN_THREADS, 0, c.stream>>>(P3, nb, ks);
      if constexpr (std::is_same<T, float>::value) mr_fold_kernel<<<nb, 256, 0, c.stream>>>(N, ks, x, xl);
      mr_replace_kernel<T><<<nb, 256, op_lds_bytes(a64), c.stream>>>(a64, ks, b, x, r1, r0, P4);
      mr_replace_finalize<<<1, FIN_THREADS, 0, c.stream>>>(P4, nb, ks);
      mr_p_update_kernel<T><<<nb, 256, 0, c.stream>>>(N, ks, r1, p1, ap, r0);
    }
    CFD_HIP(hipGetLastError());
    CFD_HIP(hipMemcpyAsync(c.hks, ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, c.stream));
    CFD_HIP(hipStreamSynchronize(c.stream));
    const KrylovScalars h = *c.hks;
    if (c.debug)
      fprintf(stderr, "[mixed] it %d rep %d brk %d restarts %d rn/nb %.3e true %.3e alpha %.3e omega %.3e rho %.3e\n",
              h.iters, h.replacements, h.breakdown, h.restarts, h.rnorm / h.norm0, h.relres, h.alpha, h.omega, h.rho);
    if (h.done) break;
    const double cur = h.rnorm / h.norm0;
    int next = batch * 2;
    if (prev_rel > 0.0 && cur > 0.0 && cur < prev_rel && h.iters > prev_it) {
      const double rate = std::log(cur / prev_rel) / (double)(h.iters - prev_it);
      const double need = std::log(tol / cur) / rate;
      if (need > 0.0 && std::isfinite(need)) next = (int)std::ceil(need * 1.05) + 1;
    }
    prev_rel = cur;
    prev_it = h.iters;
    batch = std::max(1, std::min(next, 64));
  }
  const KrylovScalars h = *c.hks;
  *replacements = h.replacements;
  *iters = h.iters + 1;
  *relres = h.relres;
  return CFD_OK;
}

}  // namespace

int mixed_reliable_solve(Ctx &c, const double *b, double *x, double tol, double delta, int maxit, int *replacements,
                         int *iters, double *relres) {
  if (c.prm.mixed_mode == CFD_MIXED_VALUES32)
    return reliable_solve<double>(c, b, x, tol, delta, maxit, replacements, iters, relres);
  return reliable_solve<float>(c, b, x, tol, delta, maxit, replacements, iters, relres);
}

}  // namespace cfd

// ---------------------------------------------------------------------------
// cfd_bicgstab_ell_noalloc: the BiCGSTAB solve of cfd_bicgstab on a
// caller-owned ELL operator and caller-owned vectors, with every buffer the
// solve needs beyond them taken from one static device array.  It calls no
// hipMalloc / hipFree and no allocating library routine, so it can run where
// allocation is not allowed (code under stream capture, tight serving loops).
//
// The array holds, per device, the solver's seven work vectors (r0, r1, p1,
// s1, ap, as and px1, the start value kept for the NaN reset) of at most
// CFD_NOALLOC_MAX_ROWS rows, the reduction partials and results, and the
// device scalars.  The host refers to it only through the device address
// hipGetSymbolAddress returns (taken once per device, with that device
// current); the array's own name is a host-side stand-in, not an address.
// Every call on a device shares that device's array and leaves its contents
// behind: calls that overlap on one device race.  Nothing in it needs to
// start at zero: each solve writes every vector and partial before it reads
// it and sets the scalars itself.
// ---------------------------------------------------------------------------
namespace cfd {
namespace {

constexpr int64_t NOALLOC_ROWS = CFD_NOALLOC_MAX_ROWS;
constexpr int NOALLOC_VECS = 7;  // r0 r1 p1 s1 ap as px1
constexpr int64_t NOALLOC_PARTIALS = (int64_t)RED_SITES * 3 * RED_MAX_BLOCKS;
constexpr int64_t NOALLOC_OUT = (int64_t)RED_SITES * 4;
constexpr int64_t NOALLOC_KS = (sizeof(KrylovScalars) + 255) / 256 * 32;  // doubles, 256-byte multiple
constexpr int64_t NOALLOC_DOUBLES = NOALLOC_VECS * NOALLOC_ROWS + NOALLOC_PARTIALS + NOALLOC_OUT + NOALLOC_KS;
static_assert(NOALLOC_PARTIALS % 32 == 0 && NOALLOC_OUT % 32 == 0, "256-byte sections");

static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// the device address of this device's `scratch`, looked up on the first call
int scratch_address(int dev, double **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(CFD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double *>(ptr);
  }
  *out = g_scratch_ptr[dev];
  return CFD_OK;
}

}  // namespace
}  // namespace cfd

extern "C" int cfd_bicgstab_ell_noalloc(const int32_t *col, const double *val, const double *b, double *x,
                                        size_t n_rows, double tol, int maxit, void *stream, int *iters,
                                        double *relres) {
  using namespace cfd;
  if (!col || !val || !b || !x || !iters || !relres) return fail(CFD_ERR_ARG, "cfd_bicgstab_ell_noalloc: NULL argument");
  if (n_rows < 1) return fail(CFD_ERR_ARG, "cfd_bicgstab_ell_noalloc: n_rows must be >= 1");
  if ((int64_t)n_rows > NOALLOC_ROWS)
    return fail(CFD_ERR_ARG, "cfd_bicgstab_ell_noalloc: n_rows " + std::to_string(n_rows) + " exceeds the maximum " +
                                 std::to_string(NOALLOC_ROWS) + " (CFD_NOALLOC_MAX_ROWS)");
  if (maxit < 0 || !(tol >= 0.0)) return fail(CFD_ERR_ARG, "cfd_bicgstab_ell_noalloc: bad tol / maxit");
  int dev = 0;
  CFD_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(CFD_ERR_ARG, "cfd_bicgstab_ell_noalloc: device index too large");
  double *base = nullptr;
  if (int rc = scratch_address(dev, &base)) return rc;
  const int64_t N = (int64_t)n_rows;
  // a context on the stack that borrows the caller's operator and the
  // scratch sections; its host mirrors are stack variables (pageable)
  Ctx c;
  c.device = dev;
  c.stream = static_cast<hipStream_t>(stream);
  c.g.N = N;
  c.g.nxc = (int)N;
  c.g.nyc = c.g.nzc = 1;
  c.ell_sliced = 0;  // the caller's arrays are in the reference's slot-major layout
  c.col = const_cast<int32_t *>(col);
  c.val = const_cast<double *>(val);
  c.ell_ready = true;
  double *vec[NOALLOC_VECS];
  for (int q = 0; q < NOALLOC_VECS; ++q) vec[q] = base + (int64_t)q * NOALLOC_ROWS;
  double *tail = base + (int64_t)NOALLOC_VECS * NOALLOC_ROWS;
  c.red.partials = tail;
  c.red.out = tail + NOALLOC_PARTIALS;
  c.red.counters = nullptr;
  c.ks = reinterpret_cast<KrylovScalars *>(tail + NOALLOC_PARTIALS + NOALLOC_OUT);
  KrylovScalars hks{};
  double hred[64] = {};
  c.hks = &hks;
  c.hred = hred;
  // a fresh bicgstabRe2's first solve: timestep_ 1 (no restart rule), px2 zero
  Vecs<double> v{x, vec[0], vec[1], vec[2], vec[3], vec[4], vec[5], b, vec[6], nullptr};
  const int rc = solve_bicg<double>(c, op_args(c, CFD_OP_ELL), v, c.ks, 1, tol, maxit, iters, relres);
  c.stream = nullptr;  // borrowed: the stack context must not release anything
  return rc;
}
