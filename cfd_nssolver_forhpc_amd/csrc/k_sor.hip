// k_sor.hip -- red-black SOR on the ELL pressure operator.
//
// The reference's SOR (SorPipeLine_seq / _omp, inc/pressure/solver/) sweeps
// lexicographically in place, which is sequential (and racy under OpenMP).
// With the 7-point stencil a red-black colouring makes every colour's
// update independent: x_r += omega (b_r - (A x)_r) / a_rr for all rows of one
// colour, then the other.  Rows (and x, b) live in a colour-permuted copy so
// that each colour launch streams a contiguous half of the matrix; the FMA
// chain per row is in slot order, as in the SpMV.
#include <algorithm>
#include <cmath>

#include "dev_common.h"
#include "dev_spmv.h"

namespace cfd {

namespace {

template <typename TV, typename TX>
__global__ void __launch_bounds__(256) sweep_kernel(int64_t N, int sliced, int64_t off, int64_t cnt,
                                                    const int32_t *__restrict__ colP, const TV *__restrict__ valP,
                                                    const TX *__restrict__ bP, TX *__restrict__ xP, double omega) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t st = ell_stride(N, sliced);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cnt; t += stride) {
    const int64_t q = off + t, b = ell_base(q, sliced);
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s < KSLOTS; ++s) {
      const int32_t cidx = ntload(colP + b + s * st);
      acc = fma((double)ntload(valP + b + s * st), (double)xP[cidx], acc);
    }
    const double d = (double)valP[b];  // slot 0: the diagonal
    xP[q] = (TX)((double)xP[q] + omega * ((double)bP[q] - acc) / d);
  }
}

template <typename TV, typename TX>
__global__ void __launch_bounds__(256) resid_kernel(int64_t N, int sliced, const int32_t *__restrict__ colP,
                                                    const TV *__restrict__ valP, const TX *__restrict__ bP,
                                                    const TX *__restrict__ xP, double *part) {
  double acc[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t st = ell_stride(N, sliced);
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < N; q += stride) {
    const int64_t b = ell_base(q, sliced);
    double ax = 0.0;
#pragma unroll
    for (int s = 0; s < KSLOTS; ++s) {
      const int32_t cidx = ntload(colP + b + s * st);
      ax = fma((double)ntload(valP + b + s * st), (double)xP[cidx], ax);
    }
    const double r = (double)bP[q] - ax;
    acc[0] += r * r;
    acc[1] += (double)bP[q] * (double)bP[q];
  }
  block_partials<2>(acc, part);
}

template <typename TI, typename TO>
__global__ void __launch_bounds__(256) permute_in_kernel(int64_t N, const int32_t *__restrict__ perm,
                                                         const TI *__restrict__ a, TO *__restrict__ aP) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < N) aP[perm[r]] = (TO)a[r];
}

template <typename T>
__global__ void __launch_bounds__(256) permute_out_kernel(int64_t N, const int32_t *__restrict__ perm,
                                                          const T *__restrict__ aP, T *__restrict__ a) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < N) a[r] = aP[perm[r]];
}

__global__ void __launch_bounds__(FIN_THREADS) finalize2(const double *part, int nblk, double *out) {
  double tot[2];
  combine_partials<2>(part, nblk, tot);
  if (threadIdx.x == 0) {
    out[0] = tot[0];
    out[1] = tot[1];
  }
}

template <typename TV, typename TX>
int rbsor_generic(Ctx &c, const TV *valP, const TX *bP, TX *xP, double omega, double tol, int maxit, int check,
                  int *iters, double *relres) {
  const int64_t N = c.g.N;
  const int nbr = blocks_for(N, 256, RED_MAX_BLOCKS);
  if (check < 1) check = 1;
  int it = 0;
  double rel = INFINITY;
  while (it < maxit) {
    const int chunk = std::min(check, maxit - it);
    for (int s = 0; s < chunk; ++s) {
      for (int colour = 0; colour < 2; ++colour) {
        const int64_t off = colour ? c.ncolour[0] : 0, cnt = c.ncolour[colour];
        sweep_kernel<TV, TX><<<blocks_for(cnt, 256, 1 << 30), 256, 0, c.stream>>>(N, c.ell_sliced, off, cnt, c.colP,
                                                                                valP, bP, xP, omega);
      }
    }
    it += chunk;
    double *part = red_part_host(c, RS_NORM);
    resid_kernel<TV, TX><<<nbr, 256, 0, c.stream>>>(N, c.ell_sliced, c.colP, valP, bP, xP, part);
    finalize2<<<1, FIN_THREADS, 0, c.stream>>>(part, nbr, c.red.out + RS_NORM * 4);
    CFD_HIP(hipGetLastError());
    CFD_HIP(hipMemcpyAsync(c.hred, c.red.out + RS_NORM * 4, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CFD_HIP(hipStreamSynchronize(c.stream));
    rel = std::sqrt(c.hred[0]) / std::sqrt(c.hred[1]);
    if (rel < tol) break;
  }
  *iters = it;
  *relres = rel;
  return CFD_OK;
}

}  // namespace

int rbsor_solve(Ctx &c, const double *b, double *x, double omega, double tol, int maxit, int check_every, int *iters,
                double *relres) {
  if (!c.ell_ready || !c.colP) return fail(CFD_ERR_STATE, "ELL operator not assembled");
  const int64_t N = c.g.N;
  const int nb = blocks_for(N, 256, 1 << 30);
  permute_in_kernel<double, double><<<nb, 256, 0, c.stream>>>(N, c.perm, b, c.bP);
  permute_in_kernel<double, double><<<nb, 256, 0, c.stream>>>(N, c.perm, x, c.xP);
  CFD_HIP(hipGetLastError());
  int rc = rbsor_generic<double, double>(c, c.valP, c.bP, c.xP, omega, tol, maxit, check_every, iters, relres);
  if (rc != CFD_OK) return rc;
  permute_out_kernel<double><<<nb, 256, 0, c.stream>>>(N, c.perm, c.xP, x);
  CFD_HIP(hipGetLastError());
  return CFD_OK;
}

int rbsor_solve_f32(Ctx &c, const float *b, float *x, double omega, double tol, int maxit, int check_every,
                    int *iters, double *relres) {
  if (!c.ell_ready || !c.valP32) return fail(CFD_ERR_STATE, "FP32 ELL operator not assembled");
  const int64_t N = c.g.N;
  const int nb = blocks_for(N, 256, 1 << 30);
  permute_in_kernel<float, float><<<nb, 256, 0, c.stream>>>(N, c.perm, b, c.bP32);
  permute_in_kernel<float, float><<<nb, 256, 0, c.stream>>>(N, c.perm, x, c.xP32);
  CFD_HIP(hipGetLastError());
  int rc = rbsor_generic<float, float>(c, c.valP32, c.bP32, c.xP32, omega, tol, maxit, check_every, iters, relres);
  if (rc != CFD_OK) return rc;
  permute_out_kernel<float><<<nb, 256, 0, c.stream>>>(N, c.perm, c.xP32, x);
  CFD_HIP(hipGetLastError());
  return CFD_OK;
}

hipError_t launch_rbsor_sweep(Ctx &c, double omega) {
  const int64_t N = c.g.N;
  for (int colour = 0; colour < 2; ++colour) {
    const int64_t off = colour ? c.ncolour[0] : 0, cnt = c.ncolour[colour];
    sweep_kernel<double, double><<<blocks_for(cnt, 256, 1 << 30), 256, 0, c.stream>>>(N, c.ell_sliced, off, cnt,
                                                                                    c.colP, c.valP, c.bP, c.xP, omega);
  }
  return hipGetLastError();
}

}  // namespace cfd

// ---------------------------------------------------------------------------
// SorPipeLine (inc/pressure/solver/SorPipeLine.seq.hpp:4-122): the
// reference's own SOR, in its volume-weighted form on the ghosted pressure,
// in place.  Per interior cell, with the face coefficients of the seq
// version (cf0..cf5 = face area / centre distance, cf7 = -1 / grid::coef,
// cf6 = mChange / dt from the predicted velocity):
//   pNEW = (-p[i+1] cf0 - p[i-1] cf1 - p[j+1] cf2 - p[j-1] cf3
//           - p[k+1] cf4 - p[k-1] cf5 + cf6) cf7
//   p += omega (pNEW - p)
// boundary neighbours read the ghost pressure as the last BC left it.  The
// sweep stops when max |pNEW - p| over the sweep is at most p_criteria
// (SOR_ERROR_MAXVAL, SorPipeLine.top.hpp:7) or after itmax sweeps, and
// reports sweeps - 1 and that maximum (simu.iters, simu.error).
//
// The reference sweeps lexicographically; here the cells go in red-black
// order ((i + j + k) even first), each colour one launch in which every
// cell reads only the other colour, so the sweep is race-free and the same
// on every run.  It converges to the same fixed point.  cf0..cf7 are not
// stored (the reference keeps 8 doubles a cell): each sweep forms them from
// the 1-D metrics in the reference's operation order, with contraction off,
// so every cell's update is the oracle's red-black restatement bit for bit.
// ---------------------------------------------------------------------------
namespace cfd {
namespace {

__global__ void __launch_bounds__(256) sorp_rhs_kernel(DevGrid g, double rdt, const double *__restrict__ u1,
                                                       const double *__restrict__ v1, const double *__restrict__ w1,
                                                       double *__restrict__ cf6) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const int64_t nyz = (int64_t)g.nyc * g.nzc;
  const int i = (int)(t / nyz) + GC, j = (int)((t / g.nzc) % g.nyc) + GC, k = (int)(t % g.nzc) + GC;
  const int64_t ic = icel(g, i, j, k), sx = (int64_t)g.ny * g.nz, sy = g.nz;
  const double mChange = (u1[ic] - u1[ic - sx]) * g.Dy[j] * g.Dz[k] + (v1[ic] - v1[ic - sy]) * g.Dx[i] * g.Dz[k] +
                         (w1[ic] - w1[ic - 1]) * g.Dx[i] * g.Dy[j];
  cf6[t] = mChange * rdt;
}

__global__ void sorp_setup_kernel(KrylovScalars *ks, double crit, int itmax) {
  ks->tol = crit;
  ks->maxit = itmax;
  ks->iters = 0;
  ks->done = 0;
  ks->relres = 0.0;
}

// one colour of a sweep; partial max |pNEW - p| per block
// wrapz: periodic z, the z neighbours of the first and last planes are the
// live pressure across the seam (not ghosts), which keeps the red-black
// colouring valid for an even nz
__global__ void __launch_bounds__(256) sorp_sweep_kernel(DevGrid g, int colour, double omega, int wrapz,
                                                         const double *__restrict__ cf6, double *__restrict__ p,
                                                         const KrylovScalars *__restrict__ ks, double *part) {
#pragma clang fp contract(off)
  if (ks->done) return;
  double mx[1] = {0.0};
  const int nzh = (g.nzc + 1) / 2;
  const int64_t cnt = (int64_t)g.nxc * g.nyc * nzh;
  const int64_t sx = (int64_t)g.ny * g.nz, sy = g.nz;
  const double *__restrict__ Dx = g.Dx, *__restrict__ Dy = g.Dy, *__restrict__ Dz = g.Dz;
  const double *__restrict__ Dxs = g.Dxs, *__restrict__ Dys = g.Dys, *__restrict__ Dzs = g.Dzs;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cnt; t += stride) {
    const int kh = (int)(t % nzh);
    const int64_t rest = t / nzh;
    const int jc = (int)(rest % g.nyc), ic_ = (int)(rest / g.nyc);
    const int kc = 2 * kh + ((ic_ + jc + colour) & 1);
    if (kc >= g.nzc) continue;
    const int i = ic_ + GC, j = jc + GC, k = kc + GC;
    const int64_t ic = icel(g, i, j, k);
    const double cf0 = Dy[j] * Dz[k] / Dxs[i], cf1 = Dy[j] * Dz[k] / Dxs[i - 1];
    const double cf2 = Dx[i] * Dz[k] / Dys[j], cf3 = Dx[i] * Dz[k] / Dys[j - 1];
    const double cf4 = Dx[i] * Dy[j] / Dzs[k], cf5 = Dx[i] * Dy[j] / Dzs[k - 1];
    // grid::coef (inc/grid/structureGrid.hpp:278-287), summed in its order
    const double coef = Dy[j] * Dz[k] / Dxs[i] + Dy[j] * Dz[k] / Dxs[i - 1] + Dx[i] * Dz[k] / Dys[j] +
                        Dx[i] * Dz[k] / Dys[j - 1] + Dx[i] * Dy[j] / Dzs[k] + Dx[i] * Dy[j] / Dzs[k - 1];
    const double cf7 = -1.0 / coef;
    const int64_t izp = (wrapz && k == g.nz - 3) ? ic - (g.nzc - 1) : ic + 1;
    const int64_t izm = (wrapz && k == GC) ? ic + (g.nzc - 1) : ic - 1;
    const double pxp = p[ic + sx], pxm = p[ic - sx], pyp = p[ic + sy], pym = p[ic - sy], pzp = p[izp],
                 pzm = p[izm], pc = p[ic];
    const int64_t row = ((int64_t)ic_ * g.nyc + jc) * g.nzc + kc;
    const double pNEW = (-pxp * cf0 - pxm * cf1 - pyp * cf2 - pym * cf3 - pzp * cf4 - pzm * cf5 + cf6[row]) * cf7;
    mx[0] = fmax(mx[0], fabs(pNEW - pc));
    p[ic] = pc + omega * (pNEW - pc);
  }
  block_partials<1, true>(mx, part);
}

// the sweep's maximum over both colours and the do-while test of SorPipeLine_seq
__global__ void __launch_bounds__(FIN_THREADS) sorp_finalize_kernel(const double *part0, const double *part1,
                                                                    int nblk, KrylovScalars *ks) {
  if (ks->done) return;
  double a[1], b[1];
  combine_partials<1, true>(part0, nblk, a);
  combine_partials<1, true>(part1, nblk, b);
  if (threadIdx.x == 0) {
    const double residual = fmax(a[0], b[0]);
    ks->relres = residual;
    ks->iters += 1;
    if (!(residual > ks->tol) || ks->iters >= ks->maxit) ks->done = 1;
  }
}

}  // namespace

int sor_pipeline_solve(Ctx &c, double omega, double crit, int itmax, int *iters, double *error) {
  const DevGrid &g = c.g;
  if (itmax < 1) return fail(CFD_ERR_ARG, "SorPipeLine: p_sor_iter_max must be >= 1");
  double *cf6 = c.f[CFD_RHS];
  sorp_rhs_kernel<<<blocks_for(g.N, 256, 1 << 30), 256, 0, c.stream>>>(g, 1.0 / c.prm.dt, c.f[CFD_T1_U],
                                                                      c.f[CFD_T1_V], c.f[CFD_T1_W], cf6);
  sorp_setup_kernel<<<1, 1, 0, c.stream>>>(c.ks, crit, itmax);
  CFD_HIP(hipGetLastError());
  const int64_t cnt = (int64_t)g.nxc * g.nyc * ((g.nzc + 1) / 2);
  const int nb = blocks_for(cnt, 256, RED_MAX_BLOCKS);
  double *part0 = red_part_host(c, RS_SPMV_DOT), *part1 = red_part_host(c, RS_SPMV_DOT2);
  int batch = 4, done_sweeps = 0;
  double prev = -1.0;
  int prev_it = 0;
  for (;;) {
    for (int s = 0; s < batch; ++s) {
      sorp_sweep_kernel<<<nb, 256, 0, c.stream>>>(g, 0, omega, c.prm.periodic_z, cf6, c.f[CFD_P], c.ks, part0);
      sorp_sweep_kernel<<<nb, 256, 0, c.stream>>>(g, 1, omega, c.prm.periodic_z, cf6, c.f[CFD_P], c.ks, part1);
      sorp_finalize_kernel<<<1, FIN_THREADS, 0, c.stream>>>(part0, part1, nb, c.ks);
    }
    CFD_HIP(hipGetLastError());
    CFD_HIP(hipMemcpyAsync(c.hks, c.ks, sizeof(KrylovScalars), hipMemcpyDeviceToHost, c.stream));
    CFD_HIP(hipStreamSynchronize(c.stream));
    const KrylovScalars h = *c.hks;
    done_sweeps = h.iters;
    if (h.done) break;
    // the sweeps still needed, from the last batch's contraction rate
    int next = batch * 2;
    if (prev > 0.0 && h.relres > 0.0 && h.relres < prev && h.iters > prev_it) {
      const double rate = std::log(h.relres / prev) / (double)(h.iters - prev_it);
      const double need = std::log(crit / h.relres) / rate;
      if (need > 0.0 && std::isfinite(need)) next = (int)std::ceil(need * 1.05) + 1;
    }
    prev = h.relres;
    prev_it = h.iters;
    batch = std::max(1, std::min(next, std::min(512, itmax - h.iters)));
  }
  *iters = done_sweeps - 1;  // SorPipeLine_seq's --simu.iters
  *error = c.hks->relres;
  return CFD_OK;
}

}  // namespace cfd

// ---------------------------------------------------------------------------
// Red-black SOR on the CSR operator (grid::createPressureMatrixCSR, the
// operator of P_SOLVER_BICG_CSR): rows in natural order, the cells of one
// colour per launch, x_r += omega (b_r - (A x)_r) / a_rr with the row's
// products summed in stored order (csr_row) and a_rr the entry whose column
// is r.  Same colouring and update as the ELL form.
// ---------------------------------------------------------------------------
namespace cfd {
namespace {

__global__ void __launch_bounds__(256) csr_sweep_kernel(int nxc, int nyc, int nzc, int colour,
                                                        const int32_t *__restrict__ ptr,
                                                        const int32_t *__restrict__ col,
                                                        const double *__restrict__ val, const double *__restrict__ b,
                                                        double *__restrict__ x, double omega) {
  const int nzh = (nzc + 1) / 2;
  const int64_t cnt = (int64_t)nxc * nyc * nzh;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cnt; t += stride) {
    const int kh = (int)(t % nzh);
    const int64_t rest = t / nzh;
    const int jc = (int)(rest % nyc), ic = (int)(rest / nyc);
    const int kc = 2 * kh + ((ic + jc + colour) & 1);
    if (kc >= nzc) continue;
    const int64_t r = ((int64_t)ic * nyc + jc) * nzc + kc;
    double acc = 0.0, diag = 0.0;
    const int32_t e1 = ptr[r + 1];
    for (int32_t e = ptr[r]; e < e1; ++e) {
      const int32_t cc = col[e];
      const double v = val[e];
      acc = fma(v, x[cc], acc);
      if (cc == r) diag = v;
    }
    x[r] = x[r] + omega * (b[r] - acc) / diag;
  }
}

__global__ void __launch_bounds__(256) csr_resid_kernel(int64_t N, const int32_t *__restrict__ ptr,
                                                        const int32_t *__restrict__ col,
                                                        const double *__restrict__ val, const double *__restrict__ b,
                                                        const double *__restrict__ x, double *part) {
  double acc[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < N; r += stride) {
    const double res = b[r] - csr_row(ptr, col, val, r, x);
    acc[0] += res * res;
    acc[1] += b[r] * b[r];
  }
  block_partials<2>(acc, part);
}

}  // namespace

int rbsor_csr_solve(Ctx &c, const double *b, double *x, double omega, double tol, int maxit, int check_every,
                    int *iters, double *relres) {
  if (!c.csr_ready) return fail(CFD_ERR_STATE, "CSR operator not assembled");
  const DevGrid &g = c.g;
  const int64_t N = g.N;
  const int64_t cnt = (int64_t)g.nxc * g.nyc * ((g.nzc + 1) / 2);
  const int nbs = blocks_for(cnt, 256, 1 << 30);
  const int nbr = blocks_for(N, 256, RED_MAX_BLOCKS);
  const int check = check_every < 1 ? 1 : check_every;
  int it = 0;
  double rel = INFINITY;
  while (it < maxit) {
    const int chunk = std::min(check, maxit - it);
    for (int s = 0; s < chunk; ++s)
      for (int colour = 0; colour < 2; ++colour)
        csr_sweep_kernel<<<nbs, 256, 0, c.stream>>>(g.nxc, g.nyc, g.nzc, colour, c.csr_ptr, c.csr_col, c.csr_val, b,
                                                    x, omega);
    it += chunk;
    double *part = red_part_host(c, RS_NORM);
    csr_resid_kernel<<<nbr, 256, 0, c.stream>>>(N, c.csr_ptr, c.csr_col, c.csr_val, b, x, part);
    finalize2<<<1, FIN_THREADS, 0, c.stream>>>(part, nbr, c.red.out + RS_NORM * 4);
    CFD_HIP(hipGetLastError());
    CFD_HIP(hipMemcpyAsync(c.hred, c.red.out + RS_NORM * 4, 2 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CFD_HIP(hipStreamSynchronize(c.stream));
    rel = std::sqrt(c.hred[0]) / std::sqrt(c.hred[1]);
    if (rel < tol) break;
  }
  *iters = it;
  *relres = rel;
  return CFD_OK;
}

}  // namespace cfd
