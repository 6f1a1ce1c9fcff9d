// k_sor.hip -- red-black SOR on the ELL pressure operator.
//
// The reference's SOR (SorPipeLine_seq / _omp, inc/pressure/solver/) sweeps
// lexicographically in place, which is sequential (and racy under OpenMP).
// With the 7-point stencil a red-black colouring makes every colour's
// update independent: x_r += omega (b_r - (A x)_r) / a_rr for all rows of one
// colour, then the other.  Rows (and x, b) live in a colour-permuted copy so
// that each colour launch streams a contiguous half of the matrix; the FMA
// chain per row is in slot order, as in the SpMV.
#include <cmath>

#include "dev_common.h"
#include "dev_spmv.h"

namespace cfd {

namespace {

template <typename TV, typename TX>
__global__ void __launch_bounds__(256) sweep_kernel(int64_t N, int64_t off, int64_t cnt,
                                                    const int32_t *__restrict__ colP, const TV *__restrict__ valP,
                                                    const TX *__restrict__ bP, TX *__restrict__ xP, double omega) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cnt; t += stride) {
    const int64_t q = off + t;
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s < KSLOTS; ++s) {
      const int32_t cidx = ntload(colP + (int64_t)s * N + q);
      acc = fma((double)ntload(valP + (int64_t)s * N + q), (double)xP[cidx], acc);
    }
    const double d = (double)valP[q];
    xP[q] = (TX)((double)xP[q] + omega * ((double)bP[q] - acc) / d);
  }
}

template <typename TV, typename TX>
__global__ void __launch_bounds__(256) resid_kernel(int64_t N, const int32_t *__restrict__ colP,
                                                    const TV *__restrict__ valP, const TX *__restrict__ bP,
                                                    const TX *__restrict__ xP, double *part) {
  double acc[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < N; q += stride) {
    double ax = 0.0;
#pragma unroll
    for (int s = 0; s < KSLOTS; ++s) {
      const int32_t cidx = ntload(colP + (int64_t)s * N + q);
      ax = fma((double)ntload(valP + (int64_t)s * N + q), (double)xP[cidx], ax);
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
        sweep_kernel<TV, TX><<<blocks_for(cnt, 256, 1 << 30), 256, 0, c.stream>>>(N, off, cnt, c.colP, valP, bP, xP,
                                                                                omega);
      }
    }
    it += chunk;
    double *part = red_part_host(c, RS_NORM);
    resid_kernel<TV, TX><<<nbr, 256, 0, c.stream>>>(N, c.colP, valP, bP, xP, part);
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
    sweep_kernel<double, double><<<blocks_for(cnt, 256, 1 << 30), 256, 0, c.stream>>>(N, off, cnt, c.colP, c.valP,
                                                                                    c.bP, c.xP, omega);
  }
  return hipGetLastError();
}

}  // namespace cfd
