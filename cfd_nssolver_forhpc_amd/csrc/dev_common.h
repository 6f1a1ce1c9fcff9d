// dev_common.h -- wave64 / workgroup reduction helpers shared by the kernels.
//
// Reductions are two-stage and deterministic: every block of a streaming
// kernel leaves its partial (wave shuffle tree, then LDS across the block's
// waves) in a fixed slot; a one-block finalize kernel combines the slots in
// a fixed order.  The kernel boundary between the two stages is what makes
// the partials of all XCDs visible -- no device-wide fences inside a kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "cfd_internal.h"

namespace cfd {

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// Sum (or max) NV per-thread values over the block; the block's result is
// written by thread 0 to part[q * RED_MAX_BLOCKS + blockIdx.x].  Block size
// must be RED_THREADS.
template <int NV, bool MAX = false>
__device__ inline void block_partials(double (&v)[NV], double *part) {
  __shared__ double lds[NV][RED_THREADS / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = MAX ? wave_max(v[q]) : wave_sum(v[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) lds[q][wid] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = lds[q][0];
      for (int w = 1; w < RED_THREADS / 64; ++w) s = MAX ? fmax(s, lds[q][w]) : s + lds[q][w];
      part[q * RED_MAX_BLOCKS + blockIdx.x] = s;
    }
  }
}

// Combine nblk partials per value in a fixed order; called by all
// FIN_THREADS threads of a single block; the totals are returned to every
// thread.
template <int NV, bool MAX = false>
__device__ inline void combine_partials(const double *part, int nblk, double (&tot)[NV]) {
  __shared__ double lds[NV][FIN_THREADS / 64];
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double *pq = part + q * RED_MAX_BLOCKS;
    if (MAX) {
      double s = 0.0;
      bool first = true;
      for (int b = threadIdx.x; b < nblk; b += FIN_THREADS) {
        const double x = pq[b];
        s = first ? x : fmax(s, x);
        first = false;
      }
      v[q] = s;
    } else {
      // eight independent accumulators keep eight loads in flight per
      // thread; the combination order is fixed, so the sum is deterministic
      double s[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      int b = threadIdx.x;
      for (; b + 7 * FIN_THREADS < nblk; b += 8 * FIN_THREADS) {
#pragma unroll
        for (int u = 0; u < 8; ++u) s[u] += pq[b + u * FIN_THREADS];
      }
      for (; b < nblk; b += FIN_THREADS) s[0] += pq[b];
      v[q] = ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = MAX ? wave_max(v[q]) : wave_sum(v[q]);
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) lds[q][wid] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = lds[q][0];
    for (int w = 1; w < FIN_THREADS / 64; ++w) s = MAX ? fmax(s, lds[q][w]) : s + lds[q][w];
    tot[q] = s;
  }
  __syncthreads();
}

__device__ inline double *red_part(const RedWork &r, int site) {
  return r.partials + (size_t)site * 3 * RED_MAX_BLOCKS;
}

}  // namespace cfd
