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
      // the slots a thread owns are loaded sixteen at a time (clamped
      // index, slots past nblk select zero) and added in a fixed pairwise
      // order, so the sum is deterministic and the finalize waits for one
      // round trip to memory per sixteen slots instead of one per slot
      constexpr int PER = RED_MAX_BLOCKS / FIN_THREADS, CH = 16;
      static_assert(PER % CH == 0 && (CH & (CH - 1)) == 0, "slots per thread");
      const int last = nblk > 0 ? nblk - 1 : 0;
      double acc = 0.0;
#pragma unroll 1
      for (int c0 = 0; c0 < PER && c0 * FIN_THREADS < nblk; c0 += CH) {
        double s[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          const int b = threadIdx.x + (c0 + u) * FIN_THREADS;
          const double x = pq[b < nblk ? b : last];
          s[u] = b < nblk ? x : 0.0;
        }
#pragma unroll
        for (int w = 1; w < CH; w *= 2) {
#pragma unroll
          for (int u = 0; u < CH; u += 2 * w) s[u] += s[u + w];
        }
        acc += s[0];
      }
      v[q] = acc;
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
