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

// One DPP lane move of a double (its two 32-bit halves); CTRL is the DPP
// control word.
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the 64 lanes of a wave, returned to every lane; all 64 lanes
// must be active.  Lane moves within and across the rows of 16 lanes (the
// DPP sequence of AMD's GCN cross-lane reduction: quad permutes, row
// rotates by 4 and 8, row broadcasts of lanes 15 and 31) leave the total in
// lane 63, which is read back with readlane; the adds run on the VALU with
// no LDS round trip per step (a shuffle is a ds_bpermute), in a fixed order.
__device__ inline double wave_sum(double v) {
  v += dpp_move<0xb1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4e>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x124>(v);  // row_ror:4
  v += dpp_move<0x128>(v);  // row_ror:8
  v += dpp_move<0x142>(v);  // row_bcast:15
  v += dpp_move<0x143>(v);  // row_bcast:31
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63), hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

// Sum (or max) NV per-thread values over the block; the block's result is
// written by thread 0 to part[q * RED_MAX_BLOCKS + blockIdx.x].  Block size
// must be BT.
template <int NV, bool MAX = false, int BT = RED_THREADS>
__device__ inline void block_partials(double (&v)[NV], double *part) {
  __shared__ double lds[NV][BT / 64];
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
      for (int w = 1; w < BT / 64; ++w) s = MAX ? fmax(s, lds[q][w]) : s + lds[q][w];
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
