// dev_spmv.h -- row kernels of the pressure operator, inlined into the
// SpMV, Krylov and SOR kernels.
//
// ELL rows are read at ell_base(row) + s * stride (the library's arrays in
// 64-row slices, a caller's slot-major): lane r of a wave reads element r of
// each slot, so every slot load of a wave64 is one contiguous 512 B (FP64) /
// 256 B (FP32) segment.  The matrix streams once
// per SpMV and is loaded non-temporally so that it does not evict the
// x vector, whose seven gathers per row (row, row +- 1, +- nz, +- ny*nz)
// are served from L2 / Infinity Cache.  The accumulation is an FMA chain in
// slot order, which is exactly what the reference build computes for
// r[i] += VAL * x[IDX] (inc/matrix/ELL_sparseMatrix.hpp:477-492).
#pragma once

#include <hip/hip_runtime.h>

#include "cfd_internal.h"

namespace cfd {

template <typename T>
__device__ inline T ntload(const T *p) {
  return __builtin_nontemporal_load(p);
}

// Vectors that no following SpMV gathers from (the operator's output, the
// x / r update's x and r) are stored non-temporally, which leaves the caches
// to the vector the SpMV gathers: 5% on the 512x256x64 SpMV (133.7 against
// 140.3 us), 2% on a BiCGSTAB timestep.
template <typename T>
__device__ inline void ntstore(T v, T *p) {
  __builtin_nontemporal_store(v, p);
}

template <typename TV, typename TX>
__device__ inline TX ell_row(const int32_t *__restrict__ col, const TV *__restrict__ val, int64_t N, int sliced,
                             int64_t r, const TX *__restrict__ x) {
  const int64_t b = ell_base(r, sliced), st = ell_stride(N, sliced);
  TX acc = TX(0);
#pragma unroll
  for (int s = 0; s < KSLOTS; ++s) {
    const int32_t c = ntload(col + b + s * st);
    const TX v = (TX)ntload(val + b + s * st);
    acc = fma(v, x[c], acc);
  }
  return acc;
}

// n / d for a 32-bit unsigned n by an invariant divisor d >= 1: one
// multiply-high, a subtract, an add and two shifts (Granlund & Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, sec. 4).
// m = floor(2^32 (2^l - d) / d) + 1 with l = ceil(log2 d).
struct FastDiv {
  uint32_t m;
  int s1, s2;
};

inline FastDiv fastdiv_make(uint32_t d) {
  int l = 0;
  while ((1ull << l) < d) ++l;
  FastDiv f;
  f.m = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
  f.s1 = l > 0 ? 1 : 0;
  f.s2 = l > 0 ? l - 1 : 0;
  return f;
}

__device__ inline uint32_t fastdiv(uint32_t n, const FastDiv &f) {
  const uint32_t t = __umulhi(f.m, n);
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

// SPE_matrix1::spmv_ompJP (inc/matrix/SPE_sparseMatrix1.hpp:126-144): the
// Jacobi-scaled operator applied matrix-free; coef = xm[nxc] xp[nxc]
// ym[nyc] yp[nyc] zm[nzc] zp[nzc]; neighbour terms in the order xm, ym, zm,
// diagonal, zp, yp, xp, each contracted to an FMA like the reference build.
// The row's (i, j, k) comes from two invariant divisions (N < 2^31).  Every
// load is issued before the sum (an absent neighbour reads the row itself,
// and its coefficient entry, which is in range, and is left out by a
// select), so a wave waits once per row instead of once per neighbour.
__device__ inline double spe_row(const double *__restrict__ coef, const double *__restrict__ jp, int nxc, int nyc,
                                 int nzc, const FastDiv &div_nz, const FastDiv &div_ny, int64_t ic,
                                 const double *__restrict__ x, bool wrapz = false) {
  const int64_t nynz = (int64_t)nyc * nzc;
  const uint32_t q = fastdiv((uint32_t)ic, div_nz);  // i nyc + j
  const int k = (int)((uint32_t)ic - q * (uint32_t)nzc);
  const uint32_t i_ = fastdiv(q, div_ny);
  const int j = (int)(q - i_ * (uint32_t)nyc);
  const int i = (int)i_;
  const double *xm = coef, *xp = xm + nxc, *ym = xp + nxc, *yp = ym + nyc, *zm = yp + nyc, *zp = zm + nzc;
  // periodic z: the z neighbours always exist, across the seam at +-(nzc - 1)
  const bool hxm = i != 0, hym = j != 0, hzm = wrapz || k != 0, hzp = wrapz || k != nzc - 1, hyp = j != nyc - 1,
             hxp = i != nxc - 1;
  const int64_t izm = k != 0 ? ic - 1 : ic + (nzc - 1), izp = k != nzc - 1 ? ic + 1 : ic - (nzc - 1);
  const double vxm = x[hxm ? ic - nynz : ic], vym = x[hym ? ic - nzc : ic], vzm = x[hzm ? izm : ic];
  const double vc = x[ic];
  const double vzp = x[hzp ? izp : ic], vyp = x[hyp ? ic + nzc : ic], vxp = x[hxp ? ic + nynz : ic];
  const double cxm = xm[i], cym = ym[j], czm = zm[k], czp = zp[k], cyp = yp[j], cxp = xp[i];
  const double J = jp[ic];
  double sum = 0.0;
  sum = hxm ? fma(-cxm * J, vxm, sum) : sum;
  sum = hym ? fma(-cym * J, vym, sum) : sum;
  sum = hzm ? fma(-czm * J, vzm, sum) : sum;
  sum += vc;
  sum = hzp ? fma(-czp * J, vzp, sum) : sum;
  sum = hyp ? fma(-cyp * J, vyp, sum) : sum;
  sum = hxp ? fma(-cxp * J, vxp, sum) : sum;
  return sum;
}

// CSR_matrix::multiply_omp (inc/matrix/CSR_sparseMatrix.hpp:278-292): the
// row's entries in stored (column) order, each product added as an FMA like
// the reference build; one thread per row
__device__ inline double csr_row(const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
                                 const double *__restrict__ val, int64_t r, const double *__restrict__ x) {
  double acc = 0.0;
  const int32_t e1 = ptr[r + 1];
  for (int32_t e = ptr[r]; e < e1; ++e) acc = fma(val[e], x[col[e]], acc);
  return acc;
}

// Operator handle passed by value to the kernels.
struct OpArgs {
  int op;                    // CFD_OP_ELL / CFD_OP_SPE / CFD_OP_CSR
  int64_t N;
  int sliced;                // ELL layout (ell_base)
  const int32_t *col;
  const double *val;
  const float *val32;
  const double *coef, *jp;   // SPE
  int nxc, nyc, nzc;
  FastDiv div_nz, div_ny;    // SPE: row -> (i, j, k)
  bool wrapz;                // SPE: periodic z (the z neighbours wrap across the seam)
  const int32_t *cptr, *ccol;  // CSR
  const double *cval;
};

// The SPE operator over plane columns.  A 1-D grid of 256-thread blocks:
// block b covers a tile of kb x jb (k, j) columns (kb a power of two, at most
// 64, jb = 256 / kb) over SPE_PLANES consecutive i-planes; thread t owns the
// column (k, j) of the tile and walks the planes.  The entry x at (i', j, k)
// is loaded once and serves plane i' as its diagonal and planes i' -+ 1 as
// their x-neighbours; the y and z coefficients are loaded once per thread and
// the x ones are plane-uniform.  Each row's terms are spe_row's, in its order,
// so every output is spe_row's bit for bit.
constexpr int SPE_PLANES = 8;

struct SpeTiles {
  int kb, jb, nkt, njt, nig;
  int64_t blocks;
};

inline SpeTiles spe_tiles(int nxc, int nyc, int nzc) {
  SpeTiles t;
  t.kb = 1;
  while (t.kb < nzc && t.kb < 64) t.kb <<= 1;
  t.jb = 256 / t.kb;
  t.nkt = (nzc + t.kb - 1) / t.kb;
  t.njt = (nyc + t.jb - 1) / t.jb;
  t.nig = (nxc + SPE_PLANES - 1) / SPE_PLANES;
  t.blocks = (int64_t)t.nkt * t.njt * t.nig;
  return t;
}

// this thread's column and first plane; false outside the grid
struct SpeColumn {
  int i0, j, k;
  bool ok;
};

__device__ inline SpeColumn spe_column(const OpArgs &a, const SpeTiles &T) {
  const int b = blockIdx.x;
  const int tk = b % T.nkt, rest = b / T.nkt;
  SpeColumn c;
  c.k = tk * T.kb + (int)(threadIdx.x % T.kb);
  c.j = (rest % T.njt) * T.jb + (int)(threadIdx.x / T.kb);
  c.i0 = (rest / T.njt) * SPE_PLANES;
  c.ok = c.k < a.nzc && c.j < a.nyc;
  return c;
}

// v at the column's row of each of its planes, loaded together (a plane
// past the grid loads the last plane's row, which is never used), so that a
// kernel that pairs the column's outputs with another vector waits for those
// loads once instead of once per plane
__device__ inline void spe_plane_values(const OpArgs &a, const SpeColumn &c, const double *__restrict__ v,
                                        double (&w)[SPE_PLANES]) {
  const int64_t nynz = (int64_t)a.nyc * a.nzc, col = (int64_t)c.j * a.nzc + c.k;
#pragma unroll
  for (int p = 0; p < SPE_PLANES; ++p) w[p] = v[(int64_t)min(c.i0 + p, a.nxc - 1) * nynz + col];
}

// y[p] = (A x)[row of plane i0 + p] and xc[p] = x at that row, for the
// planes inside the grid (np of them)
__device__ inline int spe_column_rows(const OpArgs &a, const SpeColumn &c, const double *__restrict__ x,
                                      double (&y)[SPE_PLANES], double (&xc)[SPE_PLANES]) {
  const int nxc = a.nxc, nyc = a.nyc, nzc = a.nzc;
  const int64_t nynz = (int64_t)nyc * nzc, col = (int64_t)c.j * nzc + c.k;
  const double *xm = a.coef, *xp = xm + nxc, *ym = xp + nxc, *yp = ym + nyc, *zm = yp + nyc, *zp = zm + nzc;
  const bool hym = c.j != 0, hzm = a.wrapz || c.k != 0, hzp = a.wrapz || c.k != nzc - 1, hyp = c.j != nyc - 1;
  const int dzm = c.k != 0 ? -1 : nzc - 1, dzp = c.k != nzc - 1 ? 1 : -(nzc - 1);  // z steps, across the seam
  const double cym = ym[c.j], cyp = yp[c.j], czm = zm[c.k], czp = zp[c.k];
  double cen[SPE_PLANES + 2];
#pragma unroll
  for (int p = 0; p < SPE_PLANES + 2; ++p) {
    const int ip = min(max(c.i0 - 1 + p, 0), nxc - 1);  // clamped planes are never used
    cen[p] = x[ip * nynz + col];
  }
  const int np = min(SPE_PLANES, nxc - c.i0);
  double vym[SPE_PLANES], vyp[SPE_PLANES], vzm[SPE_PLANES], vzp[SPE_PLANES], J[SPE_PLANES];
#pragma unroll
  for (int p = 0; p < SPE_PLANES; ++p) {
    const int64_t ic = (int64_t)min(c.i0 + p, nxc - 1) * nynz + col;
    vym[p] = x[hym ? ic - nzc : ic];
    vyp[p] = x[hyp ? ic + nzc : ic];
    vzm[p] = x[hzm ? ic + dzm : ic];
    vzp[p] = x[hzp ? ic + dzp : ic];
    J[p] = a.jp[ic];
  }
#pragma unroll
  for (int p = 0; p < SPE_PLANES; ++p) {
    const int i = c.i0 + p;
    const bool hxm = i != 0, hxp = i != nxc - 1;
    const double cxm = xm[min(i, nxc - 1)], cxp = xp[min(i, nxc - 1)];
    double sum = 0.0;
    sum = hxm ? fma(-cxm * J[p], cen[p], sum) : sum;
    sum = hym ? fma(-cym * J[p], vym[p], sum) : sum;
    sum = hzm ? fma(-czm * J[p], vzm[p], sum) : sum;
    sum += cen[p + 1];
    sum = hzp ? fma(-czp * J[p], vzp[p], sum) : sum;
    sum = hyp ? fma(-cyp * J[p], vyp[p], sum) : sum;
    sum = hxp ? fma(-cxp * J[p], cen[p + 2], sum) : sum;
    y[p] = sum;
    xc[p] = cen[p + 1];
  }
  return np;
}

// internal operator id: the ELL operator with FP32 values applied to an FP64
// vector (FP64 products and sum)
constexpr int OP_ELL_VAL32 = 16;

__device__ inline double op_row(const OpArgs &a, int64_t r, const double *__restrict__ x) {
  if (a.op == CFD_OP_SPE) return spe_row(a.coef, a.jp, a.nxc, a.nyc, a.nzc, a.div_nz, a.div_ny, r, x, a.wrapz);
  if (a.op == CFD_OP_CSR) return csr_row(a.cptr, a.ccol, a.cval, r, x);
  if (a.op == OP_ELL_VAL32) return ell_row<float, double>(a.col, a.val32, a.N, a.sliced, r, x);
  return ell_row<double, double>(a.col, a.val, a.N, a.sliced, r, x);
}

// CSR rows of one 256-row block staged through LDS.  The block's entries are
// contiguous, [ptr[base], ptr[base + 256]), so the block reads them with
// coalesced non-temporal loads and each thread then runs its row's FMA chain
// (the order of csr_row) out of LDS; a thread walking its own row straight
// from memory spreads every load instruction of a wave over ~3.5 KB.  Only
// CSR launches size the dynamic LDS (op_lds_bytes), so the other operators
// keep their occupancy.
constexpr int OP_BLOCK = 256;
constexpr int CSR_STAGE = OP_BLOCK * KSLOTS;

inline size_t op_lds_bytes(const OpArgs &a) {
  return a.op == CFD_OP_CSR ? (size_t)CSR_STAGE * (sizeof(double) + sizeof(int32_t)) : 0;
}

// (A x)[base + threadIdx.x], 0 past the last row.  base is the block's first
// row: the same in every thread of a 256-thread block, all of which call this.
__device__ inline double op_row_block(const OpArgs &a, int64_t base, const double *__restrict__ x) {
  const int64_t r = base + threadIdx.x;
  if (a.op != CFD_OP_CSR) return r < a.N ? op_row(a, r, x) : 0.0;
  extern __shared__ double csr_lds[];
  double *sval = csr_lds;
  int32_t *scol = reinterpret_cast<int32_t *>(csr_lds + CSR_STAGE);
  const int64_t rend = base + OP_BLOCK < a.N ? base + OP_BLOCK : a.N;
  const int32_t e0 = a.cptr[base];
  const int32_t cnt = a.cptr[rend] - e0;
  if (cnt > CSR_STAGE)  // more than 7 entries a row on average: not staged (block-uniform)
    return r < a.N ? csr_row(a.cptr, a.ccol, a.cval, r, x) : 0.0;
  // every slot a thread stages is loaded before the first LDS store (slots
  // past cnt load the block's last entry and are not stored), so staging
  // costs one round trip to memory instead of one per slot
  double sv[KSLOTS];
  int32_t sc[KSLOTS];
  if (cnt > 0) {  // block-uniform; every row holds its diagonal, so always taken
#pragma unroll
    for (int u = 0; u < KSLOTS; ++u) {
      const int32_t q = threadIdx.x + u * OP_BLOCK, qc = q < cnt ? q : cnt - 1;
      sv[u] = ntload(a.cval + e0 + qc);
      sc[u] = ntload(a.ccol + e0 + qc);
    }
  }
  __syncthreads();  // the previous block's rows have been read
#pragma unroll
  for (int u = 0; u < KSLOTS; ++u) {
    const int32_t q = threadIdx.x + u * OP_BLOCK;
    if (q < cnt) {
      sval[q] = sv[u];
      scol[q] = sc[u];
    }
  }
  __syncthreads();
  if (r >= a.N) return 0.0;
  const int32_t eb = a.cptr[r] - e0, e1 = a.cptr[r + 1] - e0;
  // the row's first KSLOTS entries: every LDS read and x gather is issued
  // before the first FMA (entries past the row read slot 0 and are not
  // added); the FMAs keep csr_row's order
  double vv[KSLOTS], xv[KSLOTS];
#pragma unroll
  for (int u = 0; u < KSLOTS; ++u) {
    const int32_t e = eb + u < e1 ? eb + u : 0;
    vv[u] = sval[e];
    xv[u] = x[scol[e]];
  }
  double acc = 0.0;
#pragma unroll
  for (int u = 0; u < KSLOTS; ++u)
    if (eb + u < e1) acc = fma(vv[u], xv[u], acc);
  for (int32_t e = eb + KSLOTS; e < e1; ++e) acc = fma(sval[e], x[scol[e]], acc);
  return acc;
}

// FP32 values and vector; products are exact in FP64 and summed in FP64, so
// the cancellation of the nearly row-sum-free operator does not cost FP32
// digits (a pure FP32 accumulation loses them for smooth x and stalls the
// Krylov iteration at 1e5+ unknowns).
__device__ inline float op_row32(const OpArgs &a, int64_t r, const float *__restrict__ x) {
  const int64_t b = ell_base(r, a.sliced), st = ell_stride(a.N, a.sliced);
  double acc = 0.0;
#pragma unroll
  for (int s = 0; s < KSLOTS; ++s) {
    const int32_t c = ntload(a.col + b + s * st);
    acc = fma((double)ntload(a.val32 + b + s * st), (double)x[c], acc);
  }
  return (float)acc;
}

inline OpArgs op_args(const Ctx &c, int op) {
  OpArgs a;
  a.op = op;
  a.N = c.g.N;
  a.sliced = c.ell_sliced;
  a.col = c.col;
  a.val = c.val;
  a.val32 = c.val32;
  a.coef = c.spe_coef;
  a.jp = c.spe_jp;
  a.nxc = c.g.nxc;
  a.nyc = c.g.nyc;
  a.nzc = c.g.nzc;
  a.div_nz = fastdiv_make((uint32_t)c.g.nzc);
  a.div_ny = fastdiv_make((uint32_t)c.g.nyc);
  a.wrapz = c.prm.periodic_z != 0;
  a.cptr = c.csr_ptr;
  a.ccol = c.csr_col;
  a.cval = c.csr_val;
  return a;
}

}  // namespace cfd
