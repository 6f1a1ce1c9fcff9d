// k_ell.hip -- pressure-operator assembly and standalone SpMV kernels.
#include "dev_common.h"
#include "dev_spmv.h"

namespace cfd {

namespace {

// createPressureMatrix (inc/matrix/PressureMatrix.hpp:5-117) one row per
// thread, written straight into the slot-major layout: slot 0 diagonal
// val*jp, then -xm*jp, -xp*jp, -ym*jp, -yp*jp, -zm*jp, -zp*jp for the
// neighbours that exist, then padding (N-1, 0) -- the slot order
// mat::ELL_matrix::set leaves (ELL_sparseMatrix.hpp:146-160, 246-320).
// Divisions and the 9-term diagonal sum are evaluated in the reference's
// order, so every stored value is bit-identical to the reference's.
struct PBC {
  int pd[6];
  int periodic_z;  // z faces wrap: no closure term, the wrapped neighbour appended after zp
};

__global__ void __launch_bounds__(256) ell_assemble_kernel(DevGrid g, PBC b, int sliced, int32_t *__restrict__ col,
                                                           double *__restrict__ val, float *__restrict__ val32,
                                                           double *__restrict__ jp) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t N = g.N;
  if (r >= N) return;
  const int64_t nynz = (int64_t)g.nyc * g.nzc;
  const int i = (int)(r / nynz) + GC;
  const int j = (int)((r / g.nzc) % g.nyc) + GC;
  const int k = (int)(r % g.nzc) + GC;
  const int xBeg = GC, yBeg = GC, zBeg = GC, xEnd = g.nx - GC - 1, yEnd = g.ny - GC - 1, zEnd = g.nz - GC - 1;
  const double xp = 1.0 / g.Dx[i] / g.Dxs[i], yp = 1.0 / g.Dy[j] / g.Dys[j], zp = 1.0 / g.Dz[k] / g.Dzs[k];
  const double xm = 1.0 / g.Dx[i] / g.Dxs[i - 1], ym = 1.0 / g.Dy[j] / g.Dys[j - 1],
               zm = 1.0 / g.Dz[k] / g.Dzs[k - 1];
  double aa = 0.0, bb = 0.0, cc = 0.0;
  if (i == xBeg && !b.pd[0]) aa = -xm;
  if (i == xEnd && !b.pd[1]) aa = -xp;
  if (j == yBeg && !b.pd[2]) bb = -ym;
  if (j == yEnd && !b.pd[3]) bb = -yp;
  if (k == zBeg && !b.pd[4] && !b.periodic_z) cc = -zm;
  if (k == zEnd && !b.pd[5] && !b.periodic_z) cc = -zp;
  double vd = xm + xp;
  vd = vd + ym;
  vd = vd + yp;
  vd = vd + zm;
  vd = vd + zp;
  vd = vd + aa;
  vd = vd + bb;
  vd = vd + cc;
  const double jr = 1. / vd;
  jp[r] = jr;
  int32_t cs[KSLOTS];
  double vs[KSLOTS];
  int s = 0;
  cs[s] = (int32_t)r; vs[s] = vd * jr; ++s;
  if (i != xBeg) { cs[s] = (int32_t)(r - nynz); vs[s] = -xm * jr; ++s; }
  if (i != xEnd) { cs[s] = (int32_t)(r + nynz); vs[s] = -xp * jr; ++s; }
  if (j != yBeg) { cs[s] = (int32_t)(r - g.nzc); vs[s] = -ym * jr; ++s; }
  if (j != yEnd) { cs[s] = (int32_t)(r + g.nzc); vs[s] = -yp * jr; ++s; }
  if (k != zBeg) { cs[s] = (int32_t)(r - 1); vs[s] = -zm * jr; ++s; }
  if (k != zEnd) { cs[s] = (int32_t)(r + 1); vs[s] = -zp * jr; ++s; }
  // periodic z: the neighbour across the seam, after the regular entries
  // (where the reference's periodic branch sets it, PressureMatrix.hpp:101-111,
  // with the wrapped cell's own index)
  if (b.periodic_z && k == zBeg) { cs[s] = (int32_t)(r + (g.nzc - 1)); vs[s] = -zm * jr; ++s; }
  if (b.periodic_z && k == zEnd) { cs[s] = (int32_t)(r - (g.nzc - 1)); vs[s] = -zp * jr; ++s; }
  for (; s < KSLOTS; ++s) { cs[s] = (int32_t)(N - 1); vs[s] = 0.0; }
  const int64_t eb = ell_base(r, sliced), est = ell_stride(N, sliced);
#pragma unroll
  for (int q = 0; q < KSLOTS; ++q) {
    col[eb + q * est] = cs[q];
    val[eb + q * est] = vs[q];
    if (val32) val32[eb + q * est] = (float)vs[q];
  }
}

// Colour-ordered copy of the ELL matrix for red-black SOR: row r moves to
// perm[r] (colour-0 rows first, natural order within a colour) and its
// columns are renumbered the same way, so each colour sweep streams a
// contiguous half of the matrix.
__global__ void __launch_bounds__(256) ell_permute_kernel(int64_t N, int sliced, const int32_t *__restrict__ perm,
                                                          const int32_t *__restrict__ col,
                                                          const double *__restrict__ val, int32_t *__restrict__ colP,
                                                          double *__restrict__ valP, float *__restrict__ valP32) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= N) return;
  const int64_t q = perm[r];
  const int64_t br = ell_base(r, sliced), bq = ell_base(q, sliced), st = ell_stride(N, sliced);
#pragma unroll
  for (int s = 0; s < KSLOTS; ++s) {
    colP[bq + s * st] = perm[col[br + s * st]];
    const double v = val[br + s * st];
    valP[bq + s * st] = v;
    if (valP32) valP32[bq + s * st] = (float)v;
  }
}

// grid::createPressureMatrixCSR (inc/grid/structureGrid.hpp:630-728) one row
// per thread at ptr[row] (row pointers counted on the host): entries in
// column order xm, ym, zm, diag, zp, yp, xp, the diagonal summed as
// xm + ym + zm + xp + yp + zp + aa + bb + cc (the CSR builder's order, not
// the ELL one), jp = 1/diag.  The builder closes every face with the Neumann
// term whatever the pressure BC table says (:658-666), and so does this.
__global__ void __launch_bounds__(256) csr_assemble_kernel(DevGrid g, const int32_t *__restrict__ ptr,
                                                           int32_t *__restrict__ col, double *__restrict__ val,
                                                           double *__restrict__ jp) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= g.N) return;
  const int64_t nynz = (int64_t)g.nyc * g.nzc;
  const int i = (int)(r / nynz), j = (int)((r / g.nzc) % g.nyc), k = (int)(r % g.nzc);
  const int I = i + GC, J = j + GC, K = k + GC;
  const double xp = 1.0 / g.Dx[I] / g.Dxs[I], yp = 1.0 / g.Dy[J] / g.Dys[J], zp = 1.0 / g.Dz[K] / g.Dzs[K];
  const double xm = 1.0 / g.Dx[I] / g.Dxs[I - 1], ym = 1.0 / g.Dy[J] / g.Dys[J - 1],
               zm = 1.0 / g.Dz[K] / g.Dzs[K - 1];
  double aa = 0.0, bb = 0.0, cc = 0.0;
  if (i == 0) aa = -xm;
  if (i == g.nxc - 1) aa = -xp;
  if (j == 0) bb = -ym;
  if (j == g.nyc - 1) bb = -yp;
  if (k == 0) cc = -zm;
  if (k == g.nzc - 1) cc = -zp;
  double ijk = xm + ym;
  ijk = ijk + zm;
  ijk = ijk + xp;
  ijk = ijk + yp;
  ijk = ijk + zp;
  ijk = ijk + aa;
  ijk = ijk + bb;
  ijk = ijk + cc;
  const double jr = 1.0 / ijk;
  jp[r] = jr;
  int32_t e = ptr[r];
  if (i != 0) { col[e] = (int32_t)(r - nynz); val[e++] = -xm * jr; }
  if (j != 0) { col[e] = (int32_t)(r - g.nzc); val[e++] = -ym * jr; }
  if (k != 0) { col[e] = (int32_t)(r - 1); val[e++] = -zm * jr; }
  col[e] = (int32_t)r; val[e++] = ijk * jr;
  if (k != g.nzc - 1) { col[e] = (int32_t)(r + 1); val[e++] = -zp * jr; }
  if (j != g.nyc - 1) { col[e] = (int32_t)(r + g.nzc); val[e++] = -yp * jr; }
  if (i != g.nxc - 1) { col[e] = (int32_t)(r + nynz); val[e++] = -xp * jr; }
}

// SPE_matrix1::setupPressure / set_Jp_pc diagonal (SPE_sparseMatrix1.hpp:78-121)
__global__ void __launch_bounds__(256) spe_jp_kernel(int nxc, int nyc, int nzc, PBC b, const double *__restrict__ coef,
                                                     double *__restrict__ jp) {
  const int64_t ic = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nynz = (int64_t)nyc * nzc;
  if (ic >= nynz * nxc) return;
  const int i = (int)(ic / nynz), j = (int)((ic / nzc) % nyc), k = (int)(ic % nzc);
  const double *xm = coef, *xp = xm + nxc, *ym = xp + nxc, *yp = ym + nyc, *zm = yp + nyc, *zp = zm + nzc;
  double ijk = xm[i] + ym[j];
  ijk = ijk + zm[k];
  ijk = ijk + xp[i];
  ijk = ijk + yp[j];
  ijk = ijk + zp[k];
  if (i == 0 && !b.pd[0]) ijk -= xm[i];
  if (i == nxc - 1 && !b.pd[1]) ijk -= xp[i];
  if (j == 0 && !b.pd[2]) ijk -= ym[j];
  if (j == nyc - 1 && !b.pd[3]) ijk -= yp[j];
  if (k == 0 && !b.pd[4] && !b.periodic_z) ijk -= zm[k];
  if (k == nzc - 1 && !b.pd[5] && !b.periodic_z) ijk -= zp[k];
  jp[ic] = 1.0 / ijk;
}

__global__ void __launch_bounds__(256) spmv_kernel(OpArgs a, const double *__restrict__ x, double *__restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < a.N; base += stride) {
    const double v = op_row_block(a, base, x);
    if (base + threadIdx.x < a.N) ntstore(v, y + base + threadIdx.x);
  }
}

// the SPE operator over plane columns (spe_column_rows): the rows of
// spmv_kernel's SPE path bit for bit, with a third of its loads
__global__ void __launch_bounds__(256) spe_spmv_kernel(OpArgs a, SpeTiles T, const double *__restrict__ x,
                                                       double *__restrict__ y) {
  const SpeColumn c = spe_column(a, T);
  if (!c.ok) return;
  double v[SPE_PLANES], xc[SPE_PLANES];
  const int np = spe_column_rows(a, c, x, v, xc);
  const int64_t nynz = (int64_t)a.nyc * a.nzc, col = (int64_t)c.j * a.nzc + c.k;
#pragma unroll
  for (int p = 0; p < SPE_PLANES; ++p)
    if (p < np) ntstore(v[p], y + (int64_t)(c.i0 + p) * nynz + col);
}

// FP32 values and vectors, FP32 accumulation (the mixed-precision inner operator)
__global__ void __launch_bounds__(256) spmv32_kernel(OpArgs a, const float *__restrict__ x, float *__restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < a.N; r += stride) y[r] = op_row32(a, r, x);
}

// FP32 matrix applied to an FP64 vector: x rounded to FP32 on load, result widened
__global__ void __launch_bounds__(256) spmv32_from64_kernel(OpArgs a, const double *__restrict__ x,
                                                            double *__restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < a.N; r += stride) {
    const int64_t b = ell_base(r, a.sliced), st = ell_stride(a.N, a.sliced);
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s < KSLOTS; ++s) {
      const int32_t c = ntload(a.col + b + s * st);
      acc = fma((double)ntload(a.val32 + b + s * st), (double)(float)x[c], acc);
    }
    y[r] = acc;
  }
}

__global__ void __launch_bounds__(256) cast32_kernel(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (float)a[t];
}

PBC pbc_of(const Ctx &c) {
  PBC b;
  for (int q = 0; q < 6; ++q) b.pd[q] = c.prm.p_dirichlet[q];
  b.periodic_z = c.prm.periodic_z;
  return b;
}

}  // namespace

// grid size of the streaming kernels: one row per thread.  Measured at
// 8.4M rows, a one-row-per-thread ELL SpMV streams at 6.0 TB/s where a
// grid-stride loop over 4096-16384 blocks reaches 5.4-5.7 TB/s.
int stream_blocks(int64_t n) { return blocks_for(n, 256, 1 << 30); }

hipError_t launch_ell_assemble(Ctx &c) {
  ell_assemble_kernel<<<blocks_for(c.g.N, 256, 1 << 30), 256, 0, c.stream>>>(c.g, pbc_of(c), c.ell_sliced, c.col,
                                                                             c.val, c.val32, c.f[CFD_JP]);
  return hipGetLastError();
}

hipError_t launch_ell_permute(Ctx &c) {
  ell_permute_kernel<<<blocks_for(c.g.N, 256, 1 << 30), 256, 0, c.stream>>>(c.g.N, c.ell_sliced, c.perm, c.col,
                                                                            c.val, c.colP, c.valP, c.valP32);
  return hipGetLastError();
}

hipError_t launch_cast32(Ctx &c, const double *a, float *b, int64_t n) {
  cast32_kernel<<<blocks_for(n, 256, 1 << 30), 256, 0, c.stream>>>(a, b, n);
  return hipGetLastError();
}

hipError_t launch_csr_assemble(Ctx &c) {
  csr_assemble_kernel<<<blocks_for(c.g.N, 256, 1 << 30), 256, 0, c.stream>>>(c.g, c.csr_ptr, c.csr_col,
                                                                             c.csr_val, c.csr_jp);
  return hipGetLastError();
}

hipError_t launch_spe_assemble(Ctx &c) {
  spe_jp_kernel<<<blocks_for(c.g.N, 256, 1 << 30), 256, 0, c.stream>>>(c.g.nxc, c.g.nyc, c.g.nzc, pbc_of(c),
                                                                       c.spe_coef, c.spe_jp);
  return hipGetLastError();
}

hipError_t launch_spmv(Ctx &c, int op, int precision, const double *x, double *y) {
  const OpArgs a = op_args(c, op);
  const SpeTiles T = spe_tiles(c.g.nxc, c.g.nyc, c.g.nzc);
  if (precision == CFD_MIXED && op == CFD_OP_ELL)
    spmv32_from64_kernel<<<stream_blocks(c.g.N), 256, 0, c.stream>>>(a, x, y);
  else if (op == CFD_OP_SPE && T.blocks <= (1 << 30))
    spe_spmv_kernel<<<(unsigned)T.blocks, 256, 0, c.stream>>>(a, T, x, y);
  else
    spmv_kernel<<<stream_blocks(c.g.N), 256, op_lds_bytes(a), c.stream>>>(a, x, y);
  return hipGetLastError();
}

hipError_t launch_spmv_ell_f32(Ctx &c, const float *x, float *y) {
  const OpArgs a = op_args(c, CFD_OP_ELL);
  spmv32_kernel<<<stream_blocks(c.g.N), 256, 0, c.stream>>>(a, x, y);
  return hipGetLastError();
}

}  // namespace cfd
