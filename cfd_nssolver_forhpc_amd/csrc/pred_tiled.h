// pred_tiled.h -- the LDS-tiled QUICK convection-diffusion predictor
// (ConvectionDifussion with QuickUnitA / QuickUnitB,
// inc/source/convection_and_difussion.hpp:20-84,168-446).
//
// The cell-by-cell form divides every face gradient (f[n+1] - f[n]) / D and
// every geometric QUICK weight 0.125 D D / Ds again in each cell that reads
// it: 144 FP64 divisions per cell (seven instructions each), which bound the
// kernel.  Here a workgroup owns a tj x TK patch of (j, k) columns and marches
// over PT_TI consecutive i-planes:
//   - the (tj+4) x (TK+4) slabs of the plane's u, v and w sit in LDS; the
//     next plane's slabs are fetched into registers before the plane's face
//     gradients are divided, so their loads overlap that work;
//   - the y and z face gradients of the plane's three fields are divided
//     once per face into LDS, the x face gradients once per face in each
//     thread's registers (its own column's values ride along as the march
//     advances);
//   - the QUICK weights, the interpolation factor IT, the axis widths and
//     the face divisors are tabulated in LDS at the start, per plane, row and
//     column of the workgroup.
// A cell then costs about 30 divisions.  Every quotient, weight and sum is
// the cell-by-cell form's, operand for operand.
//
// The kernel is a sequence of phases separated by workgroup barriers; each
// phase is a function of the thread index, so the same code runs one thread
// at a time on the host.
#pragma once

#include "cfd_internal.h"

namespace cfd {

constexpr int PT_THREADS = 256;
constexpr int PT_TI = 16;  // i-planes a workgroup marches over
constexpr int PT_NV = 11;  // tabulated values per axis position

struct PtShape {
  int tj, tk;          // patch extents, tj * tk = PT_THREADS
  int njt, nkt, nit;   // patches along j and k, plane ranges along i
};

// The patch geometry for depth TK: every index split of the phases divides
// by these compile-time constants (a multiply and a shift, not the
// ~35-instruction sequence of a run-time integer division)
template <int TK>
struct PtGeom {
  static constexpr int tk = TK, tj = PT_THREADS / TK;
  static constexpr int SJ = tj + 4, SK = TK + 4, SL = SJ * SK;
  static constexpr int NPT = (SL + PT_THREADS - 1) / PT_THREADS;  // slab cells a thread fetches
  static constexpr int GY = (tj + 3) * TK, GZ = tj * (TK + 3), GF = GY + GZ;
  static constexpr int NTAB = PT_NV * (PT_TI + tj + TK) + 2 * (PT_TI + 3 + tj + 3 + TK + 3);
  static constexpr size_t doubles = 6 * (size_t)SL + 3 * (size_t)GF + (size_t)NTAB;
};
static_assert(PtGeom<4>::tj * 4 == PT_THREADS && PtGeom<64>::tj * 64 == PT_THREADS, "patch shapes");

// patch depth tk: the power of two in [4, 64] that wastes the fewest lanes
// and halo cells per column row (ceil(nzc / tk) patches of tk + 4 slab cells)
inline PtShape pt_shape(const DevGrid &g) {
  int best = 64;
  int64_t cost_best = -1;
  for (int tk = 64; tk >= 4; tk >>= 1) {
    const int64_t cost = (int64_t)((g.nzc + tk - 1) / tk) * (tk + 4);
    if (cost_best < 0 || cost < cost_best) {
      cost_best = cost;
      best = tk;
    }
  }
  PtShape s;
  s.tk = best;
  s.tj = PT_THREADS / best;
  s.nkt = (g.nzc + s.tk - 1) / s.tk;
  s.njt = (g.nyc + s.tj - 1) / s.tj;
  s.nit = (g.nxc + PT_TI - 1) / PT_TI;
  return s;
}

inline size_t pt_lds_doubles(const PtShape &s) {
  switch (s.tk) {
    case 4: return PtGeom<4>::doubles;
    case 8: return PtGeom<8>::doubles;
    case 16: return PtGeom<16>::doubles;
    case 32: return PtGeom<32>::doubles;
    default: return PtGeom<64>::doubles;
  }
}

// Workgroup LDS, in doubles:
//   slab[6][SL]      u, v, w (slots 2 f and 2 f + 1 by plane parity), each
//                    (tj+4) x (TK+4): the patch with 2 cells of halo
//   grad[3][GF]      per field: y faces (tj+3 rows j0-2..j0+tj, TK columns),
//                    then z faces (tj rows, TK+3 faces k0-2..k0+TK)
//   cx[PT_NV][PT_TI], cy[PT_NV][tj], cz[PT_NV][TK]   per axis position: the
//                    QUICK weights, IT, D and Ds (pt_axis_value)
//   dx[2][PT_TI+3], dy[2][tj+3], dz[2][TK+3]   per face n from the block's
//                    first face (ib-2, j0-2, k0-2): the A divisor D[n+1],
//                    the B divisor Ds[n]
struct PtLds {
  double *slab, *grad, *cx, *cy, *cz, *dx, *dy, *dz;
};

template <int TK>
__host__ __device__ inline PtLds pt_lds(double *base) {
  using G = PtGeom<TK>;
  PtLds L;
  L.slab = base;
  L.grad = base + 6 * G::SL;
  L.cx = L.grad + 3 * G::GF;
  L.cy = L.cx + PT_NV * PT_TI;
  L.cz = L.cy + PT_NV * G::tj;
  L.dx = L.cz + PT_NV * TK;
  L.dy = L.dx + 2 * (PT_TI + 3);
  L.dz = L.dy + 2 * (G::tj + 3);
  return L;
}

// the patch and plane range of workgroup b (k patches fastest, then j, then i)
struct PtBlock {
  int j0, k0, ib, ie;
};

constexpr int PT_XCDS = 8;  // MI355X deals workgroup ids round-robin to its 8 XCDs, each with its own L2

// Workgroup id b runs on XCD b % 8.  The patch it takes is numbered so that
// every XCD gets one contiguous run of patches (XCD x the x-th eighth): the
// patches beside each other in j, whose slabs share their halo rows, then
// meet in the same L2 instead of fetching those rows once per XCD.
__host__ __device__ inline int pt_xcd_order(int b, int nblk) {
  const int q = nblk / PT_XCDS, r = nblk % PT_XCDS, x = b % PT_XCDS;
  return x * q + (x < r ? x : r) + b / PT_XCDS;
}

__host__ __device__ inline PtBlock pt_block(const DevGrid &g, const PtShape &s, int b) {
  b = pt_xcd_order(b, s.njt * s.nkt * s.nit);
  PtBlock B;
  const int kt = b % s.nkt;
  b /= s.nkt;
  const int jt = b % s.njt, it = b / s.njt;
  B.j0 = GC + jt * s.tj;
  B.k0 = GC + kt * s.tk;
  B.ib = GC + it * PT_TI;
  B.ie = B.ib + PT_TI < g.nx - GC ? B.ib + PT_TI : g.nx - GC;
  return B;
}

// Tabulated values of axis position I (D: cell widths, Ds: staggered
// distances), the weights in the cell-by-cell form's operand order:
//   0-3 QuickUnitA (the component's own axis): the P face for U > 0 and
//       U <= 0, the N face for U > 0 and U <= 0;
//   4-7 QuickUnitB (across it), the same four;
//   8   IT = (Ds[I] * 0.5) / D[I], the cross-velocity interpolation factor;
//   9   D[I];  10  Ds[I].
__host__ __device__ inline double pt_axis_value(int var, const double *D, const double *Ds, int I) {
  switch (var) {
    case 0: return 0.125 * D[I + 1] * D[I + 1] / Ds[I];
    case 1: return 0.125 * D[I + 1] * D[I + 1] / Ds[I + 1];
    case 2: return 0.125 * D[I] * D[I] / Ds[I - 1];
    case 3: return 0.125 * D[I] * D[I] / Ds[I];
    case 4: return 0.125 * Ds[I] * Ds[I] / D[I];
    case 5: return 0.125 * Ds[I] * Ds[I] / D[I + 1];
    case 6: return 0.125 * Ds[I - 1] * Ds[I - 1] / D[I - 1];
    case 7: return 0.125 * Ds[I - 1] * Ds[I - 1] / D[I];
    case 8: return (Ds[I] * 0.5) / D[I];
    case 9: return D[I];
    default: return Ds[I];
  }
}

// one axis' tables: values of positions first..first+n-1 (those before
// `end`), divisors of faces first-2..first+n (those whose far cell is before
// `len`, the ghosted extent; 1 elsewhere, where no active cell reads)
__host__ __device__ inline void pt_axis_tables(const double *D, const double *Ds, int first, int n, int end, int len,
                                               double *val, double *div, int tid) {
  for (int q = tid; q < PT_NV * n + 2 * (n + 3); q += PT_THREADS) {
    if (q < PT_NV * n) {
      const int var = q / n, I = first + q % n;
      val[q] = I < end ? pt_axis_value(var, D, Ds, I) : 0.0;
    } else {
      const int r = q - PT_NV * n;
      const bool A = r < n + 3;
      const int face = first - 2 + (A ? r : r - (n + 3));
      div[r] = face + 1 < len ? (A ? D[face + 1] : Ds[face]) : 1.0;
    }
  }
}

// phase 0: the tables of the block's planes, rows and columns
template <int TK>
__host__ __device__ inline void pt_tables(const DevGrid &g, const PtBlock &B, const PtLds &L, int tid) {
  pt_axis_tables(g.Dx, g.Dxs, B.ib, PT_TI, B.ie, g.nx, L.cx, L.dx, tid);
  pt_axis_tables(g.Dy, g.Dys, B.j0, PtGeom<TK>::tj, g.ny - GC, g.ny, L.cy, L.dy, tid);
  pt_axis_tables(g.Dz, g.Dzs, B.k0, TK, g.nz - GC, g.nz, L.cz, L.dz, tid);
}

// slot of a field's slab for plane p: u 0/1, v 2/3, w 4/5 by the parity of p
template <int TK>
__host__ __device__ inline double *pt_slot(const PtLds &L, int field, int p) {
  return L.slab + (size_t)(2 * field + (p & 1)) * PtGeom<TK>::SL;
}

// plane p's slabs of u, v, w into registers (cells past the grid's ghost
// layers read as 0 and reach no active cell) ...
template <int TK>
__host__ __device__ inline void pt_fetch(const DevGrid &g, const PtBlock &B, const double *__restrict__ u,
                                         const double *__restrict__ v, const double *__restrict__ w, int p, int tid,
                                         double (&r)[3][PtGeom<TK>::NPT]) {
  using G = PtGeom<TK>;
  const double *F[3] = {u, v, w};
  const int64_t base = icel(g, p, B.j0 - 2, B.k0 - 2);
  for (int m = 0; m < G::NPT; ++m) {
    const int q = tid + m * PT_THREADS;
    const int row = q / G::SK, cz = q % G::SK;
    const bool ok = q < G::SL && B.j0 - 2 + row < g.ny && B.k0 - 2 + cz < g.nz;
    const int64_t at = base + (int64_t)row * g.nz + cz;
    for (int f = 0; f < 3; ++f) r[f][m] = ok ? F[f][at] : 0.0;
  }
}

// ... and from the registers into the plane's slots
template <int TK>
__host__ __device__ inline void pt_store(const PtLds &L, int p, int tid, const double (&r)[3][PtGeom<TK>::NPT]) {
  using G = PtGeom<TK>;
  for (int m = 0; m < G::NPT; ++m) {
    const int q = tid + m * PT_THREADS;
    if (q < G::SL)
      for (int f = 0; f < 3; ++f) pt_slot<TK>(L, f, p)[q] = r[f][m];
  }
}

// phase: the y and z face gradients G[n] = (f[n+1] - f[n]) / divisor of plane
// p's u, v, w: A faces (the component's own axis: v in y, w in z) divide by
// D[n+1], B faces (u in y and z, v in z, w in y) by Ds[n]; faces whose far
// cell lies past the ghost layers are 0 and reach no active cell
template <int TK>
__host__ __device__ inline void pt_grads(const DevGrid &g, const PtBlock &B, const PtLds &L, int p, int tid) {
  using G = PtGeom<TK>;
  for (int field = 0; field < 3; ++field) {
    const double *S = pt_slot<TK>(L, field, p);
    double *Gf = L.grad + field * G::GF;
    const double *dyv = L.dy + (field == 1 ? 0 : G::tj + 3), *dzv = L.dz + (field == 2 ? 0 : TK + 3);
    for (int r = tid; r < G::GY; r += PT_THREADS) {  // y face j0 - 2 + row, between rows row and row + 1
      const int row = r / TK, col = r % TK;
      const bool ok = B.j0 - 1 + row < g.ny && B.k0 + col < g.nz;
      const double f0 = S[row * G::SK + col + 2], f1 = S[(row + 1) * G::SK + col + 2];
      Gf[r] = ok ? (f1 - f0) / dyv[row] : 0.0;
    }
    for (int r = tid; r < G::GZ; r += PT_THREADS) {  // z face k0 - 2 + fz, between columns fz and fz + 1
      const int row = r / (TK + 3), fz = r % (TK + 3);
      const bool ok = B.k0 - 1 + fz < g.nz && B.j0 + row < g.ny;
      const double f0 = S[(row + 2) * G::SK + fz], f1 = S[(row + 2) * G::SK + fz + 1];
      Gf[G::GY + r] = ok ? (f1 - f0) / dzv[fz] : 0.0;
    }
  }
}

// A thread's own column: u, v, w at planes i-1..i+2 and the x face
// gradients at faces i-2..i+1 (u: A, divisor Dx[n+1]; v, w: B, Dxs[n])
struct PtCol {
  double f[3][4];
  double gx[3][4];
};

__host__ __device__ inline void pt_col_init(const DevGrid &g, const double *__restrict__ u, const double *__restrict__ v,
                                            const double *__restrict__ w, int i, int j, int k, PtCol &c) {
  const double *F[3] = {u, v, w};
  double fm2[3];
  for (int f = 0; f < 3; ++f) {
    fm2[f] = F[f][icel(g, i - 2, j, k)];
    for (int m = 0; m < 4; ++m) c.f[f][m] = F[f][icel(g, i - 1 + m, j, k)];
  }
  for (int f = 0; f < 3; ++f) {
    c.gx[f][0] = (c.f[f][0] - fm2[f]) / (f == 0 ? g.Dx[i - 1] : g.Dxs[i - 2]);
    for (int m = 1; m < 4; ++m) {
      const int n = i - 2 + m;
      c.gx[f][m] = (c.f[f][m] - c.f[f][m - 1]) / (f == 0 ? g.Dx[n + 1] : g.Dxs[n]);
    }
  }
}

// the column's plane i + 3 values, fetched before the plane's cells ...
__host__ __device__ inline void pt_col_fetch(const DevGrid &g, const double *__restrict__ u,
                                             const double *__restrict__ v, const double *__restrict__ w, int i, int j,
                                             int k, double (&nxt)[3]) {
  const int64_t at = icel(g, i + 3, j, k);
  nxt[0] = u[at];
  nxt[1] = v[at];
  nxt[2] = w[at];
}

// ... and the advance from plane i to i + 1 (new x face i + 2)
__host__ __device__ inline void pt_col_shift(const PtBlock &B, const PtLds &L, int i, const double (&nxt)[3],
                                             PtCol &c) {
  const int fx = i + 2 - (B.ib - 2);  // face i + 2 in the block's x divisor tables
  for (int f = 0; f < 3; ++f) {
    for (int m = 0; m < 3; ++m) {
      c.f[f][m] = c.f[f][m + 1];
      c.gx[f][m] = c.gx[f][m + 1];
    }
    c.f[f][3] = nxt[f];
    c.gx[f][3] = (c.f[f][3] - c.f[f][2]) / (f == 0 ? L.dx[fx] : L.dx[PT_TI + 3 + fx]);
  }
}

// QuickUnitA / QuickUnitB from the face gradients at faces I-2, I-1, I, I+1
// and the four weights (P face for U > 0, U <= 0; N face for U > 0, U <= 0)
__host__ __device__ inline void pt_quick(double Up, double Un, double fc, double fp, double fn, double gm2, double gm1,
                                         double g0, double g1, const double *c, int stride, double &phiP,
                                         double &phiN) {
  if (Up > 0.0)
    phiP = 0.5 * (fc + fp) - c[0] * (g0 - gm1);
  else
    phiP = 0.5 * (fc + fp) - c[stride] * (g1 - g0);
  if (Un > 0.0)
    phiN = 0.5 * (fn + fc) - c[2 * stride] * (gm1 - gm2);
  else
    phiN = 0.5 * (fn + fc) - c[3 * stride] * (g0 - gm1);
}

struct PtArgs {
  double nu, dt;
  const double *u, *v, *w;
  double *u1, *v1, *w1;
  const double *src;
  int src_mode;
  double *pu, *pv, *pw;
  const double *vis;
  int kw_end;
};

// phase: the three components at (i, j, k) of the thread's column
// (:196-269 u, :278-351 v, :360-434 w).  The four values one plane off
// and one cell aside (v and w of plane i+1 for u, u of plane i-1 for v and
// w) come from global memory; everything else from LDS and the column.
template <int TK>
__host__ __device__ inline void pt_cell(const DevGrid &g, const PtBlock &B, const PtLds &L, const PtArgs &a,
                                        const PtCol &c, int i, int tid) {
  using G = PtGeom<TK>;
  constexpr int SK = G::SK, tk = TK, tj = G::tj;
  const int lk = tid % TK, lj = tid / TK;
  const int j = B.j0 + lj, k = B.k0 + lk;
  if (j >= g.ny - GC || k >= g.nz - GC) return;
  const int p = i - B.ib;
  const int64_t ic = icel(g, i, j, k), t = icelcal(g, i, j, k), N = g.N;
  const int64_t sx = (int64_t)g.ny * g.nz, sy = g.nz;
  const double v_xp_ym = a.v[ic + sx - sy], w_xp_zm = a.w[ic + sx - 1];
  const double u_xm_yp = a.u[ic - sx + sy], u_xm_zp = a.u[ic - sx + 1];
  const int at = (lj + 2) * SK + lk + 2;  // the cell in a slab
  const double *u0 = pt_slot<TK>(L, 0, i), *v0 = pt_slot<TK>(L, 1, i), *w0 = pt_slot<TK>(L, 2, i);
  const int gyat = (lj + 2) * tk + lk, gzat = G::GY + lj * (tk + 3) + lk + 2;  // face j, face k of the cell
  const double *Gu = L.grad, *Gv = L.grad + G::GF, *Gw = L.grad + 2 * G::GF;
  const double *cx = L.cx + p, *cy = L.cy + lj, *cz = L.cz + lk;
  const double Dx = cx[9 * PT_TI], Dxs = cx[10 * PT_TI], Dy = cy[9 * tj], Dys = cy[10 * tj];
  const double Dz = cz[9 * tk], Dzs = cz[10 * tk];
  const double nudt = a.vis ? (a.nu + a.vis[t]) * a.dt : a.nu * a.dt;
  const double dt = a.dt;
  const double uc = c.f[0][1], vc = c.f[1][1], wc = c.f[2][1];
  const double u_xm = c.f[0][0];

  if (i < g.nx - GC - 1) {  // u
    const double IT = cx[8 * PT_TI];
    const double Up = 0.5 * (c.f[0][2] + uc);
    const double Un = 0.5 * (u_xm + uc);
    const double Vp = vc + (c.f[1][2] - vc) * IT;
    const double Vn = v0[at - SK] + (v_xp_ym - v0[at - SK]) * IT;
    const double Wp = wc + (c.f[2][2] - wc) * IT;
    const double Wn = w0[at - 1] + (w_xp_zm - w0[at - 1]) * IT;
    const double *gx = c.gx[0];
    double uPx, uNx, uPy, uNy, uPz, uNz;
    pt_quick(Up, Un, uc, c.f[0][2], u_xm, gx[0], gx[1], gx[2], gx[3], cx, PT_TI, uPx, uNx);
    pt_quick(Vp, Vn, uc, u0[at + SK], u0[at - SK], Gu[gyat - 2 * tk], Gu[gyat - tk], Gu[gyat], Gu[gyat + tk],
             cy + 4 * tj, tj, uPy, uNy);
    pt_quick(Wp, Wn, uc, u0[at + 1], u0[at - 1], Gu[gzat - 2], Gu[gzat - 1], Gu[gzat], Gu[gzat + 1], cz + 4 * tk, tk,
             uPz, uNz);
    const double conv = -dt * ((uPx * uPx - uNx * uNx) / Dxs + (uPy * Vp - uNy * Vn) / Dy + (uPz * Wp - uNz * Wn) / Dz);
    const double diff =
        nudt * ((gx[2] - gx[1]) / Dxs + (Gu[gyat] - Gu[gyat - tk]) / Dy + (Gu[gzat] - Gu[gzat - 1]) / Dz);
    double r = uc + conv + diff;
    if (a.pu) a.pu[ic] = r;
    if (a.src_mode == 1) r += a.src[t];
    else if (a.src_mode == 2 && ic < N) r += a.src[ic];
    a.u1[ic] = r;
  }
  if (j < g.ny - GC - 1) {  // v
    const double IT = cy[8 * tj];
    const double Vp = 0.5 * (v0[at + SK] + vc);
    const double Vn = 0.5 * (v0[at - SK] + vc);
    const double Up = uc + (u0[at + SK] - uc) * IT;
    const double Un = u_xm + (u_xm_yp - u_xm) * IT;
    const double Wp = wc + (w0[at + SK] - wc) * IT;
    const double Wn = w0[at - 1] + (w0[at + SK - 1] - w0[at - 1]) * IT;
    const double *gx = c.gx[1];
    double vPx, vNx, vPy, vNy, vPz, vNz;
    pt_quick(Up, Un, vc, c.f[1][2], c.f[1][0], gx[0], gx[1], gx[2], gx[3], cx + 4 * PT_TI, PT_TI, vPx, vNx);
    pt_quick(Vp, Vn, vc, v0[at + SK], v0[at - SK], Gv[gyat - 2 * tk], Gv[gyat - tk], Gv[gyat], Gv[gyat + tk], cy, tj,
             vPy, vNy);
    pt_quick(Wp, Wn, vc, v0[at + 1], v0[at - 1], Gv[gzat - 2], Gv[gzat - 1], Gv[gzat], Gv[gzat + 1], cz + 4 * tk, tk,
             vPz, vNz);
    const double conv = -dt * ((Up * vPx - Un * vNx) / Dx + (vPy * vPy - vNy * vNy) / Dys + (Wp * vPz - Wn * vNz) / Dz);
    const double diff =
        nudt * ((gx[2] - gx[1]) / Dx + (Gv[gyat] - Gv[gyat - tk]) / Dys + (Gv[gzat] - Gv[gzat - 1]) / Dz);
    double r = vc + conv + diff;
    if (a.pv) a.pv[ic] = r;
    if (a.src_mode == 1) r += a.src[N + t];
    else if (a.src_mode == 2 && ic < N) r += a.src[N + ic];
    a.v1[ic] = r;
  }
  if (k < a.kw_end) {  // w; with periodic z its last face too
    const double IT = cz[8 * tk];
    const double Wp = 0.5 * (w0[at + 1] + wc);
    const double Wn = 0.5 * (w0[at - 1] + wc);
    const double Up = uc + (u0[at + 1] - uc) * IT;
    const double Un = u_xm + (u_xm_zp - u_xm) * IT;
    const double Vp = vc + (v0[at + 1] - vc) * IT;
    const double Vn = v0[at - SK] + (v0[at - SK + 1] - v0[at - SK]) * IT;
    const double *gx = c.gx[2];
    double wPx, wNx, wPy, wNy, wPz, wNz;
    pt_quick(Up, Un, wc, c.f[2][2], c.f[2][0], gx[0], gx[1], gx[2], gx[3], cx + 4 * PT_TI, PT_TI, wPx, wNx);
    pt_quick(Vp, Vn, wc, w0[at + SK], w0[at - SK], Gw[gyat - 2 * tk], Gw[gyat - tk], Gw[gyat], Gw[gyat + tk],
             cy + 4 * tj, tj, wPy, wNy);
    pt_quick(Wp, Wn, wc, w0[at + 1], w0[at - 1], Gw[gzat - 2], Gw[gzat - 1], Gw[gzat], Gw[gzat + 1], cz, tk, wPz,
             wNz);
    const double conv = -dt * ((Up * wPx - Un * wNx) / Dx + (Vp * wPy - Vn * wNy) / Dy + (wPz * wPz - wNz * wNz) / Dzs);
    const double diff =
        nudt * ((gx[2] - gx[1]) / Dx + (Gw[gyat] - Gw[gyat - tk]) / Dy + (Gw[gzat] - Gw[gzat - 1]) / Dzs);
    double r = wc + conv + diff;
    if (a.pw) a.pw[ic] = r;
    if (a.src_mode == 1) r += a.src[2 * N + t];
    else if (a.src_mode == 2 && ic < N) r += a.src[2 * N + ic];
    a.w1[ic] = r;
  }
}

// the thread's column lies on the grid (j, k interior)
template <int TK>
__host__ __device__ inline bool pt_active(const DevGrid &g, const PtBlock &B, int tid) {
  return B.j0 + tid / TK < g.ny - GC && B.k0 + tid % TK < g.nz - GC;
}

}  // namespace cfd
