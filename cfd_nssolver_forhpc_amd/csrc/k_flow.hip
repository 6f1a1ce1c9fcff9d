// k_flow.hip -- gfx950 kernels of the velocity side of the timestep:
// convection-diffusion predictor (QUICK, LUD or UD face values), source term,
// pressure RHS, DFIB forcing/correction, boundary conditions, IB mask,
// initial state.
//
// All fields use the reference's ghosted layout icel = (i*ny + j)*nz + k
// (k fastest), so consecutive lanes of a wave64 touch consecutive k and every
// stencil read of a wave is a contiguous (coalesced) segment; i/j neighbours
// are reused through L2.  These kernels are HBM-bound (0.3-2 flop/B), so
// they are written for one pass over memory per phase, with the phases of a
// pass fused where the data dependence allows.
#include "dev_common.h"
#include "pred_tiled.h"

namespace cfd {

namespace {

struct Cell {
  int i, j, k;
};

// interior flat index t in [0, N) -> ghosted (i, j, k)
__device__ inline Cell decode_interior(const DevGrid &g, int64_t t) {
  Cell c;
  const int64_t nzc = g.nzc, nyc = g.nyc;
  const int64_t jk = t % (nyc * nzc);
  c.i = (int)(t / (nyc * nzc)) + GC;
  c.j = (int)(jk / nzc) + GC;
  c.k = (int)(jk % nzc) + GC;
  return c;
}

// LudUnitA / LudUnitB (inc/source/convection_and_difussion.hpp:90-144; the
// two bodies are the same): the upwind cell plus half its width ratio times
// its one-sided difference
__device__ inline void lud(double Up, double Un, int64_t pp, int64_t p, int64_t ic, int64_t n, int64_t nn,
                           const double *__restrict__ Di, const double *__restrict__ Ds, int I,
                           const double *__restrict__ phi, double &phiP, double &phiN) {
  const double fc = phi[ic], fp = phi[p], fn = phi[n];
  if (Up > 0.0) phiP = fc + 0.5 * Ds[I] * (fc - fn) / Di[I];
  else phiP = fp + 0.5 * Ds[I + 1] * (fp - phi[pp]) / Di[I + 1];
  if (Un > 0.0) phiN = fn + 0.5 * Ds[I - 1] * (fn - phi[nn]) / Di[I - 1];
  else phiN = fc + 0.5 * Ds[I] * (fc - fp) / Di[I + 1];
}

// UdUnit (:148-165): the upwind cell's value
__device__ inline void ud(double Up, double Un, int64_t p, int64_t ic, int64_t n, const double *__restrict__ phi,
                          double &phiP, double &phiN) {
  phiP = Up > 0.0 ? phi[ic] : phi[p];
  phiN = Un > 0.0 ? phi[n] : phi[ic];
}

// the face-value unit of scheme S (LUD or UD: one unit for the component's
// own direction and across it; QUICK runs in quick_predictor_kernel)
template <int S, bool OWN>
__device__ inline void face_unit(double Up, double Un, int64_t pp, int64_t p, int64_t ic, int64_t n, int64_t nn,
                                 const double *__restrict__ Di, const double *__restrict__ Ds, int I,
                                 const double *__restrict__ phi, double &phiP, double &phiN) {
  static_assert(S == CFD_CONV_LUD || S == CFD_CONV_UD, "QUICK has its own kernel");
  if constexpr (S == CFD_CONV_LUD) lud(Up, Un, pp, p, ic, n, nn, Di, Ds, I, phi, phiP, phiN);
  else ud(Up, Un, p, ic, n, phi, phiP, phiN);
}

// SmagorinskyModel (inc/source/smagorinskyModel.hpp:30-79) with
// GradientGauss (:6-27): central differences phi[-] - phi[+] over the
// metrics (Dxs, Dy, Dz) for u, (Dx, Dys, Dz) for v, (Dx, Dy, Dzs) for w,
// |2 ux^2 + vy^2 + wz^2 + (uy + vx)^2 + (uz + wx)^2 + (vz + wy)^2| (the
// reference's grouping), nu_t = (C_s (Dx Dy Dz)^(1/3))^2 sqrt(.), C_s 0.18
__global__ void __launch_bounds__(256) smagorinsky_kernel(DevGrid g, const double *__restrict__ u,
                                                          const double *__restrict__ v,
                                                          const double *__restrict__ w, double *__restrict__ vis) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const Cell c = decode_interior(g, t);
  const int i = c.i, j = c.j, k = c.k;
  const int64_t ic = icel(g, i, j, k);
  const int64_t sx = (int64_t)g.ny * g.nz, sy = g.nz;
  const double ux = 0.5 / g.Dxs[i] * (u[ic - sx] - u[ic + sx]), uy = 0.5 / g.Dy[j] * (u[ic - sy] - u[ic + sy]),
               uz = 0.5 / g.Dz[k] * (u[ic - 1] - u[ic + 1]);
  const double vx = 0.5 / g.Dx[i] * (v[ic - sx] - v[ic + sx]), vy = 0.5 / g.Dys[j] * (v[ic - sy] - v[ic + sy]),
               vz = 0.5 / g.Dz[k] * (v[ic - 1] - v[ic + 1]);
  const double wx = 0.5 / g.Dx[i] * (w[ic - sx] - w[ic + sx]), wy = 0.5 / g.Dy[j] * (w[ic - sy] - w[ic + sy]),
               wz = 0.5 / g.Dzs[k] * (w[ic - 1] - w[ic + 1]);
  const double s = fabs(2.0 * (ux * ux) + vy * vy + wz * wz + (uy + vx) * (uy + vx) + (uz + wx) * (uz + wx) +
                        (vz + wy) * (vz + wy));
  const double cd = 0.18 * pow(g.Dx[i] * g.Dy[j] * g.Dz[k], 1.0 / 3.0);
  vis[t] = (cd * cd) * sqrt(s);
}

// The LUD / UD predictor, one cell a thread.
// src_mode: 0 none, 1 add src[icelCal] at the cell, 2 reference index
// mapping (src[m] added at ghosted m = icel when m < N); S: CFD_CONV_LUD or
// CFD_CONV_UD; vis: the eddy viscosity added to nu
// (convection_and_difussion.hpp:246-250), or null
template <int S>
__global__ void __launch_bounds__(256) predictor_kernel(DevGrid g, double nu, double dt, const double *__restrict__ u,
                                                        const double *__restrict__ v, const double *__restrict__ w,
                                                        double *__restrict__ u1, double *__restrict__ v1,
                                                        double *__restrict__ w1, const double *__restrict__ src,
                                                        int src_mode, double *__restrict__ pu,
                                                        double *__restrict__ pv, double *__restrict__ pw,
                                                        const double *__restrict__ vis, int kw_end) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const Cell c = decode_interior(g, t);
  const int i = c.i, j = c.j, k = c.k;
  const int64_t ic = icel(g, i, j, k);
  const int64_t sx = (int64_t)g.ny * g.nz, sy = g.nz, sz = 1;
  const int64_t xm = ic - sx, xp = ic + sx, ym = ic - sy, yp = ic + sy, zm = ic - sz, zp = ic + sz;
  const int64_t xmm = ic - 2 * sx, xpp = ic + 2 * sx, ymm = ic - 2 * sy, ypp = ic + 2 * sy, zmm = ic - 2, zpp = ic + 2;
  const double *__restrict__ Dx = g.Dx, *__restrict__ Dy = g.Dy, *__restrict__ Dz = g.Dz;
  const double *__restrict__ Dxs = g.Dxs, *__restrict__ Dys = g.Dys, *__restrict__ Dzs = g.Dzs;
  const double nudt = vis ? (nu + vis[t]) * dt : nu * dt;
  const int64_t N = g.N;

  if (i < g.nx - GC - 1) {  // u (:196-269)
    const double IT = (Dxs[i] * 0.5) / Dx[i];
    const double uc = u[ic];
    const double Up = 0.5 * (u[xp] + uc);
    const double Un = 0.5 * (u[xm] + uc);
    const double Vp = v[ic] + (v[xp] - v[ic]) * IT;
    const double Vn = v[ym] + (v[ym + sx] - v[ym]) * IT;
    const double Wp = w[ic] + (w[xp] - w[ic]) * IT;
    const double Wn = w[zm] + (w[zm + sx] - w[zm]) * IT;
    double uPx, uNx, uPy, uNy, uPz, uNz;
    face_unit<S, true>(Up, Un, xpp, xp, ic, xm, xmm, Dx, Dxs, i, u, uPx, uNx);
    face_unit<S, false>(Vp, Vn, ypp, yp, ic, ym, ymm, Dy, Dys, j, u, uPy, uNy);
    face_unit<S, false>(Wp, Wn, zpp, zp, ic, zm, zmm, Dz, Dzs, k, u, uPz, uNz);
    const double conv =
        -dt * ((uPx * uPx - uNx * uNx) / Dxs[i] + (uPy * Vp - uNy * Vn) / Dy[j] + (uPz * Wp - uNz * Wn) / Dz[k]);
    const double diff = nudt * (((u[xp] - uc) / Dx[i + 1] - (uc - u[xm]) / Dx[i]) / Dxs[i] +
                                ((u[yp] - uc) / Dys[j] - (uc - u[ym]) / Dys[j - 1]) / Dy[j] +
                                ((u[zp] - uc) / Dzs[k] - (uc - u[zm]) / Dzs[k - 1]) / Dz[k]);
    double r = uc + conv + diff;
    if (pu) pu[ic] = r;
    if (src_mode == 1) r += src[t];
    else if (src_mode == 2 && ic < N) r += src[ic];
    u1[ic] = r;
  }
  if (j < g.ny - GC - 1) {  // v (:278-351)
    const double IT = (Dys[j] * 0.5) / Dy[j];
    const double vc = v[ic];
    const double Vp = 0.5 * (v[yp] + vc);
    const double Vn = 0.5 * (v[ym] + vc);
    const double Up = u[ic] + (u[yp] - u[ic]) * IT;
    const double Un = u[xm] + (u[xm + sy] - u[xm]) * IT;
    const double Wp = w[ic] + (w[yp] - w[ic]) * IT;
    const double Wn = w[zm] + (w[zm + sy] - w[zm]) * IT;
    double vPx, vNx, vPy, vNy, vPz, vNz;
    face_unit<S, false>(Up, Un, xpp, xp, ic, xm, xmm, Dx, Dxs, i, v, vPx, vNx);
    face_unit<S, true>(Vp, Vn, ypp, yp, ic, ym, ymm, Dy, Dys, j, v, vPy, vNy);
    face_unit<S, false>(Wp, Wn, zpp, zp, ic, zm, zmm, Dz, Dzs, k, v, vPz, vNz);
    const double conv =
        -dt * ((Up * vPx - Un * vNx) / Dx[i] + (vPy * vPy - vNy * vNy) / Dys[j] + (Wp * vPz - Wn * vNz) / Dz[k]);
    const double diff = nudt * (((v[xp] - vc) / Dxs[i] - (vc - v[xm]) / Dxs[i - 1]) / Dx[i] +
                                ((v[yp] - vc) / Dy[j + 1] - (vc - v[ym]) / Dy[j]) / Dys[j] +
                                ((v[zp] - vc) / Dzs[k] - (vc - v[zm]) / Dzs[k - 1]) / Dz[k]);
    double r = vc + conv + diff;
    if (pv) pv[ic] = r;
    if (src_mode == 1) r += src[N + t];
    else if (src_mode == 2 && ic < N) r += src[N + ic];
    v1[ic] = r;
  }
  if (k < kw_end) {  // w (:360-434); with periodic z its last face too
    const double IT = (Dzs[k] * 0.5) / Dz[k];
    const double wc = w[ic];
    const double Wp = 0.5 * (w[zp] + wc);
    const double Wn = 0.5 * (w[zm] + wc);
    const double Up = u[ic] + (u[zp] - u[ic]) * IT;
    const double Un = u[xm] + (u[xm + 1] - u[xm]) * IT;
    const double Vp = v[ic] + (v[zp] - v[ic]) * IT;
    const double Vn = v[ym] + (v[ym + 1] - v[ym]) * IT;
    double wPx, wNx, wPy, wNy, wPz, wNz;
    face_unit<S, false>(Up, Un, xpp, xp, ic, xm, xmm, Dx, Dxs, i, w, wPx, wNx);
    face_unit<S, false>(Vp, Vn, ypp, yp, ic, ym, ymm, Dy, Dys, j, w, wPy, wNy);
    face_unit<S, true>(Wp, Wn, zpp, zp, ic, zm, zmm, Dz, Dzs, k, w, wPz, wNz);
    const double conv =
        -dt * ((Up * wPx - Un * wNx) / Dx[i] + (Vp * wPy - Vn * wNy) / Dy[j] + (wPz * wPz - wNz * wNz) / Dzs[k]);
    const double diff = nudt * (((w[xp] - wc) / Dxs[i] - (wc - w[xm]) / Dxs[i - 1]) / Dx[i] +
                                ((w[yp] - wc) / Dys[j] - (wc - w[ym]) / Dys[j - 1]) / Dy[j] +
                                ((w[zp] - wc) / Dz[k + 1] - (wc - w[zm]) / Dz[k]) / Dzs[k]);
    double r = wc + conv + diff;
    if (pw) pw[ic] = r;
    if (src_mode == 1) r += src[2 * N + t];
    else if (src_mode == 2 && ic < N) r += src[2 * N + ic];
    w1[ic] = r;
  }
}

// The QUICK predictor, LDS-tiled (pred_tiled.h): a workgroup marches a
// tj x TK patch of (j, k) columns over PT_TI i-planes.  Two barriers a
// plane: the plane's face gradients (with the next plane's slabs fetched
// before and stored after them), then the plane's cells (with the column's
// next values fetched before them).  Same arguments and outputs as
// predictor_kernel.
template <int TK>
__global__ void __launch_bounds__(PT_THREADS) quick_predictor_kernel(DevGrid g, PtShape s, PtArgs a) {
  extern __shared__ __align__(16) double pt_mem[];
  const PtLds L = pt_lds<TK>(pt_mem);
  const PtBlock B = pt_block(g, s, (int)blockIdx.x);
  const int tid = (int)threadIdx.x;
  const bool act = pt_active<TK>(g, B, tid);
  const int j = B.j0 + tid / TK, k = B.k0 + tid % TK;
  double r[3][PtGeom<TK>::NPT];
  pt_fetch<TK>(g, B, a.u, a.v, a.w, B.ib, tid, r);
  PtCol c;
  if (act) pt_col_init(g, a.u, a.v, a.w, B.ib, j, k, c);
  pt_tables<TK>(g, B, L, tid);
  pt_store<TK>(L, B.ib, tid, r);
  __syncthreads();
  for (int i = B.ib; i < B.ie; ++i) {
    const bool more = i + 1 < B.ie;
    if (more) pt_fetch<TK>(g, B, a.u, a.v, a.w, i + 1, tid, r);
    pt_grads<TK>(g, B, L, i, tid);
    if (more) pt_store<TK>(L, i + 1, tid, r);
    __syncthreads();
    double nxt[3];
    if (act && more) pt_col_fetch(g, a.u, a.v, a.w, i, j, k, nxt);
    if (act) pt_cell<TK>(g, B, L, a, c, i, tid);
    if (act && more) pt_col_shift(B, L, i, nxt, c);
    __syncthreads();
  }
}

// T1 = P + source_termF at the predictor's positions, from the prediction P
// kept by the step's first pass (T0, hence P, is the same in every pass; the
// sum is the one predictor_kernel forms, so the result is bit-identical)
__global__ void __launch_bounds__(256) pred_add_kernel(DevGrid g, const double *__restrict__ pu,
                                                       const double *__restrict__ pv, const double *__restrict__ pw,
                                                       const double *__restrict__ src, double *__restrict__ u1,
                                                       double *__restrict__ v1, double *__restrict__ w1, int kw_end) {
#pragma clang fp contract(off)  // the oracle's operation order, no fused multiply-adds
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const Cell c = decode_interior(g, t);
  const int64_t ic = icel(g, c.i, c.j, c.k);
  const int64_t N = g.N;
  if (c.i < g.nx - GC - 1) u1[ic] = pu[ic] + src[t];
  if (c.j < g.ny - GC - 1) v1[ic] = pv[ic] + src[N + t];
  if (c.k < kw_end) w1[ic] = pw[ic] + src[2 * N + t];
}

// T1 += source_termF at positions the predictor did not write (reference
// index mapping): ghosted m in [0, N) outside each component's update range.
__global__ void __launch_bounds__(256) source_quirk_fixup_kernel(DevGrid g, double *__restrict__ u1,
                                                                 double *__restrict__ v1, double *__restrict__ w1,
                                                                 const double *__restrict__ src) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.N) return;
  const int64_t nzny = (int64_t)g.ny * g.nz;
  const int i = (int)(m / nzny);
  const int j = (int)((m / g.nz) % g.ny);
  const int k = (int)(m % g.nz);
  const bool in_i = i >= GC && i < g.nx - GC, in_j = j >= GC && j < g.ny - GC, in_k = k >= GC && k < g.nz - GC;
  const bool cd_u = in_j && in_k && i >= GC && i < g.nx - GC - 1;
  const bool cd_v = in_i && in_k && j >= GC && j < g.ny - GC - 1;
  const bool cd_w = in_i && in_j && k >= GC && k < g.nz - GC - 1;
  if (!cd_u) u1[m] += src[m];
  if (!cd_v) v1[m] += src[g.N + m];
  if (!cd_w) w1[m] += src[2 * g.N + m];
}

// T1 += source_termF over all interior cells (RUN_CPU_PC_METHOD.hpp:309-316)
__global__ void __launch_bounds__(256) source_add_kernel(DevGrid g, double *__restrict__ u1, double *__restrict__ v1,
                                                         double *__restrict__ w1, const double *__restrict__ src,
                                                         int quirk) {
#pragma clang fp contract(off)  // the oracle's operation order, no fused multiply-adds
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  int64_t m = t;
  if (!quirk) {
    const Cell c = decode_interior(g, t);
    m = icel(g, c.i, c.j, c.k);
  }
  u1[m] += src[t];
  v1[m] += src[g.N + t];
  w1[m] += src[2 * g.N + t];
}

// createBMatrix_omp, JACOBI_PC branch (inc/matrix/b_matrix.hpp:51-80)
__global__ void __launch_bounds__(256) rhs_kernel(DevGrid g, double ddt, const double *__restrict__ u1,
                                                  const double *__restrict__ v1, const double *__restrict__ w1,
                                                  const double *__restrict__ jp, double *__restrict__ b) {
#pragma clang fp contract(off)  // the oracle's operation order, no fused multiply-adds
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const Cell c = decode_interior(g, t);
  const int64_t ic = icel(g, c.i, c.j, c.k);
  const int64_t sx = (int64_t)g.ny * g.nz;
  b[t] = -jp[t] * ddt *
         ((u1[ic] - u1[ic - sx]) / g.Dx[c.i] + (v1[ic] - v1[ic - g.nz]) / g.Dy[c.j] +
          (w1[ic] - w1[ic - 1]) / g.Dz[c.k]);
}

// Pressure_transform_X_result (inc/pressure/Transform.hpp:8-24)
__global__ void __launch_bounds__(256) scatter_kernel(DevGrid g, const double *__restrict__ X, double *__restrict__ p) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const Cell c = decode_interior(g, t);
  p[icel(g, c.i, c.j, c.k)] = X[t];
}

// update_UandF_omp (inc/dfib/updateUandF.hpp:95-192) with the phases that
// follow it in a pass fused in: source accumulation (RUN_CPU_PC_METHOD.hpp:
// 393-401), drag partial sums (:403-435), the steady-state max difference
// (CheckSteadystate.hpp:165-219) and, in the corrector, the write of the new
// velocity straight into T0 (the T0 <- T3 copy, physicalVariables.hpp:84-102).
// gain scales the forcing (u_new = (1 - gain*eta_f) T2; 1 is the reference,
// bit for bit); t2out, when set, records T2 for the discrete adjoint.
// Faces the update covers beyond the reference's ranges: the outflow face
// of u (convective outflow, projected with p = 0 outside and free of
// immersed-boundary forcing) and the last z face of w (periodic z, whose
// +z pressure neighbour wraps to the first plane).
struct UpdFaces {
  int iu_end, kw_end;  // exclusive ends of the u / w ranges
  int conv, wrapz;
};

__global__ void __launch_bounds__(256) dfib_update_kernel(DevGrid g, UpdFaces uf, double dt, double gain,
                                                          double *__restrict__ t2out, int mode,
                                                          const double *__restrict__ eta,
                                                          const double *__restrict__ p,
                                                          const double *__restrict__ u1,
                                                          const double *__restrict__ v1,
                                                          const double *__restrict__ w1, double *__restrict__ u3,
                                                          double *__restrict__ v3, double *__restrict__ w3,
                                                          double *__restrict__ u0, double *__restrict__ v0,
                                                          double *__restrict__ w0, double *__restrict__ f,
                                                          double *__restrict__ src, double *part_drag,
                                                          double *part_max) {
#pragma clang fp contract(off)  // the oracle's operation order, no fused multiply-adds
  double drag[2] = {0.0, 0.0};
  double mx[3] = {0.0, 0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < g.N; t += stride) {
    const Cell c = decode_interior(g, t);
    const int i = c.i, j = c.j, k = c.k;
    const int64_t ic = icel(g, i, j, k);
    const int64_t sx = (int64_t)g.ny * g.nz, sy = g.nz;
    const int64_t N = g.N;
    const bool from_x = mode & UPD_P_FROM_X;
    // pressure at the cell and its +x/+y/+z neighbours
    const double pc = from_x ? p[t] : p[ic];
    const double eC = eta[ic];
    const double comp_in[3] = {u1[ic], v1[ic], w1[ic]};
    double *const outs3[3] = {u3, v3, w3};
    double *const outs0[3] = {u0, v0, w0};
    const bool active[3] = {i < uf.iu_end, j < g.ny - GC - 1, k < uf.kw_end};
    const int64_t step_ghost[3] = {sx, sy, 1};
    const int64_t step_int[3] = {(int64_t)g.nyc * g.nzc, g.nzc, 1};
    const double *Ds[3] = {g.Dxs, g.Dys, g.Dzs};
    const double *Dd[3] = {g.Dx, g.Dy, g.Dz};
    const int idx[3] = {i, j, k};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (active[q]) {
        // the +q pressure neighbour: across the outflow face p = 0 (Dirichlet),
        // across the periodic z seam the first plane
        const bool outflow = q == 0 && uf.conv && i == g.nx - 3;
        const bool seam = q == 2 && uf.wrapz && k == g.nz - 3;
        double pP;
        if (from_x)
          pP = outflow ? 0.0 : p[seam ? t - (g.nzc - 1) : t + step_int[q]];
        else
          pP = p[ic + step_ghost[q]];
        const double ds = Ds[q][idx[q]];
        const double T2 = comp_in[q] - dt * (pP - pc) / ds;
        const double e = outflow ? 0.0 : eC + (eta[ic + step_ghost[q]] - eC) * ((ds * 0.5) / Dd[q][idx[q]]);
        const double nv = eC * 0.0 + (1.0 - gain * e) * T2;
        const double fv = (nv - T2) / dt;
        f[q * N + t] = fv;
        if (t2out) t2out[q * N + t] = T2;
        if (mode & UPD_SRC_SET) src[q * N + t] = dt * fv;
        else if (mode & UPD_SRC_ADD) src[q * N + t] += dt * fv;
        if (mode & UPD_MAXDIFF) mx[q] = fmax(mx[q], outs0[q][ic] - nv);
        if (mode & UPD_WRITE_T3) outs3[q][ic] = nv;
        if (mode & UPD_WRITE_T0) outs0[q][ic] = nv;
        if (q < 2) drag[q] += fv * (g.Dx[i] * g.Dy[j] * g.Dz[k]);
      } else {
        // last face: the update never writes T3 here, and the reference's
        // T0 <- T3 copy carries that value into T0, where the x- and y-face
        // ghost fills of BC_staggered_main read it before the face's own BC
        // resets it
        if (mode & UPD_MAXDIFF) mx[q] = fmax(mx[q], outs0[q][ic] - outs3[q][ic]);
        if (mode & UPD_WRITE_T0) outs0[q][ic] = outs3[q][ic];
      }
    }
  }
  if (mode & UPD_DRAG) block_partials<2>(drag, part_drag);
  if (mode & UPD_MAXDIFF) block_partials<3, true>(mx, part_max);
}

// source_termF += dt * f (RUN_CPU_PC_METHOD.hpp:393-401)
__global__ void __launch_bounds__(256) source_accumulate_kernel(int64_t n3, double dt, const double *__restrict__ f,
                                                                double *__restrict__ src) {
#pragma clang fp contract(off)  // the oracle's operation order, no fused multiply-adds
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n3) src[t] += dt * f[t];
}

// Boundary faces normal to x (faces 0 and 1) of BC_staggered_main
// (Boundary_Condition.hpp:72-93, 148-154) and BC_staggered_copy
// (updateBC.hpp:32-54); one thread per (j, k).
struct BCArgs {
  double dir[6][3], num[6][3];
  int pd[6];
  const double *uin;  // face-0 u Dirichlet value per interior row j (NULL: dir[0][0])
  int conv;           // face 1 is a convective outflow
  double conv_udt;    // U dt of the convective outflow
  int periodic_z;     // faces 4 and 5 are periodic
};

__device__ inline double nrm(const BCArgs &a, int b, int c, double x) { return a.dir[b][c] + a.num[b][c] * x; }
__device__ inline double tan_(const BCArgs &a, int b, int c, double x) {
  return a.dir[b][c] * 2. + (2.0 * a.num[b][c] - 1.0) * x;
}

__global__ void __launch_bounds__(256) bc_x_kernel(DevGrid g, BCArgs a, int do_main, int do_copy, int do_p,
                                                   double *__restrict__ u, double *__restrict__ v,
                                                   double *__restrict__ w, double *__restrict__ p,
                                                   double *__restrict__ u1, double *__restrict__ v1,
                                                   double *__restrict__ w1) {
#pragma clang fp contract(off)  // the oracle's operation order, no fused multiply-adds
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nj = g.nyc, nk = g.nzc;
  if (t >= nj * nk) return;
  const int j = (int)(t / nk) + GC, k = (int)(t % nk) + GC;
  const int nx = g.nx;
  const int64_t m0 = icel(g, 0, j, k), m1 = icel(g, 1, j, k), m2 = icel(g, 2, j, k);
  const int64_t n1 = icel(g, nx - 1, j, k), n2 = icel(g, nx - 2, j, k), n3 = icel(g, nx - 3, j, k),
                n4 = icel(g, nx - 4, j, k);
  if (do_main) {
    double x;
    x = (a.uin ? a.uin[j - GC] : a.dir[0][0]) + a.num[0][0] * u[m2]; u[m1] = x; u[m0] = x;
    x = tan_(a, 0, 1, v[m2]); v[m1] = x; v[m0] = x;
    x = tan_(a, 0, 2, w[m2]); w[m1] = x; w[m0] = x;
    if (a.conv) {
      // convective outflow, one implicit upwind step from the boundary's
      // previous value: the outflow face of u from the last interior face
      // (one cell width away), the ghosts of v and w from the last cell
      // (one centre distance away); both ghost layers take the new value
      const double su = a.conv_udt / g.Dx[nx - 3], st = a.conv_udt / g.Dxs[nx - 3];
      x = (u[n3] + su * u[n4]) / (1.0 + su); u[n3] = x; u[n2] = x; u[n1] = x;
      x = (v[n2] + st * v[n3]) / (1.0 + st); v[n2] = x; v[n1] = x;
      x = (w[n2] + st * w[n3]) / (1.0 + st); w[n2] = x; w[n1] = x;
    } else {
      x = nrm(a, 1, 0, u[n4]); u[n3] = x; u[n2] = x; u[n1] = x;
      x = tan_(a, 1, 1, v[n3]); v[n2] = x; v[n1] = x;
      x = tan_(a, 1, 2, w[n3]); w[n2] = x; w[n1] = x;
    }
  }
  if (do_p) {
    const double lo = a.pd[0] ? 0.0 : p[m2], hi = a.pd[1] ? 0.0 : p[n3];
    p[m1] = lo; p[m0] = lo; p[n2] = hi; p[n1] = hi;
  }
  if (do_copy) {
    const int64_t ms[4] = {m0, m1, n2, n1};
#pragma unroll
    for (int q = 0; q < 4; ++q) { u1[ms[q]] = u[ms[q]]; v1[ms[q]] = v[ms[q]]; w1[ms[q]] = w[ms[q]]; }
    u1[n3] = u[n3];
  }
}

__global__ void __launch_bounds__(256) bc_y_kernel(DevGrid g, BCArgs a, int do_main, int do_copy, int do_p,
                                                   double *__restrict__ u, double *__restrict__ v,
                                                   double *__restrict__ w, double *__restrict__ p,
                                                   double *__restrict__ u1, double *__restrict__ v1,
                                                   double *__restrict__ w1) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ni = g.nxc, nk = g.nzc;
  if (t >= ni * nk) return;
  const int i = (int)(t / nk) + GC, k = (int)(t % nk) + GC;
  const int ny = g.ny;
  const int64_t m0 = icel(g, i, 0, k), m1 = icel(g, i, 1, k), m2 = icel(g, i, 2, k);
  const int64_t n1 = icel(g, i, ny - 1, k), n2 = icel(g, i, ny - 2, k), n3 = icel(g, i, ny - 3, k),
                n4 = icel(g, i, ny - 4, k);
  if (do_main) {
    double x;
    x = tan_(a, 2, 0, u[m2]); u[m1] = x; u[m0] = x;
    x = nrm(a, 2, 1, v[m2]); v[m1] = x; v[m0] = x;
    x = tan_(a, 2, 2, w[m2]); w[m1] = x; w[m0] = x;
    x = nrm(a, 3, 1, v[n4]); v[n3] = x; v[n2] = x; v[n1] = x;
    x = tan_(a, 3, 0, u[n3]); u[n2] = x; u[n1] = x;
    x = tan_(a, 3, 2, w[n3]); w[n2] = x; w[n1] = x;
  }
  if (do_p) {
    const double lo = a.pd[2] ? 0.0 : p[m2], hi = a.pd[3] ? 0.0 : p[n3];
    p[m1] = lo; p[m0] = lo; p[n2] = hi; p[n1] = hi;
  }
  if (do_copy) {
    const int64_t ms[4] = {m0, m1, n2, n1};
#pragma unroll
    for (int q = 0; q < 4; ++q) { u1[ms[q]] = u[ms[q]]; v1[ms[q]] = v[ms[q]]; w1[ms[q]] = w[ms[q]]; }
    v1[n3] = v[n3];
  }
}

// periodic z: the two ghost layers on each side are the images of the
// interior planes one period away (k = 0, 1 <- nz-4, nz-3; nz-2, nz-1 <- 2,
// 3) for every (i, j) of the ghosted plane, after the x and y faces, so the
// edge and corner ghosts wrap too; the same for u, v, w and p
__global__ void __launch_bounds__(256) bc_z_periodic_kernel(DevGrid g, int do_main, int do_copy, int do_p,
                                                            double *__restrict__ u, double *__restrict__ v,
                                                            double *__restrict__ w, double *__restrict__ p,
                                                            double *__restrict__ u1, double *__restrict__ v1,
                                                            double *__restrict__ w1) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.nx * g.ny) return;
  const int i = (int)(t / g.ny), j = (int)(t % g.ny);
  const int nz = g.nz;
  const int64_t base = icel(g, i, j, 0);
  const int dst[4] = {0, 1, nz - 2, nz - 1}, src[4] = {nz - 4, nz - 3, 2, 3};
  double *const f[3] = {u, v, w};
  double *const f1[3] = {u1, v1, w1};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (do_main)
      for (int c = 0; c < 3; ++c) f[c][base + dst[q]] = f[c][base + src[q]];
    if (do_p) p[base + dst[q]] = p[base + src[q]];
    if (do_copy)
      for (int c = 0; c < 3; ++c) f1[c][base + dst[q]] = f[c][base + dst[q]];
  }
}

__global__ void __launch_bounds__(256) bc_z_kernel(DevGrid g, BCArgs a, int do_main, int do_copy, int do_p,
                                                   double *__restrict__ u, double *__restrict__ v,
                                                   double *__restrict__ w, double *__restrict__ p,
                                                   double *__restrict__ u1, double *__restrict__ v1,
                                                   double *__restrict__ w1) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ni = g.nxc, nj = g.nyc;
  if (t >= ni * nj) return;
  const int i = (int)(t / nj) + GC, j = (int)(t % nj) + GC;
  const int nz = g.nz;
  const int64_t m0 = icel(g, i, j, 0), m1 = icel(g, i, j, 1), m2 = icel(g, i, j, 2);
  const int64_t n1 = icel(g, i, j, nz - 1), n2 = icel(g, i, j, nz - 2), n3 = icel(g, i, j, nz - 3),
                n4 = icel(g, i, j, nz - 4);
  if (do_main) {
    double x;
    x = tan_(a, 4, 0, u[m2]); u[m1] = x; u[m0] = x;
    x = tan_(a, 4, 1, v[m2]); v[m1] = x; v[m0] = x;
    x = nrm(a, 4, 2, w[m2]); w[m1] = x; w[m0] = x;
    x = nrm(a, 5, 2, w[n4]); w[n3] = x; w[n2] = x; w[n1] = x;
    x = tan_(a, 5, 0, u[n3]); u[n2] = x; u[n1] = x;
    x = tan_(a, 5, 1, v[n3]); v[n2] = x; v[n1] = x;
  }
  if (do_p) {
    const double lo = a.pd[4] ? 0.0 : p[m2], hi = a.pd[5] ? 0.0 : p[n3];
    p[m1] = lo; p[m0] = lo; p[n2] = hi; p[n1] = hi;
  }
  if (do_copy) {
    const int64_t ms[4] = {m0, m1, n2, n1};
#pragma unroll
    for (int q = 0; q < 4; ++q) { u1[ms[q]] = u[ms[q]]; v1[ms[q]] = v[ms[q]]; w1[ms[q]] = w[ms[q]]; }
    w1[n3] = w[n3];
  }
}

// DFIB_CylinderZ (inc/dfib/setEta.hpp:136-204), plane k = 0; one thread per
// (i, j).  The multiply-adds the reference build contracts are written as fma.
__global__ void __launch_bounds__(64) eta_plane_kernel(DevGrid g, double cx, double cy, double R, int nsub,
                                                       double inv_n2, double *__restrict__ eta) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ni = g.nxc, nj = g.nyc;
  if (t >= ni * nj) return;
  const int i = (int)(t / nj) + GC, j = (int)(t % nj) + GC;
  const double Dxi = g.Dx[i], Dyj = g.Dy[j], Xi = g.X[i], Yj = g.Y[j];
  const double Diagonal = sqrt(fma(Dxi, Dxi, Dyj * Dyj)) * 0.5;
  const double ax = ((Xi + g.X[i + 1]) * 0.5) - cx;
  const double ay = ((Yj + g.Y[j + 1]) * 0.5) - cy;
  const double Distance = sqrt(fma(ax, ax, ay * ay));
  double e;
  if (fabs(Distance - R) < Diagonal) {
    const double dyg = Dyj / (double)nsub, dxg = Dxi / (double)nsub;
    long long xi = 0;
    for (int ii = 0; ii < nsub; ++ii) {
      const double sx0 = fma((double)ii, dxg, Xi), sx1 = fma((double)(ii + 1), dxg, Xi);
      const double bx = (sx0 + sx1) * 0.5 - cx;
      const double bx2 = bx * bx;
      for (int jj = 0; jj < nsub; ++jj) {
        const double sy0 = fma((double)jj, dyg, Yj), sy1 = fma((double)(jj + 1), dyg, Yj);
        const double by = (sy0 + sy1) * 0.5 - cy;
        const double Sdist = sqrt(fma(bx, bx, by * by));
        xi += (Sdist <= R) ? 1 : 0;
      }
      (void)bx2;
    }
    e = (double)xi * inv_n2;
  } else if (Distance > R) {
    e = 0.0;
  } else {
    e = 1.0;
  }
  eta[icel(g, i, j, 0)] = e;
}

// eta(i, j, k) = eta(i, j, 0) for i < nx-1, j < ny-1, 1 <= k < nz-1 (setEta.hpp:200-203)
__global__ void __launch_bounds__(256) eta_extrude_kernel(DevGrid g, double *__restrict__ eta) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 2);
  if (t >= n) return;
  const int64_t nk = g.nz - 2, nj = g.ny - 1;
  const int k = (int)(t % nk) + 1;
  const int j = (int)((t / nk) % nj);
  const int i = (int)(t / (nk * nj));
  eta[icel(g, i, j, k)] = eta[icel(g, i, j, 0)];
}

__device__ inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// iniU (physicalVariables.hpp:53-75) plus seeded interior noise; T1 = T3 = T0
__global__ void __launch_bounds__(256) init_vel_kernel(DevGrid g, double u0, double v0, double w0, double amp,
                                                       uint64_t seed, double *__restrict__ U0, double *__restrict__ V0,
                                                       double *__restrict__ W0, double *__restrict__ U1,
                                                       double *__restrict__ V1, double *__restrict__ W1,
                                                       double *__restrict__ U3, double *__restrict__ V3,
                                                       double *__restrict__ W3) {
#pragma clang fp contract(off)  // noise = amp * (2u - 1) rounded, then added: identical on every host
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.G) return;
  double val[3] = {u0, v0, w0};
  if (amp != 0.0) {
    const int64_t nzny = (int64_t)g.ny * g.nz;
    const int i = (int)(m / nzny), j = (int)((m / g.nz) % g.ny), k = (int)(m % g.nz);
    if (i >= GC && i < g.nx - GC && j >= GC && j < g.ny - GC && k >= GC && k < g.nz - GC) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const uint64_t z = splitmix64(seed ^ ((uint64_t)c << 56) ^ (uint64_t)m);
        const double u01 = (double)(z >> 11) * 0x1.0p-53;
        val[c] += amp * (2.0 * u01 - 1.0);
      }
    }
  }
  U0[m] = val[0]; V0[m] = val[1]; W0[m] = val[2];
  U1[m] = val[0]; V1[m] = val[1]; W1[m] = val[2];
  U3[m] = val[0]; V3[m] = val[1]; W3[m] = val[2];
}

__global__ void __launch_bounds__(256) copy_kernel(const double *__restrict__ a, double *__restrict__ b, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = a[t];
}

__global__ void __launch_bounds__(256) fill_kernel(double *__restrict__ a, double v, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = v;
}

// standalone drag / max-difference reductions
__global__ void __launch_bounds__(256) drag_kernel(DevGrid g, const double *__restrict__ f, double *part) {
  double v[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < g.N; t += stride) {
    const Cell c = decode_interior(g, t);
    const double dv = g.Dx[c.i] * g.Dy[c.j] * g.Dz[c.k];
    v[0] += f[t] * dv;
    v[1] += f[g.N + t] * dv;
  }
  block_partials<2>(v, part);
}

// sum of one force component times the cell volume (virtualF_Int's sum for
// one direction, virtualForceIntergrator.hpp:5-37)
__global__ void __launch_bounds__(256) force_sum_kernel(DevGrid g, const double *__restrict__ fc, double *part) {
  double v[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < g.N; t += stride) {
    const Cell c = decode_interior(g, t);
    v[0] += fc[t] * (g.Dx[c.i] * g.Dy[c.j] * g.Dz[c.k]);
  }
  block_partials<1>(v, part);
}

__global__ void __launch_bounds__(256) maxdiff_kernel(DevGrid g, const double *__restrict__ u0,
                                                      const double *__restrict__ v0, const double *__restrict__ w0,
                                                      const double *__restrict__ u3, const double *__restrict__ v3,
                                                      const double *__restrict__ w3, double *part) {
  double v[3] = {0.0, 0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < g.N; t += stride) {
    const Cell c = decode_interior(g, t);
    const int64_t ic = icel(g, c.i, c.j, c.k);
    v[0] = fmax(v[0], u0[ic] - u3[ic]);
    v[1] = fmax(v[1], v0[ic] - v3[ic]);
    v[2] = fmax(v[2], w0[ic] - w3[ic]);
  }
  block_partials<3, true>(v, part);
}

// FP64 divergence of a staggered velocity over the interior cells, in the
// finite-volume form of the pressure right-hand side (createBMatrix_omp,
// inc/matrix/b_matrix.hpp:68-80, without its -jp/dt scaling):
//   d = (u[ic] - u[i-1]) / Dx + (v[ic] - v[j-1]) / Dy + (w[ic] - w[k-1]) / Dz.
// The boundary faces (index 1 and n-3 along each component's own axis) are
// read from (ub, vb, wb), the others from (u, v, w): after the corrector's
// update that is the projected velocity on the boundary faces the projection
// saw, so d is what the pressure solve left.  Partials: max |d| and sum d^2
// over every cell, and the same over the fluid cells -- eta zero at the
// cell and its six neighbours, so that no face of the cell carries
// immersed-boundary forcing, and no face on a Dirichlet pressure boundary,
// whose velocity update_UandF leaves uncorrected (its index range stops
// before the last face, inc/dfib/updateUandF.hpp:107-111).
struct DivMask {
  int pd[6];
  int conv, wrapz;  // the outflow face and the z seam are projected faces
};

__global__ void __launch_bounds__(256) divergence_kernel(DevGrid g, DivMask dm, const double *__restrict__ u,
                                                         const double *__restrict__ v, const double *__restrict__ w,
                                                         const double *__restrict__ ub, const double *__restrict__ vb,
                                                         const double *__restrict__ wb,
                                                         const double *__restrict__ eta, double *part_max,
                                                         double *part_sum) {
  double mx[2] = {0.0, 0.0}, sm[2] = {0.0, 0.0};
  const int64_t sx = (int64_t)g.ny * g.nz, sy = g.nz;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < g.N; t += stride) {
    const Cell c = decode_interior(g, t);
    const int64_t ic = icel(g, c.i, c.j, c.k);
    // periodic z: the -z face of the first plane is the last plane's +z face
    const double uP = (c.i == g.nx - 3 && !dm.conv ? ub : u)[ic], uM = (c.i == GC ? ub : u)[ic - sx];
    const double vP = (c.j == g.ny - 3 ? vb : v)[ic], vM = (c.j == GC ? vb : v)[ic - sy];
    const double wP = (c.k == g.nz - 3 && !dm.wrapz ? wb : w)[ic];
    const double wM = dm.wrapz ? w[c.k == GC ? ic + (g.nzc - 1) : ic - 1] : (c.k == GC ? wb : w)[ic - 1];
    const double d = (uP - uM) / g.Dx[c.i] + (vP - vM) / g.Dy[c.j] + (wP - wM) / g.Dz[c.k];
    const double a = fabs(d);
    mx[0] = fmax(mx[0], a);
    sm[0] += d * d;
    const bool dir_face = (dm.pd[0] && c.i == GC) || (dm.pd[1] && !dm.conv && c.i == g.nx - 3) ||
                          (dm.pd[2] && c.j == GC) || (dm.pd[3] && c.j == g.ny - 3) ||
                          (dm.pd[4] && !dm.wrapz && c.k == GC) || (dm.pd[5] && !dm.wrapz && c.k == g.nz - 3);
    const bool fluid = !dir_face && eta[ic] == 0.0 && eta[ic - sx] == 0.0 && eta[ic + sx] == 0.0 &&
                       eta[ic - sy] == 0.0 && eta[ic + sy] == 0.0 && eta[ic - 1] == 0.0 && eta[ic + 1] == 0.0;
    if (fluid) {
      mx[1] = fmax(mx[1], a);
      sm[1] += d * d;
    }
  }
  block_partials<2, true>(mx, part_max);
  block_partials<2>(sm, part_sum);
}

template <int NV, bool MAX>
__global__ void __launch_bounds__(FIN_THREADS) finalize_kernel(const double *part, int nblk, double *out) {
  double tot[NV];
  combine_partials<NV, MAX>(part, nblk, tot);
  if (threadIdx.x == 0)
    for (int q = 0; q < NV; ++q) out[q] = tot[q];
}

BCArgs bc_args(const Ctx &c) {
  BCArgs a;
  for (int b = 0; b < 6; ++b) {
    for (int q = 0; q < 3; ++q) {
      a.dir[b][q] = c.prm.bc_dir[b][q];
      a.num[b][q] = c.prm.bc_num[b][q];
    }
    a.pd[b] = c.prm.p_dirichlet[b];
  }
  a.uin = c.uin;
  a.conv = c.prm.outflow_convective;
  a.conv_udt = c.prm.outflow_speed * c.prm.dt;
  a.periodic_z = c.prm.periodic_z;
  return a;
}

UpdFaces upd_faces(const Ctx &c) {
  UpdFaces f;
  f.conv = c.prm.outflow_convective;
  f.wrapz = c.prm.periodic_z;
  f.iu_end = c.g.nx - GC - (f.conv ? 0 : 1);
  f.kw_end = c.g.nz - GC - (f.wrapz ? 0 : 1);
  return f;
}

}  // namespace

hipError_t launch_init_fields(Ctx &c, double noise_amp, uint64_t seed) {
  const DevGrid &g = c.g;
  init_vel_kernel<<<blocks_for(g.G, 256, 1 << 30), 256, 0, c.stream>>>(
      g, c.prm.u_init[0], c.prm.u_init[1], c.prm.u_init[2], noise_amp, seed, c.f[CFD_T0_U], c.f[CFD_T0_V],
      c.f[CFD_T0_W], c.f[CFD_T1_U], c.f[CFD_T1_V], c.f[CFD_T1_W], c.f[CFD_T3_U], c.f[CFD_T3_V], c.f[CFD_T3_W]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if ((e = launch_fill(c, c.f[CFD_P], 0.0, g.G)) != hipSuccess) return e;
  if ((e = launch_fill(c, c.f[CFD_X], 0.0, g.N)) != hipSuccess) return e;
  if ((e = launch_fill(c, c.f[CFD_F], 0.0, 3 * g.N)) != hipSuccess) return e;
  if ((e = launch_fill(c, c.f[CFD_SRC], 0.0, 3 * g.N)) != hipSuccess) return e;
  return launch_bc(c, true, true, true);
}

hipError_t launch_eta_cylinder_z(Ctx &c, double inv_n2) {
  const DevGrid &g = c.g;
  hipError_t e = launch_fill(c, c.f[CFD_ETA], 0.0, g.G);
  if (e != hipSuccess) return e;
  if (c.prm.ib_on) {
    const int64_t n = (int64_t)g.nxc * g.nyc;
    eta_plane_kernel<<<blocks_for(n, 64, 1 << 30), 64, 0, c.stream>>>(g, c.prm.ib_cx, c.prm.ib_cy, c.prm.ib_r,
                                                                      c.prm.ib_nsub, inv_n2, c.f[CFD_ETA]);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int64_t m = (int64_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 2);
    eta_extrude_kernel<<<blocks_for(m, 256, 1 << 30), 256, 0, c.stream>>>(g, c.f[CFD_ETA]);
    e = hipGetLastError();
  }
  return e;
}

hipError_t launch_predictor(Ctx &c, const double *u0, const double *v0, const double *w0, double *u1, double *v1,
                            double *w1, const double *src, int src_mode, double *const keep[3]) {
  const int nb = blocks_for(c.g.N, 256, 1 << 30);
  double *k0 = keep ? keep[0] : nullptr, *k1 = keep ? keep[1] : nullptr, *k2 = keep ? keep[2] : nullptr;
  const double nu = 1.0 / c.prm.Re;
  const int kw_end = upd_faces(c).kw_end;
  const double *vis = nullptr;
  if (c.prm.les) {  // the eddy viscosity of this T0 (RUN_CPU_PC_METHOD.hpp:298-300, 449-451)
    if (!c.viseff) {
      hipError_t e = hipMalloc((void **)&c.viseff, (size_t)c.g.N * sizeof(double));
      if (e != hipSuccess) {
        c.viseff = nullptr;
        return e;
      }
    }
    smagorinsky_kernel<<<nb, 256, 0, c.stream>>>(c.g, u0, v0, w0, c.viseff);
    vis = c.viseff;
  }
  if (c.prm.convection == CFD_CONV_LUD)
    predictor_kernel<CFD_CONV_LUD><<<nb, 256, 0, c.stream>>>(c.g, nu, c.prm.dt, u0, v0, w0, u1, v1, w1, src,
                                                             src_mode, k0, k1, k2, vis, kw_end);
  else if (c.prm.convection == CFD_CONV_UD)
    predictor_kernel<CFD_CONV_UD><<<nb, 256, 0, c.stream>>>(c.g, nu, c.prm.dt, u0, v0, w0, u1, v1, w1, src,
                                                            src_mode, k0, k1, k2, vis, kw_end);
  else {
    const PtShape s = pt_shape(c.g);
    const PtArgs a{nu, c.prm.dt, u0, v0, w0, u1, v1, w1, src, src_mode, k0, k1, k2, vis, kw_end};
    const int64_t nblk = (int64_t)s.njt * s.nkt * s.nit;
    if (nblk > INT32_MAX) return hipErrorInvalidValue;
    const size_t lds = pt_lds_doubles(s) * sizeof(double);
    switch (s.tk) {
      case 4: quick_predictor_kernel<4><<<(unsigned)nblk, PT_THREADS, lds, c.stream>>>(c.g, s, a); break;
      case 8: quick_predictor_kernel<8><<<(unsigned)nblk, PT_THREADS, lds, c.stream>>>(c.g, s, a); break;
      case 16: quick_predictor_kernel<16><<<(unsigned)nblk, PT_THREADS, lds, c.stream>>>(c.g, s, a); break;
      case 32: quick_predictor_kernel<32><<<(unsigned)nblk, PT_THREADS, lds, c.stream>>>(c.g, s, a); break;
      default: quick_predictor_kernel<64><<<(unsigned)nblk, PT_THREADS, lds, c.stream>>>(c.g, s, a); break;
    }
  }
  return hipGetLastError();
}

hipError_t launch_pred_add(Ctx &c, double *const keep[3], const double *src, double *u1, double *v1, double *w1) {
  pred_add_kernel<<<blocks_for(c.g.N, 256, 1 << 30), 256, 0, c.stream>>>(c.g, keep[0], keep[1], keep[2], src, u1,
                                                                         v1, w1, upd_faces(c).kw_end);
  return hipGetLastError();
}

hipError_t launch_source_quirk_fixup(Ctx &c, double *u1, double *v1, double *w1, const double *src) {
  source_quirk_fixup_kernel<<<blocks_for(c.g.N, 256, 1 << 30), 256, 0, c.stream>>>(c.g, u1, v1, w1, src);
  return hipGetLastError();
}

hipError_t launch_source_add(Ctx &c, double *u1, double *v1, double *w1, const double *src, int quirk) {
  source_add_kernel<<<blocks_for(c.g.N, 256, 1 << 30), 256, 0, c.stream>>>(c.g, u1, v1, w1, src, quirk);
  return hipGetLastError();
}

hipError_t launch_rhs(Ctx &c, const double *u1, const double *v1, const double *w1, const double *jp, double *b) {
  rhs_kernel<<<blocks_for(c.g.N, 256, 1 << 30), 256, 0, c.stream>>>(c.g, 1.0 / c.prm.dt, u1, v1, w1, jp, b);
  return hipGetLastError();
}

hipError_t launch_scatter_p(Ctx &c, const double *X, double *p) {
  scatter_kernel<<<blocks_for(c.g.N, 256, 1 << 30), 256, 0, c.stream>>>(c.g, X, p);
  return hipGetLastError();
}

hipError_t launch_dfib_update_into(Ctx &c, int mode, const double *p_or_x, double *drag_out, double *max_out,
                                   double *t2out) {
  if (!drag_out) mode &= ~UPD_DRAG;
  if (!max_out) mode &= ~UPD_MAXDIFF;
  const int nblk = (mode & (UPD_DRAG | UPD_MAXDIFF)) ? blocks_for(c.g.N, 256, RED_MAX_BLOCKS)
                                                     : blocks_for(c.g.N, 256, 1 << 30);
  dfib_update_kernel<<<nblk, 256, 0, c.stream>>>(
      c.g, upd_faces(c), c.prm.dt, c.gain, t2out, mode, c.f[CFD_ETA], p_or_x, c.f[CFD_T1_U], c.f[CFD_T1_V], c.f[CFD_T1_W], c.f[CFD_T3_U],
      c.f[CFD_T3_V], c.f[CFD_T3_W], c.f[CFD_T0_U], c.f[CFD_T0_V], c.f[CFD_T0_W], c.f[CFD_F], c.f[CFD_SRC],
      red_part_host(c, RS_DRAG), red_part_host(c, RS_MAX));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (mode & UPD_DRAG) {
    finalize_kernel<2, false><<<1, FIN_THREADS, 0, c.stream>>>(red_part_host(c, RS_DRAG), nblk, drag_out);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (mode & UPD_MAXDIFF) {
    finalize_kernel<3, true><<<1, FIN_THREADS, 0, c.stream>>>(red_part_host(c, RS_MAX), nblk, max_out);
    e = hipGetLastError();
  }
  return e;
}

hipError_t launch_source_accumulate(Ctx &c) {
  const int64_t n3 = 3 * c.g.N;
  source_accumulate_kernel<<<blocks_for(n3, 256, 1 << 30), 256, 0, c.stream>>>(n3, c.prm.dt, c.f[CFD_F],
                                                                               c.f[CFD_SRC]);
  return hipGetLastError();
}

hipError_t launch_bc(Ctx &c, bool main, bool copy, bool pressure) {
  const DevGrid &g = c.g;
  const BCArgs a = bc_args(c);
  double *u = c.f[CFD_T0_U], *v = c.f[CFD_T0_V], *w = c.f[CFD_T0_W], *p = c.f[CFD_P];
  double *u1 = c.f[CFD_T1_U], *v1 = c.f[CFD_T1_V], *w1 = c.f[CFD_T1_W];
  const int m = main ? 1 : 0, cp = copy ? 1 : 0, pp = pressure ? 1 : 0;
  bc_x_kernel<<<blocks_for((int64_t)g.nyc * g.nzc, 256, 1 << 30), 256, 0, c.stream>>>(g, a, m, cp, pp, u, v, w, p,
                                                                                       u1, v1, w1);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  bc_y_kernel<<<blocks_for((int64_t)g.nxc * g.nzc, 256, 1 << 30), 256, 0, c.stream>>>(g, a, m, cp, pp, u, v, w, p,
                                                                                       u1, v1, w1);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (c.prm.periodic_z)
    bc_z_periodic_kernel<<<blocks_for((int64_t)g.nx * g.ny, 256, 1 << 30), 256, 0, c.stream>>>(g, m, cp, pp, u, v, w,
                                                                                              p, u1, v1, w1);
  else
    bc_z_kernel<<<blocks_for((int64_t)g.nxc * g.nyc, 256, 1 << 30), 256, 0, c.stream>>>(g, a, m, cp, pp, u, v, w, p,
                                                                                         u1, v1, w1);
  return hipGetLastError();
}

hipError_t launch_periodic_refresh_p(Ctx &c, double *p) {
  if (!c.prm.periodic_z) return hipSuccess;
  const DevGrid &g = c.g;
  bc_z_periodic_kernel<<<blocks_for((int64_t)g.nx * g.ny, 256, 1 << 30), 256, 0, c.stream>>>(
      g, 0, 0, 1, nullptr, nullptr, nullptr, p, nullptr, nullptr, nullptr);
  return hipGetLastError();
}

hipError_t launch_periodic_refresh(Ctx &c, double *u, double *v, double *w) {
  if (!c.prm.periodic_z) return hipSuccess;
  const DevGrid &g = c.g;
  bc_z_periodic_kernel<<<blocks_for((int64_t)g.nx * g.ny, 256, 1 << 30), 256, 0, c.stream>>>(
      g, 1, 0, 0, u, v, w, nullptr, nullptr, nullptr, nullptr);
  return hipGetLastError();
}

hipError_t launch_drag(Ctx &c) {
  const int nblk = blocks_for(c.g.N, 256, RED_MAX_BLOCKS);
  drag_kernel<<<nblk, 256, 0, c.stream>>>(c.g, c.f[CFD_F], red_part_host(c, RS_DRAG));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  finalize_kernel<2, false><<<1, FIN_THREADS, 0, c.stream>>>(red_part_host(c, RS_DRAG), nblk, c.red.out + RS_DRAG * 4);
  return hipGetLastError();
}

hipError_t launch_force_sum(Ctx &c, int dir, double *out) {
  const int nblk = blocks_for(c.g.N, 256, RED_MAX_BLOCKS);
  force_sum_kernel<<<nblk, 256, 0, c.stream>>>(c.g, c.f[CFD_F] + (int64_t)dir * c.g.N, red_part_host(c, RS_DRAG));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  finalize_kernel<1, false><<<1, FIN_THREADS, 0, c.stream>>>(red_part_host(c, RS_DRAG), nblk, out);
  return hipGetLastError();
}

hipError_t launch_maxdiff(Ctx &c) {
  const int nblk = blocks_for(c.g.N, 256, RED_MAX_BLOCKS);
  maxdiff_kernel<<<nblk, 256, 0, c.stream>>>(c.g, c.f[CFD_T0_U], c.f[CFD_T0_V], c.f[CFD_T0_W], c.f[CFD_T3_U],
                                             c.f[CFD_T3_V], c.f[CFD_T3_W], red_part_host(c, RS_MAX));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  finalize_kernel<3, true><<<1, FIN_THREADS, 0, c.stream>>>(red_part_host(c, RS_MAX), nblk, c.red.out + RS_MAX * 4);
  return hipGetLastError();
}

hipError_t launch_divergence(Ctx &c, const double *const vel[3], const double *const bnd[3], double *out) {
  const int nblk = blocks_for(c.g.N, 256, RED_MAX_BLOCKS);
  double *pmax = red_part_host(c, RS_MISC), *psum = red_part_host(c, RS_NORM);
  DivMask dm;
  for (int f = 0; f < 6; ++f) dm.pd[f] = c.prm.p_dirichlet[f];
  dm.conv = c.prm.outflow_convective;
  dm.wrapz = c.prm.periodic_z;
  divergence_kernel<<<nblk, 256, 0, c.stream>>>(c.g, dm, vel[0], vel[1], vel[2], bnd[0], bnd[1], bnd[2],
                                                c.f[CFD_ETA], pmax, psum);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // out: max |d| (all), max |d| (fluid), sum d^2 (all), sum d^2 (fluid)
  finalize_kernel<2, true><<<1, FIN_THREADS, 0, c.stream>>>(pmax, nblk, out);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  finalize_kernel<2, false><<<1, FIN_THREADS, 0, c.stream>>>(psum, nblk, out + 2);
  return hipGetLastError();
}

hipError_t launch_copy(Ctx &c, const double *src, double *dst, int64_t n) {
  copy_kernel<<<blocks_for(n, 256, 1 << 30), 256, 0, c.stream>>>(src, dst, n);
  return hipGetLastError();
}

// bandwidth calibration: a grid-stride copy of n doubles (16 B moved per element)
__global__ void __launch_bounds__(256) stream_copy_kernel(const double *__restrict__ a, double *__restrict__ b,
                                                          int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(a + t), b + t);
}

hipError_t launch_copy_kernel_bench(Ctx &c, const double *src, double *dst, int64_t n) {
  stream_copy_kernel<<<blocks_for(n, 256, 8192), 256, 0, c.stream>>>(src, dst, n);
  return hipGetLastError();
}

hipError_t launch_fill(Ctx &c, double *dst, double v, int64_t n) {
  fill_kernel<<<blocks_for(n, 256, 1 << 30), 256, 0, c.stream>>>(dst, v, n);
  return hipGetLastError();
}

}  // namespace cfd
