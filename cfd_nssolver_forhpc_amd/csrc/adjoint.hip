// adjoint.hip -- discrete adjoint of the prediction-correction timestep on
// gfx950 (an addition: the reference has no adjoint).
//
// Objective: J = mean over steps [avg_from, nsteps) of the corrector pass's
// drag coefficient cD = -2 sum(f_x dV) / lz (virtualF_Int,
// inc/dfib/virtualForceIntergrator.hpp:5-37).  Parameters: the inflow
// profile u_in(y) (the face-0 u Dirichlet value of BC_staggered_main per
// interior row) and the DFIB forcing gain (u_new = (1 - gain eta_f) T2).
//
// The forward run keeps a checkpoint (T0, p, X and the solver counters)
// every `every` steps.  The reverse sweep restores each checkpoint, re-runs
// its segment recording every step's starting T0 and every pass's T2 (the
// only forward data the transposed step needs), and walks the segment
// backwards.  The forward step is deterministic (fixed reduction order), so
// the recomputed states are bit-identical to the first run.
//
// Per step, in reverse: BC_staggered_main^T (face kernels z, y, x, faces in
// reverse order), then per pass from the corrector down: DFIB update^T (with
// the gain gradient as a deterministic two-stage reduction), the pressure
// gather, the transposed Poisson solve, createBMatrix^T fused with the
// source-term and T1 bookkeeping, and finally the QUICK predictor's Jacobian
// transpose.  The transposed solve reuses the forward operator: A = J V^-1 S
// with J = diag(jp), V = diag(cell volume) and S symmetric, so A^T y = g is
// solved as A z = J V^-1 g, y = V z / jp.  The predictor transpose avoids
// atomics: each output cell writes its 21 local partial products
// (contrib[component][slot][cell]) and every input cell gathers the slots
// that point at it, in a fixed order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "dev_common.h"

namespace cfd {

struct AdjCkpt {
  double *buf = nullptr;  // T0 u, v, w (3G), p (G), X (N)
  int loop = 1, solves = 0, solves32 = 0;
  double time = 0.0;
  bool p_from_x = false;
};

struct AdjState {
  int nsteps = 0, avg_from = 0, every = 0, passes = 0;
  std::vector<double> cd;
  std::vector<AdjCkpt> ck;
  AdjCkpt fin;
  bool ready = false;
  unsigned long long generation = 0;  // counts objective runs; the last one's number while ready
  // reverse-sweep workspace
  int ws_every = 0, ws_passes = 0;
  double *lx = nullptr, *lsum = nullptr;          // 3G each: adjoint of T0 / summed adjoint of T1
  double *Lam = nullptr, *lt2 = nullptr;          // 3N each: adjoint of the source term / of T2
  double *lX = nullptr, *rhs = nullptr, *z = nullptr, *lb = nullptr;  // N each
  double *contrib = nullptr;                      // 63G
  double *guin = nullptr;                         // ny_cal * nz_cal
  double *gacc = nullptr;                         // gain gradient accumulator
  double *xseg = nullptr, *t2seg = nullptr;       // every * 3G, every * passes * 3N
  std::vector<int> solved;
  int adj_solves = 0;
  long long adj_iters = 0;
  double fwd_ms = 0.0, bwd_ms = 0.0;
};

namespace {

template <typename T>
int amalloc(T **p, size_t n) {
  *p = nullptr;
  hipError_t e = hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(CFD_ERR_NOMEM, std::string("adjoint: hipMalloc of ") + std::to_string(n * sizeof(T)) +
                                   " bytes failed: " + hipGetErrorString(e));
  }
  return CFD_OK;
}

template <typename T>
void afree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

#define ADJ_HIP(call)                                                                             \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return fail(CFD_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct ICell {
  int i, j, k;
};

__device__ inline ICell decode_cell(const DevGrid &g, int64_t t) {
  ICell c;
  const int64_t nzc = g.nzc, nyc = g.nyc;
  const int64_t jk = t % (nyc * nzc);
  c.i = (int)(t / (nyc * nzc)) + GC;
  c.j = (int)(jk / nzc) + GC;
  c.k = (int)(jk % nzc) + GC;
  return c;
}

__device__ inline ICell decode_ghosted(const DevGrid &g, int64_t m) {
  ICell c;
  const int64_t nynz = (int64_t)g.ny * g.nz;
  c.i = (int)(m / nynz);
  c.j = (int)((m / g.nz) % g.ny);
  c.k = (int)(m % g.nz);
  return c;
}

// ---- BC_staggered_main^T -------------------------------------------------
// forward: y[m0] = y[m1] (= y[m2]) = a x + c  ->  lx += a (ly[m0] + ly[m1] ..); ly = 0
struct BCNum {
  double num[6][3];
};

__device__ inline void adj2(double *L, double a, int64_t m0, int64_t m1, int64_t m2) {
  const double s = L[m0] + L[m1];
  L[m0] = 0.0;
  L[m1] = 0.0;
  L[m2] += a * s;
}
__device__ inline void adj3(double *L, double a, int64_t m0, int64_t m1, int64_t m2, int64_t m3) {
  const double s = L[m0] + L[m1] + L[m2];
  L[m0] = 0.0;
  L[m1] = 0.0;
  L[m2] = 0.0;
  L[m3] += a * s;
}
__device__ inline double an(const BCNum &b, int f, int c) { return b.num[f][c]; }
__device__ inline double at(const BCNum &b, int f, int c) { return 2.0 * b.num[f][c] - 1.0; }

__global__ void __launch_bounds__(256) bc_z_adj_kernel(DevGrid g, BCNum b, double *__restrict__ lu,
                                                       double *__restrict__ lv, double *__restrict__ lw) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.nxc * g.nyc) return;
  const int i = (int)(t / g.nyc) + GC, j = (int)(t % g.nyc) + GC;
  const int nz = g.nz;
  const int64_t m0 = icel(g, i, j, 0), m1 = icel(g, i, j, 1), m2 = icel(g, i, j, 2);
  const int64_t n1 = icel(g, i, j, nz - 1), n2 = icel(g, i, j, nz - 2), n3 = icel(g, i, j, nz - 3),
                n4 = icel(g, i, j, nz - 4);
  adj2(lv, at(b, 5, 1), n1, n2, n3);  // BC_5
  adj2(lu, at(b, 5, 0), n1, n2, n3);
  adj3(lw, an(b, 5, 2), n1, n2, n3, n4);
  adj2(lw, an(b, 4, 2), m0, m1, m2);  // BC_4
  adj2(lv, at(b, 4, 1), m0, m1, m2);
  adj2(lu, at(b, 4, 0), m0, m1, m2);
}

__global__ void __launch_bounds__(256) bc_y_adj_kernel(DevGrid g, BCNum b, double *__restrict__ lu,
                                                       double *__restrict__ lv, double *__restrict__ lw) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.nxc * g.nzc) return;
  const int i = (int)(t / g.nzc) + GC, k = (int)(t % g.nzc) + GC;
  const int ny = g.ny;
  const int64_t m0 = icel(g, i, 0, k), m1 = icel(g, i, 1, k), m2 = icel(g, i, 2, k);
  const int64_t n1 = icel(g, i, ny - 1, k), n2 = icel(g, i, ny - 2, k), n3 = icel(g, i, ny - 3, k),
                n4 = icel(g, i, ny - 4, k);
  adj2(lw, at(b, 3, 2), n1, n2, n3);  // BC_3
  adj2(lu, at(b, 3, 0), n1, n2, n3);
  adj3(lv, an(b, 3, 1), n1, n2, n3, n4);
  adj2(lw, at(b, 2, 2), m0, m1, m2);  // BC_2
  adj2(lv, an(b, 2, 1), m0, m1, m2);
  adj2(lu, at(b, 2, 0), m0, m1, m2);
}

// convective outflow y[out] = (y[old] + s y[in]) / (1 + s) into the
// positions outs (y[old] among them)  ->  l[old] = S / (1 + s),
// l[in] += s S / (1 + s), the other outputs 0, with S the outputs' sum
__device__ inline void adj_conv2(double *L, double s, int64_t n1, int64_t n2, int64_t in) {
  const double S = L[n1] + L[n2];
  L[n1] = 0.0;
  L[n2] = S / (1.0 + s);
  L[in] += s * S / (1.0 + s);
}

// guin[(j-2)*nz_cal + (k-2)] accumulates the adjoint of the face-0 u value;
// conv_udt > 0: face 1 is the convective outflow of bc_x_kernel (U dt)
__global__ void __launch_bounds__(256) bc_x_adj_kernel(DevGrid g, BCNum b, double conv_udt, double *__restrict__ lu,
                                                       double *__restrict__ lv, double *__restrict__ lw,
                                                       double *__restrict__ guin) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.nyc * g.nzc) return;
  const int j = (int)(t / g.nzc) + GC, k = (int)(t % g.nzc) + GC;
  const int nx = g.nx;
  const int64_t m0 = icel(g, 0, j, k), m1 = icel(g, 1, j, k), m2 = icel(g, 2, j, k);
  const int64_t n1 = icel(g, nx - 1, j, k), n2 = icel(g, nx - 2, j, k), n3 = icel(g, nx - 3, j, k),
                n4 = icel(g, nx - 4, j, k);
  if (conv_udt > 0.0) {
    // u: the outflow face from its old value and the last interior face;
    // v, w: the ghosts from their old value and the last cell
    const double su = conv_udt / g.Dx[nx - 3], st = conv_udt / g.Dxs[nx - 3];
    adj_conv2(lw, st, n1, n2, n3);
    adj_conv2(lv, st, n1, n2, n3);
    const double S = lu[n1] + lu[n2] + lu[n3];
    lu[n1] = 0.0;
    lu[n2] = 0.0;
    lu[n3] = S / (1.0 + su);
    lu[n4] += su * S / (1.0 + su);
  } else {
    adj2(lw, at(b, 1, 2), n1, n2, n3);  // BC_1
    adj2(lv, at(b, 1, 1), n1, n2, n3);
    adj3(lu, an(b, 1, 0), n1, n2, n3, n4);
  }
  adj2(lw, at(b, 0, 2), m0, m1, m2);  // BC_0
  adj2(lv, at(b, 0, 1), m0, m1, m2);
  const double s = lu[m0] + lu[m1];
  guin[t] += s;
  lu[m0] = 0.0;
  lu[m1] = 0.0;
  lu[m2] += an(b, 0, 0) * s;
}

// periodic z (bc_z_periodic_kernel): the ghost planes {0, 1, nz-2, nz-1} are
// copies of {nz-4, nz-3, 2, 3} over the whole ghosted (i, j) plane, so their
// adjoints fold into those images.  fold_into_images: the same for one array.
__device__ inline void fold_z(double *L, int64_t base, int nz) {
  const int dst[4] = {0, 1, nz - 2, nz - 1}, src[4] = {nz - 4, nz - 3, 2, 3};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    L[base + src[q]] += L[base + dst[q]];
    L[base + dst[q]] = 0.0;
  }
}

__global__ void __launch_bounds__(256) bc_z_periodic_adj_kernel(DevGrid g, double *__restrict__ lu,
                                                                double *__restrict__ lv, double *__restrict__ lw) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)g.nx * g.ny) return;
  const int64_t base = icel(g, (int)(t / g.ny), (int)(t % g.ny), 0);
  fold_z(lu, base, g.nz);
  fold_z(lv, base, g.nz);
  fold_z(lw, base, g.nz);
}

// ---- update_UandF^T ------------------------------------------------------
// u3 = (1 - gain e) T2, f = (u3 - T2)/dt with lf = dt*Lam (+ the drag weight
// on f_x in the corrector): writes lt2 (3N, 0 where a component is not
// updated) and the per-block partial of dJ/dgain.
// Faces updated beyond the table ranges (convective outflow, periodic z):
// the outflow face of u (iu_end = nx - 2; no forcing there, e = 0) and the
// seam face of w (kw_end = nz - 2).
struct AdjFaces {
  int iu_end, kw_end, conv, wrapz;
};

__global__ void __launch_bounds__(256) dfib_adj_kernel(DevGrid g, AdjFaces af, double dt, double gain, double drag_w,
                                                       const double *__restrict__ eta, const double *__restrict__ t2,
                                                       const double *__restrict__ lu3, const double *__restrict__ lv3,
                                                       const double *__restrict__ lw3, const double *__restrict__ Lam,
                                                       double *__restrict__ lt2, double *part) {
  double acc[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t N = g.N;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < N; t += stride) {
    const ICell c = decode_cell(g, t);
    const int64_t ic = icel(g, c.i, c.j, c.k);
    const int idx[3] = {c.i, c.j, c.k};
    const int lim[3] = {af.iu_end, g.ny - GC - 1, af.kw_end};
    const int64_t sg[3] = {(int64_t)g.ny * g.nz, g.nz, 1};
    const double *Ds[3] = {g.Dxs, g.Dys, g.Dzs};
    const double *Dd[3] = {g.Dx, g.Dy, g.Dz};
    const double *L3[3] = {lu3, lv3, lw3};
    const double eC = eta[ic];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      if (idx[q] < lim[q]) {
        const int I = idx[q];
        const double ds = Ds[q][I];
        const bool outflow = q == 0 && af.conv && I == g.nx - 3;
        const double e = outflow ? 0.0 : eC + (eta[ic + sg[q]] - eC) * ((ds * 0.5) / Dd[q][I]);
        double lf = dt * Lam[q * N + t];
        if (q == 0 && drag_w != 0.0) lf += drag_w * (-2.0 / g.lz) * (g.Dx[c.i] * g.Dy[c.j] * g.Dz[c.k]);
        const double lu3t = (L3[q] ? L3[q][ic] : 0.0) + lf / dt;
        lt2[q * N + t] = (1.0 - gain * e) * lu3t - lf / dt;
        acc[0] += -e * t2[q * N + t] * lu3t;
      } else {
        lt2[q * N + t] = 0.0;
      }
    }
  }
  block_partials<1>(acc, part);
}

__global__ void __launch_bounds__(FIN_THREADS) gain_finalize(const double *part, int nblk, double *gacc) {
  double tot[1];
  combine_partials<1>(part, nblk, tot);
  if (threadIdx.x == 0) gacc[0] += tot[0];
}

// adjoint of the interior pressure: T2 = T1 - dt (p[+] - p) / ds per component
// (with periodic z the seam face's +z neighbour is the first plane)
__global__ void __launch_bounds__(256) lx_gather_kernel(DevGrid g, int wrapz, double dt,
                                                        const double *__restrict__ lt2, double *__restrict__ lX) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const ICell c = decode_cell(g, t);
  const int idx[3] = {c.i, c.j, c.k};
  const int64_t si[3] = {(int64_t)g.nyc * g.nzc, g.nzc, 1};
  const double *Ds[3] = {g.Dxs, g.Dys, g.Dzs};
  const int64_t N = g.N;
  double acc = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int I = idx[q];
    acc += dt / Ds[q][I] * lt2[q * N + t];
    if (I - 1 >= GC) acc -= dt / Ds[q][I - 1] * lt2[q * N + t - si[q]];
    else if (q == 2 && wrapz) acc -= dt / Ds[2][g.nz - 3] * lt2[2 * N + t + (g.nzc - 1)];
  }
  lX[t] = acc;
}

// A^T y = lX  <=>  A z = jp lX / vol,  y = vol z / jp
__global__ void __launch_bounds__(256) adj_rhs_form_kernel(DevGrid g, const double *__restrict__ jp,
                                                           const double *__restrict__ lX, double *__restrict__ rhs) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const ICell c = decode_cell(g, t);
  rhs[t] = jp[t] * lX[t] / (g.Dx[c.i] * g.Dy[c.j] * g.Dz[c.k]);
}

__global__ void __launch_bounds__(256) adj_rhs_back_kernel(DevGrid g, const double *__restrict__ jp,
                                                           const double *__restrict__ z, double *__restrict__ lb) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const ICell c = decode_cell(g, t);
  lb[t] = (g.Dx[c.i] * g.Dy[c.j] * g.Dz[c.k]) * z[t] / jp[t];
}

// Adjoint of T1 in one pass, over every ghosted position: the DFIB part
// (lt2, zero where the component is not updated) plus createBMatrix^T
// (b = -jp/dt (du/dx + dv/dy + dw/dz) per interior cell), then added to the
// per-step sum and, at interior cells, to the source-term adjoint (T1 = P + S
// there).  With periodic z the pass refreshes T1's z ghost planes from their
// interior images before the right-hand side reads them, so a ghost plane's
// adjoint is added to its image's and the ghost keeps none.
__device__ inline double t1_val(const DevGrid &g, int ci, int cj, int ck, int q, double ddt, int solved,
                                const double *__restrict__ jp, const double *__restrict__ lb,
                                const double *__restrict__ lt2) {
  const int idx[3] = {ci, cj, ck};
  const int lo = GC, hi[3] = {g.nx - GC, g.ny - GC, g.nz - GC};
  const bool in[3] = {ci >= lo && ci < hi[0], cj >= lo && cj < hi[1], ck >= lo && ck < hi[2]};
  const bool interior = in[0] && in[1] && in[2];
  const int64_t N = g.N;
  const double *Dd[3] = {g.Dx, g.Dy, g.Dz};
  const int64_t t = interior ? icelcal(g, ci, cj, ck) : 0;
  double val = 0.0;
  if (interior) val = lt2[q * N + t];
  if (solved) {
    if (interior) val += -jp[t] * ddt * lb[t] / Dd[q][idx[q]];
    // the cell one step up along q reads this position as its lower face
    const int up = idx[q] + 1;
    bool cell_up = up >= lo && up < hi[q];
#pragma unroll
    for (int r = 0; r < 3; ++r)
      if (r != q) cell_up = cell_up && in[r];
    if (cell_up) {
      int ii = ci, jj = cj, kk = ck;
      if (q == 0) ii = up;
      else if (q == 1) jj = up;
      else kk = up;
      const int64_t tu = icelcal(g, ii, jj, kk);
      val -= -jp[tu] * ddt * lb[tu] / Dd[q][up];
    }
  }
  return val;
}

__global__ void __launch_bounds__(256) t1_adj_kernel(DevGrid g, int wrapz, double ddt, int solved,
                                                     const double *__restrict__ jp, const double *__restrict__ lb,
                                                     const double *__restrict__ lt2, double *__restrict__ lsu,
                                                     double *__restrict__ lsv, double *__restrict__ lsw,
                                                     double *__restrict__ Lam) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= g.G) return;
  const ICell c = decode_ghosted(g, m);
  const int nz = g.nz;
  if (wrapz && (c.k < GC || c.k >= nz - GC)) return;  // a refreshed ghost: its adjoint goes to its image
  // with periodic z, the ghost plane that is this plane's image, if any
  int img = -1;
  if (wrapz) img = c.k == nz - 4 ? 0 : c.k == nz - 3 ? 1 : c.k == 2 ? nz - 2 : c.k == 3 ? nz - 1 : -1;
  const bool interior = c.i >= GC && c.i < g.nx - GC && c.j >= GC && c.j < g.ny - GC && c.k >= GC && c.k < nz - GC;
  const int64_t N = g.N;
  double *LS[3] = {lsu, lsv, lsw};
  const int64_t t = interior ? icelcal(g, c.i, c.j, c.k) : 0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double val = t1_val(g, c.i, c.j, c.k, q, ddt, solved, jp, lb, lt2);
    if (img >= 0) val += t1_val(g, c.i, c.j, img, q, ddt, solved, jp, lb, lt2);
    LS[q][m] += val;
    if (interior) Lam[q * N + t] += val;
  }
}

// ---- QUICK predictor^T ----------------------------------------------------
// Coefficients of a QUICK face pair (QuickUnitA / QuickUnitB,
// inc/source/convection_and_difussion.hpp:20-84) with respect to phi at
// (pp, p, c, n, nn) for the upwind branches taken.
__device__ inline void quick_coef(bool typeA, double Up, double Un, const double *__restrict__ Di,
                                  const double *__restrict__ Ds, int I, double cP[5], double cN[5]) {
#pragma unroll
  for (int m = 0; m < 5; ++m) cP[m] = cN[m] = 0.0;
  if (Up > 0.0) {
    const double a = typeA ? 0.125 * Di[I + 1] * Di[I + 1] / Ds[I] : 0.125 * Ds[I] * Ds[I] / Di[I];
    const double h1 = typeA ? Di[I + 1] : Ds[I], h0 = typeA ? Di[I] : Ds[I - 1];
    cP[1] = 0.5 - a / h1;
    cP[2] = 0.5 + a / h1 + a / h0;
    cP[3] = -a / h0;
  } else {
    const double a = typeA ? 0.125 * Di[I + 1] * Di[I + 1] / Ds[I + 1] : 0.125 * Ds[I] * Ds[I] / Di[I + 1];
    const double h2 = typeA ? Di[I + 2] : Ds[I + 1], h1 = typeA ? Di[I + 1] : Ds[I];
    cP[0] = -a / h2;
    cP[1] = 0.5 + a / h2 + a / h1;
    cP[2] = 0.5 - a / h1;
  }
  if (Un > 0.0) {
    const double b = typeA ? 0.125 * Di[I] * Di[I] / Ds[I - 1] : 0.125 * Ds[I - 1] * Ds[I - 1] / Di[I - 1];
    const double h0 = typeA ? Di[I] : Ds[I - 1], hm = typeA ? Di[I - 1] : Ds[I - 2];
    cN[2] = 0.5 - b / h0;
    cN[3] = 0.5 + b / h0 + b / hm;
    cN[4] = -b / hm;
  } else {
    const double b = typeA ? 0.125 * Di[I] * Di[I] / Ds[I] : 0.125 * Ds[I - 1] * Ds[I - 1] / Di[I];
    const double h1 = typeA ? Di[I + 1] : Ds[I], h0 = typeA ? Di[I] : Ds[I - 1];
    cN[1] = -b / h1;
    cN[2] = 0.5 + b / h1 + b / h0;
    cN[3] = 0.5 - b / h0;
  }
}

constexpr int PSLOTS = 21;  // 13 own-component + 2 x 4 convecting-velocity slots per output
__host__ __device__ inline int first_other(int c) { return c == 0 ? 1 : 0; }

// Own slots: 0 centre, then per direction a: 1+4a +s_a, 2+4a -s_a,
// 3+4a +2 s_a, 4+4a -2 s_a.  Convecting-velocity slots of velocity a in
// output c (13.. for first_other(c), 17.. for the other): centre, +s_c,
// -s_a, -s_a + s_c.
__global__ void __launch_bounds__(256) pred_contrib_kernel(DevGrid g, int kw_end, double nu, double dt,
                                                           const double *__restrict__ u, const double *__restrict__ v,
                                                           const double *__restrict__ w, const double *__restrict__ lsu,
                                                           const double *__restrict__ lsv,
                                                           const double *__restrict__ lsw, double *__restrict__ contrib) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.N) return;
  const ICell cl = decode_cell(g, t);
  const int idx[3] = {cl.i, cl.j, cl.k};
  const int64_t ic = icel(g, cl.i, cl.j, cl.k);
  const int64_t sa[3] = {(int64_t)g.ny * g.nz, g.nz, 1};
  const int lim[3] = {g.nx - GC - 1, g.ny - GC - 1, kw_end};
  const double *Di[3] = {g.Dx, g.Dy, g.Dz}, *Ds[3] = {g.Dxs, g.Dys, g.Dzs};
  const double *F[3] = {u, v, w};
  const double *LY[3] = {lsu, lsv, lsw};
  const double nudt = nu * dt;
  const int64_t G = g.G;
  for (int c = 0; c < 3; ++c) {
    if (idx[c] >= lim[c]) continue;
    const double ly = LY[c][ic];
    double own[13], crA[4], crB[4];
#pragma unroll
    for (int s = 0; s < 13; ++s) own[s] = 0.0;
#pragma unroll
    for (int s = 0; s < 4; ++s) crA[s] = crB[s] = 0.0;
    own[0] = ly;
    const double *phi = F[c];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int64_t s = sa[a];
      const int I = idx[a];
      const bool is_own = a == c;
      const int base = 1 + 4 * a;
      const int slot5[5] = {base + 2, base, 0, base + 1, base + 3};  // pp p c n nn
      const double f5[5] = {phi[ic + 2 * s], phi[ic + s], phi[ic], phi[ic - s], phi[ic - 2 * s]};
      double Vp, Vn, IT = 0.0;
      if (is_own) {
        Vp = 0.5 * (f5[1] + f5[2]);
        Vn = 0.5 * (f5[3] + f5[2]);
      } else {
        const int Ic = idx[c];
        IT = (Ds[c][Ic] * 0.5) / Di[c][Ic];
        const int64_t sc = sa[c];
        const double *V = F[a];
        Vp = V[ic] + (V[ic + sc] - V[ic]) * IT;
        Vn = V[ic - s] + (V[ic - s + sc] - V[ic - s]) * IT;
      }
      double cP[5], cN[5];
      quick_coef(is_own, Vp, Vn, Di[a], Ds[a], I, cP, cN);
      double phP = 0.0, phN = 0.0;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        phP += cP[q] * f5[q];
        phN += cN[q] * f5[q];
      }
      const double H = is_own ? Ds[a][I] : Di[a][I];
      double lP, lN;
      if (is_own) {
        lP = -dt / H * 2.0 * phP * ly;
        lN = dt / H * 2.0 * phN * ly;
      } else {
        lP = -dt / H * Vp * ly;
        lN = dt / H * Vn * ly;
        const double lVp = -dt / H * phP * ly, lVn = dt / H * phN * ly;
        double *X = a == first_other(c) ? crA : crB;
        X[0] += (1.0 - IT) * lVp;
        X[1] += IT * lVp;
        X[2] += (1.0 - IT) * lVn;
        X[3] += IT * lVn;
      }
#pragma unroll
      for (int q = 0; q < 5; ++q) own[slot5[q]] += cP[q] * lP + cN[q] * lN;
      double hp, hm, H2;
      if (is_own) {
        hp = Di[a][I + 1]; hm = Di[a][I]; H2 = Ds[a][I];
      } else {
        hp = Ds[a][I]; hm = Ds[a][I - 1]; H2 = Di[a][I];
      }
      const double k1 = nudt * ly / H2;
      own[base] += k1 / hp;
      own[0] -= k1 / hp + k1 / hm;
      own[base + 1] += k1 / hm;
    }
    double *out = contrib + (int64_t)c * PSLOTS * G + ic;
#pragma unroll
    for (int s = 0; s < 13; ++s) out[(int64_t)s * G] = own[s];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      out[(int64_t)(13 + s) * G] = crA[s];
      out[(int64_t)(17 + s) * G] = crB[s];
    }
  }
}

// positions BC_staggered_copy fills in T1 from T0 (updateBC.hpp:6-109); with
// periodic z the predictor writes w's last face itself
__device__ inline bool copied(const DevGrid &g, int wrapz, int comp, const ICell &c) {
  const bool ii = c.i >= GC && c.i < g.nx - GC, jj = c.j >= GC && c.j < g.ny - GC, kk = c.k >= GC && c.k < g.nz - GC;
  const bool gi = c.i < GC || c.i >= g.nx - GC, gj = c.j < GC || c.j >= g.ny - GC, gk = c.k < GC || c.k >= g.nz - GC;
  if (gi && jj && kk) return true;
  if (gj && ii && kk) return true;
  if (gk && ii && jj) return true;
  if (comp == 0 && c.i == g.nx - GC - 1 && jj && kk) return true;
  if (comp == 1 && c.j == g.ny - GC - 1 && ii && kk) return true;
  if (comp == 2 && !wrapz && c.k == g.nz - GC - 1 && ii && jj) return true;
  return false;
}

__global__ void __launch_bounds__(256) pred_gather_kernel(DevGrid g, int wrapz, const double *__restrict__ contrib,
                                                          const double *__restrict__ lsu, const double *__restrict__ lsv,
                                                          const double *__restrict__ lsw, double *__restrict__ lu,
                                                          double *__restrict__ lv, double *__restrict__ lw) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t G = g.G;
  if (m >= G) return;
  const ICell c = decode_ghosted(g, m);
  const int64_t sa[3] = {(int64_t)g.ny * g.nz, g.nz, 1};
  const double *LS[3] = {lsu, lsv, lsw};
  double *LX[3] = {lu, lv, lw};
#pragma unroll
  for (int cp = 0; cp < 3; ++cp) {
    double acc = 0.0;
    const double *own = contrib + (int64_t)cp * PSLOTS * G;
    // own-component slots of output cp
    {
      const int64_t o = m;
      acc += own[o];
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int64_t offs[4] = {sa[a], -sa[a], 2 * sa[a], -2 * sa[a]};
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t o = m - offs[r];
        if (o >= 0 && o < G) acc += own[(int64_t)(1 + 4 * a + r) * G + o];
      }
    }
    // convecting-velocity slots of the other outputs
#pragma unroll
    for (int oc = 0; oc < 3; ++oc) {
      if (oc == cp) continue;
      const int base = cp == first_other(oc) ? 13 : 17;
      const int64_t offs[4] = {0, sa[oc], -sa[cp], -sa[cp] + sa[oc]};
      const double *cr = contrib + ((int64_t)oc * PSLOTS + base) * G;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t o = m - offs[r];
        if (o >= 0 && o < G) acc += cr[(int64_t)r * G + o];
      }
    }
    if (copied(g, wrapz, cp, c)) acc += LS[cp][m];
    // a ghost of T0 keeps its value through the step (the update writes the
    // interior only), so the adjoint its end-of-step reads left there (the
    // convective outflow's old ghost value) carries over
    const bool ghost = c.i < GC || c.i >= g.nx - GC || c.j < GC || c.j >= g.ny - GC || c.k < GC || c.k >= g.nz - GC;
    LX[cp][m] = ghost ? acc + LX[cp][m] : acc;
  }
}

int blocks(int64_t n) { return blocks_for(n, 256, 1 << 30); }

BCNum bc_num(const Ctx &c) {
  BCNum b;
  for (int f = 0; f < 6; ++f)
    for (int q = 0; q < 3; ++q) b.num[f][q] = c.prm.bc_num[f][q];
  return b;
}

hipError_t launch_bc_adjoint(Ctx &c, AdjState &A) {
  const DevGrid &g = c.g;
  const BCNum b = bc_num(c);
  const int64_t G = g.G;
  double *lu = A.lx, *lv = A.lx + G, *lw = A.lx + 2 * G;
  // reverse of launch_bc: z (periodic images, or the table), y, x
  if (c.prm.periodic_z)
    bc_z_periodic_adj_kernel<<<blocks((int64_t)g.nx * g.ny), 256, 0, c.stream>>>(g, lu, lv, lw);
  else
    bc_z_adj_kernel<<<blocks((int64_t)g.nxc * g.nyc), 256, 0, c.stream>>>(g, b, lu, lv, lw);
  bc_y_adj_kernel<<<blocks((int64_t)g.nxc * g.nzc), 256, 0, c.stream>>>(g, b, lu, lv, lw);
  const double conv_udt = c.prm.outflow_convective ? c.prm.outflow_speed * c.prm.dt : 0.0;
  bc_x_adj_kernel<<<blocks((int64_t)g.nyc * g.nzc), 256, 0, c.stream>>>(g, b, conv_udt, lu, lv, lw, A.guin);
  return hipGetLastError();
}

int save_ckpt(Ctx &c, AdjCkpt &k) {
  const size_t G = c.g.G, N = c.g.N;
  if (!k.buf) {
    int rc = amalloc(&k.buf, 4 * G + N);
    if (rc) return rc;
  }
  for (int q = 0; q < 3; ++q)
    ADJ_HIP(hipMemcpyAsync(k.buf + q * G, c.f[CFD_T0_U + q], G * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  ADJ_HIP(hipMemcpyAsync(k.buf + 3 * G, c.f[CFD_P], G * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  ADJ_HIP(hipMemcpyAsync(k.buf + 4 * G, c.f[CFD_X], N * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  k.loop = c.loop;
  k.time = c.time;
  k.solves = c.solves;
  k.solves32 = c.solves32;
  k.p_from_x = c.p_from_x;
  return CFD_OK;
}

int restore_ckpt(Ctx &c, const AdjCkpt &k) {
  const size_t G = c.g.G, N = c.g.N;
  for (int q = 0; q < 3; ++q)
    ADJ_HIP(hipMemcpyAsync(c.f[CFD_T0_U + q], k.buf + q * G, G * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  ADJ_HIP(hipMemcpyAsync(c.f[CFD_P], k.buf + 3 * G, G * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  ADJ_HIP(hipMemcpyAsync(c.f[CFD_X], k.buf + 4 * G, N * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  ADJ_HIP(launch_bc(c, false, true, false));  // T1's boundary faces = T0's, as after the checkpointed step
  c.loop = k.loop;
  c.time = k.time;
  c.solves = k.solves;
  c.solves32 = k.solves32;
  c.p_from_x = k.p_from_x;
  return CFD_OK;
}

int ensure_workspace(Ctx &c, AdjState &A) {
  const size_t G = c.g.G, N = c.g.N;
  if (A.lx && A.ws_every == A.every && A.ws_passes == A.passes) return CFD_OK;
  afree(A.xseg);
  afree(A.t2seg);
  int rc;
  if (!A.lx) {
    if ((rc = amalloc(&A.lx, 3 * G)) || (rc = amalloc(&A.lsum, 3 * G)) || (rc = amalloc(&A.Lam, 3 * N)) ||
        (rc = amalloc(&A.lt2, 3 * N)) || (rc = amalloc(&A.lX, N)) || (rc = amalloc(&A.rhs, N)) ||
        (rc = amalloc(&A.z, N)) || (rc = amalloc(&A.lb, N)) || (rc = amalloc(&A.contrib, (size_t)3 * PSLOTS * G)) ||
        (rc = amalloc(&A.guin, (size_t)c.g.nyc * c.g.nzc)) || (rc = amalloc(&A.gacc, 1)))
      return rc;
    // the predictor transpose writes only at output cells; every other slot stays 0
    ADJ_HIP(launch_fill(c, A.contrib, 0.0, (int64_t)3 * PSLOTS * G));
  }
  if ((rc = amalloc(&A.xseg, (size_t)A.every * 3 * G)) ||
      (rc = amalloc(&A.t2seg, (size_t)A.every * A.passes * 3 * N)))
    return rc;
  ADJ_HIP(launch_fill(c, A.t2seg, 0.0, (int64_t)A.every * A.passes * 3 * N));
  A.ws_every = A.every;
  A.ws_passes = A.passes;
  return CFD_OK;
}

// the transposed timestep: A.lx holds the adjoint of the next step's T0 on
// entry and this step's on return
int adjoint_step(Ctx &c, AdjState &A, const double *x, const double *t2, bool solved, double drag_w) {
  const DevGrid &g = c.g;
  const int64_t G = g.G, N = g.N;
  const cfd_params &P = c.prm;
  const int L = A.passes - 1;
  ADJ_HIP(launch_bc_adjoint(c, A));
  ADJ_HIP(launch_fill(c, A.lsum, 0.0, 3 * G));
  ADJ_HIP(launch_fill(c, A.Lam, 0.0, 3 * N));
  const int nb = blocks_for(N, 256, RED_MAX_BLOCKS);
  double *part = red_part_host(c, RS_MISC);
  AdjFaces af;
  af.conv = P.outflow_convective;
  af.wrapz = P.periodic_z;
  af.iu_end = g.nx - GC - (af.conv ? 0 : 1);
  af.kw_end = g.nz - GC - (af.wrapz ? 0 : 1);
  for (int q = L; q >= 0; --q) {
    const bool last = q == L;
    dfib_adj_kernel<<<nb, 256, 0, c.stream>>>(g, af, P.dt, c.gain, last ? drag_w : 0.0, c.f[CFD_ETA],
                                              t2 + (size_t)q * 3 * N, last ? A.lx : nullptr,
                                              last ? A.lx + G : nullptr, last ? A.lx + 2 * G : nullptr, A.Lam,
                                              A.lt2, part);
    gain_finalize<<<1, FIN_THREADS, 0, c.stream>>>(part, nb, A.gacc);
    ADJ_HIP(hipGetLastError());
    if (solved) {
      lx_gather_kernel<<<blocks(N), 256, 0, c.stream>>>(g, af.wrapz, P.dt, A.lt2, A.lX);
      adj_rhs_form_kernel<<<blocks(N), 256, 0, c.stream>>>(g, c.f[CFD_JP], A.lX, A.rhs);
      ADJ_HIP(hipGetLastError());
      int it = 0;
      double rel = 0.0;
      A.adj_solves++;
      int rc;
      if (P.precond == CFD_PRECOND_MG)
        rc = pbicgstab_solve_aux(c, A.rhs, A.z, P.tol, P.maxit,
                                 P.precision == CFD_MIXED && P.mixed_mode == CFD_MIXED_PC32, &it, &rel);
      else
        rc = bicgstab_solve_aux(c, A.rhs, A.z, P.tol, P.maxit, A.adj_solves, &it, &rel);
      if (rc) return rc;
      A.adj_iters += it;
      adj_rhs_back_kernel<<<blocks(N), 256, 0, c.stream>>>(g, c.f[CFD_JP], A.z, A.lb);
    }
    t1_adj_kernel<<<blocks(G), 256, 0, c.stream>>>(g, af.wrapz, 1.0 / P.dt, solved ? 1 : 0, c.f[CFD_JP], A.lb,
                                                   A.lt2, A.lsum, A.lsum + G, A.lsum + 2 * G, A.Lam);
    ADJ_HIP(hipGetLastError());
  }
  pred_contrib_kernel<<<blocks(N), 256, 0, c.stream>>>(g, af.kw_end, 1.0 / P.Re, P.dt, x, x + G, x + 2 * G, A.lsum,
                                                       A.lsum + G, A.lsum + 2 * G, A.contrib);
  pred_gather_kernel<<<blocks(G), 256, 0, c.stream>>>(g, af.wrapz, A.contrib, A.lsum, A.lsum + G, A.lsum + 2 * G,
                                                      A.lx, A.lx + G, A.lx + 2 * G);
  ADJ_HIP(hipGetLastError());
  return CFD_OK;
}

int adjoint_checks(const Ctx &c) {
  if (c.prm.op != CFD_OP_ELL) return fail(CFD_ERR_ARG, "adjoint: needs the ELL pressure operator");
  if (c.prm.solver != CFD_SOLVER_BICGSTAB) return fail(CFD_ERR_ARG, "adjoint: needs the BiCGSTAB pressure solve");
  if (c.prm.source_index_quirk) return fail(CFD_ERR_ARG, "adjoint: needs source_index_quirk = 0");
  if (c.prm.convection != CFD_CONV_QUICK) return fail(CFD_ERR_ARG, "adjoint: needs the QUICK convection scheme");
  if (c.prm.les) return fail(CFD_ERR_ARG, "adjoint: needs les = 0");
  if (c.prm.periodic_z) {
    // the transposed solve relies on A = diag(jp) V^-1 S with S symmetric;
    // the wrapped z coupling is symmetric on a uniform z axis
    const double dz = c.hDz[0], dzs = c.hDzs[0];
    for (size_t k = 0; k < c.hDz.size(); ++k)
      if (std::fabs(c.hDz[k] - dz) > 1e-12 * dz || std::fabs(c.hDzs[k] - dzs) > 1e-12 * dzs)
        return fail(CFD_ERR_ARG, "adjoint: periodic z needs a uniform z axis");
  }
  if (c.g.nxc < 2 || c.g.nyc < 2 || c.g.nzc < 2) return fail(CFD_ERR_ARG, "adjoint: needs >= 2 cells per axis");
  if (!c.ell_ready) return fail(CFD_ERR_STATE, "adjoint: ELL operator not assembled");
  return CFD_OK;
}

}  // namespace

// the recorded objective run no longer describes the context: its state,
// parameters, forcing gain or inflow profile changed after it
void adjoint_invalidate(Ctx &c) {
  if (c.adj) c.adj->ready = false;
}

void adjoint_release(Ctx &c) {
  AdjState *A = c.adj;
  if (!A) return;
  for (auto &k : A->ck) afree(k.buf);
  afree(A->fin.buf);
  afree(A->lx); afree(A->lsum); afree(A->Lam); afree(A->lt2); afree(A->lX); afree(A->rhs); afree(A->z);
  afree(A->lb); afree(A->contrib); afree(A->guin); afree(A->gacc); afree(A->xseg); afree(A->t2seg);
  delete A;
  c.adj = nullptr;
}

}  // namespace cfd

using namespace cfd;

#define ADJ_CTX(ctx)                                                                \
  do {                                                                              \
    if (!(ctx)) return fail(CFD_ERR_ARG, "context is NULL");                        \
    if (hipSetDevice(((Ctx *)(ctx))->device) != hipSuccess)                         \
      return fail(CFD_ERR_HIP, "hipSetDevice failed");                              \
  } while (0)

extern "C" {

int cfd_set_inflow_profile(cfd_ctx *ctx, const double *u_in, size_t n) {
  ADJ_CTX(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!u_in) {
    afree(c.uin);
    adjoint_invalidate(c);
    return CFD_OK;
  }
  if (n != (size_t)c.g.nyc) return fail(CFD_ERR_ARG, "inflow profile needs ny_cal values");
  for (size_t j = 0; j < n; ++j)
    if (!std::isfinite(u_in[j])) return fail(CFD_ERR_ARG, "inflow profile is not finite");
  if (!c.uin) {
    int rc = amalloc(&c.uin, n);
    if (rc) return rc;
  }
  ADJ_HIP(hipMemcpyAsync(c.uin, u_in, n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  ADJ_HIP(hipStreamSynchronize(c.stream));
  adjoint_invalidate(c);
  return CFD_OK;
}

int cfd_set_forcing_gain(cfd_ctx *ctx, double gain) {
  ADJ_CTX(ctx);
  if (!std::isfinite(gain)) return fail(CFD_ERR_ARG, "gain is not finite");
  reinterpret_cast<Ctx *>(ctx)->gain = gain;
  adjoint_invalidate(*reinterpret_cast<Ctx *>(ctx));
  return CFD_OK;
}

int cfd_adjoint_generation(cfd_ctx *ctx, unsigned long long *generation) {
  ADJ_CTX(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!generation) return fail(CFD_ERR_ARG, "generation is NULL");
  *generation = (c.adj && c.adj->ready) ? c.adj->generation : 0ull;
  return CFD_OK;
}

int cfd_adjoint_forward(cfd_ctx *ctx, int nsteps, int avg_from, int ckpt_every, double *J) {
  ADJ_CTX(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!J) return fail(CFD_ERR_ARG, "J is NULL");
  if (nsteps < 1 || avg_from < 0 || avg_from >= nsteps || ckpt_every < 1)
    return fail(CFD_ERR_ARG, "adjoint: need 0 <= avg_from < nsteps and ckpt_every >= 1");
  int rc = adjoint_checks(c);
  if (rc) return rc;
  if (!c.adj) c.adj = new AdjState();
  AdjState &A = *c.adj;
  const int ncp = (nsteps + ckpt_every - 1) / ckpt_every;
  if ((int)A.ck.size() > ncp) {
    for (size_t q = ncp; q < A.ck.size(); ++q) afree(A.ck[q].buf);
    A.ck.resize(ncp);
  }
  A.ck.resize(ncp);
  A.ready = false;
  A.nsteps = nsteps;
  A.avg_from = avg_from;
  A.every = ckpt_every;
  A.passes = c.prm.inner_loops + 1;
  A.cd.assign(nsteps, 0.0);
  ADJ_HIP(hipEventRecord(c.ev0, c.stream));
  for (int s = 0; s < nsteps; ++s) {
    if (s % ckpt_every == 0 && (rc = save_ckpt(c, A.ck[s / ckpt_every]))) return rc;
    cfd_step_stats st;
    if ((rc = step_impl(c, &st))) return rc;
    A.cd[s] = st.cd[A.passes - 1];
  }
  if ((rc = save_ckpt(c, A.fin))) return rc;
  ADJ_HIP(hipEventRecord(c.ev1, c.stream));
  ADJ_HIP(hipEventSynchronize(c.ev1));
  float ms = 0.0f;
  ADJ_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  A.fwd_ms = ms;
  double s = 0.0;
  for (int q = avg_from; q < nsteps; ++q) s += A.cd[q];
  *J = s / (double)(nsteps - avg_from);
  A.ready = true;
  A.generation += 1;
  return CFD_OK;
}

int cfd_adjoint_backward(cfd_ctx *ctx, double *grad_uin, size_t n, double *grad_gain, cfd_adjoint_stats *stats) {
  ADJ_CTX(ctx);
  Ctx &c = *reinterpret_cast<Ctx *>(ctx);
  if (!grad_uin || !grad_gain) return fail(CFD_ERR_ARG, "NULL gradient output");
  if (n != (size_t)c.g.nyc) return fail(CFD_ERR_ARG, "grad_uin needs ny_cal values");
  if (!c.adj || !c.adj->ready) return fail(CFD_ERR_STATE, "adjoint: no objective run (cfd_adjoint_forward)");
  int rc = adjoint_checks(c);
  if (rc) return rc;
  AdjState &A = *c.adj;
  if ((rc = ensure_workspace(c, A))) return rc;
  const int64_t G = c.g.G, N = c.g.N;
  ADJ_HIP(launch_fill(c, A.lx, 0.0, 3 * G));
  ADJ_HIP(launch_fill(c, A.z, 0.0, N));
  ADJ_HIP(launch_fill(c, A.lb, 0.0, N));
  ADJ_HIP(launch_fill(c, A.guin, 0.0, (int64_t)c.g.nyc * c.g.nzc));
  ADJ_HIP(launch_fill(c, A.gacc, 0.0, 1));
  A.adj_solves = 0;
  A.adj_iters = 0;
  A.solved.assign(A.every, 0);
  const int ncp = (int)A.ck.size();
  const double wJ = 1.0 / (double)(A.nsteps - A.avg_from);
  int recomputed = 0;
  ADJ_HIP(hipEventRecord(c.ev0, c.stream));
  for (int seg = ncp - 1; seg >= 0; --seg) {
    if ((rc = restore_ckpt(c, A.ck[seg]))) return rc;
    const int s0 = seg * A.every, s1 = std::min(A.nsteps, s0 + A.every);
    for (int s = s0; s < s1; ++s) {
      double *xs = A.xseg + (size_t)(s - s0) * 3 * G;
      for (int q = 0; q < 3; ++q)
        ADJ_HIP(hipMemcpyAsync(xs + q * G, c.f[CFD_T0_U + q], G * sizeof(double), hipMemcpyDeviceToDevice,
                               c.stream));
      A.solved[s - s0] = c.loop > 1;
      c.t2_rec = A.t2seg + (size_t)(s - s0) * A.passes * 3 * N;
      rc = step_impl(c, nullptr);
      c.t2_rec = nullptr;
      if (rc) return rc;
      ++recomputed;
    }
    for (int s = s1 - 1; s >= s0; --s) {
      const double *xs = A.xseg + (size_t)(s - s0) * 3 * G;
      const double *t2 = A.t2seg + (size_t)(s - s0) * A.passes * 3 * N;
      if ((rc = adjoint_step(c, A, xs, t2, A.solved[s - s0] != 0, s >= A.avg_from ? wJ : 0.0))) return rc;
    }
  }
  ADJ_HIP(launch_bc_adjoint(c, A));  // BC_staggered_main of the initial state
  if ((rc = restore_ckpt(c, A.fin))) return rc;
  ADJ_HIP(hipEventRecord(c.ev1, c.stream));
  std::vector<double> h((size_t)c.g.nyc * c.g.nzc + 1);
  ADJ_HIP(hipMemcpyAsync(h.data(), A.guin, (h.size() - 1) * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  ADJ_HIP(hipMemcpyAsync(h.data() + h.size() - 1, A.gacc, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  ADJ_HIP(hipStreamSynchronize(c.stream));
  float ms = 0.0f;
  ADJ_HIP(hipEventElapsedTime(&ms, c.ev0, c.ev1));
  A.bwd_ms = ms;
  for (int j = 0; j < c.g.nyc; ++j) {
    double s = 0.0;
    for (int k = 0; k < c.g.nzc; ++k) s += h[(size_t)j * c.g.nzc + k];
    grad_uin[j] = s;
  }
  *grad_gain = h.back();
  if (stats) {
    stats->steps = A.nsteps;
    stats->checkpoints = ncp;
    stats->recomputed_steps = recomputed;
    stats->adjoint_solves = A.adj_solves;
    stats->adjoint_iters = (int)std::min<long long>(A.adj_iters, 0x7fffffff);
    stats->forward_ms = A.fwd_ms;
    stats->backward_ms = A.bwd_ms;
  }
  return CFD_OK;
}

}  // extern "C"
