// k_mg.hip -- cell-centred geometric multigrid V-cycle used as a
// preconditioner of the pressure solve (an addition: the reference solves
// with Jacobi-scaled BiCGSTAB only).
//
// The Jacobi-scaled operator is A = J L with J = diag(jp) and L the
// finite-volume Laplacian of createPressureMatrix (inc/matrix/
// PressureMatrix.hpp:5-117): per axis, the coefficient to the -/+ neighbour
// is 1/(D[i] Ds[i-1]) / 1/(D[i] Ds[i]), Neumann faces drop theirs, Dirichlet
// faces keep theirs in the diagonal.  M^-1 r = V-cycle(L, r / jp) is a cheap
// approximation of A^-1 r.
//
// Every level is matrix-free: the grids are tensor products, so a level is
// its extents plus per-axis arrays (coefficients to the -/+ neighbour, their
// diagonal share, cell widths); a smoothing sweep streams only x and b.
// Coarse levels halve every axis that is even and at least 4 cells long;
// their coefficients are rediscretised from the merged cell widths (a
// Dirichlet face keeps its mirror-ghost coefficient).  With periodic z every
// level wraps in z like the operator (the first and last planes are
// neighbours across the seam); z is halved only while its length stays even,
// so the red-black colouring stays consistent across the seam.  Restriction is the
// volume-weighted sum of the children's residuals, prolongation adds the
// parent's correction.  Smoother: red-black Gauss-Seidel, nu sweeps before
// and after.  Levels of at most MG_SMALL cells run inside one 512-thread
// workgroup (one launch for the whole bottom of the cycle); the coarsest is
// solved by its dense inverse or smoothed MG_COARSE_SWEEPS times.
//
// The finest level, two thirds of a cycle's bytes, keeps x and b
// colour-split (ColourSplit) and has kernels of its own: colour sweeps whose
// threads walk MG_PLANES planes and load each other-colour entry once for
// its three uses, the post-smoothing's first sweep with the prolongation
// folded in and its last one writing the result, and a restriction that
// reads the children as colour pairs.  The next levels run the same kernels
// in the same layout while the level above halves every axis and they are
// not part of the one-workgroup bottom, up to MG_SPLIT_MAX levels
// (split_levels).  Each does the arithmetic of the generic per-cell
// functions term for term, so the cycle's result is the same bit for bit in
// either layout.
//
// Every kernel returns at once when the Krylov solver's `done` flag is set,
// so preconditioned iterations can be launched in batches like the plain
// ones.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "dev_common.h"

namespace cfd {

constexpr int MG_MAX_LEVELS = 16;
constexpr int64_t MG_SMALL = 4096;
constexpr int MG_COARSE_SWEEPS = 48;
// colour-split levels at most (the finest one and the next ones).  At
// 512x256x64, two (512x256x64, 256x128x32) ran the multigrid steps 1-2%
// faster than the finest alone; a third level (128x64x16) gave that back
// and a fourth cost 2% more: the small levels' sweeps are bound by their
// launches, and fewer threads walking planes hide less latency.
constexpr int MG_SPLIT_MAX = 2;

// per-level device view: ax = cm[nx] cp[nx] dd[nx] w[nx] (and ay, az)
struct MGView {
  int nx, ny, nz, fx, fy, fz;
  int wz;  // periodic z: k = 0 and k = nz - 1 are neighbours (nz even)
  int64_t n;
  const double *ax, *ay, *az;
};

struct MGLevel {
  MGView v{};
  double *coef = nullptr;               // ax | ay | az
  double *x64 = nullptr, *b64 = nullptr;
  float *x32 = nullptr, *b32 = nullptr;
  // a colour-split level below the finest: its x in natural layout, written
  // by its last post-smoothing half-sweep for the prolongation above it
  double *xn64 = nullptr;
  float *xn32 = nullptr;
};

struct MGState {
  std::vector<MGLevel> lv;
  int small_from = 0;   // first level handled by the one-workgroup kernel
  int nu = 2;
  std::string shape;
  int split_levels = 0; // levels 0 .. split_levels - 1 are colour-split (they run the split kernels)
  int dense_n = 0;      // > 0: the coarsest level (dense_n cells) is solved by its dense inverse
  double *cinv64 = nullptr;
  float *cinv32 = nullptr;
};

namespace {

template <typename T>
int mg_alloc(T **p, size_t n) {
  *p = nullptr;
  hipError_t e = hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(CFD_ERR_NOMEM, std::string("multigrid: hipMalloc failed: ") + hipGetErrorString(e));
  }
  return CFD_OK;
}

template <typename T>
void mg_free(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

__device__ inline int64_t idx3(const MGView &L, int i, int j, int k) { return ((int64_t)i * L.ny + j) * L.nz + k; }

// the z index one step down / up, across the seam with periodic z (meaningful
// where that neighbour exists)
__device__ inline int kdn(const MGView &L, int k) { return k > 0 ? k - 1 : L.nz - 1; }
__device__ inline int kup(const MGView &L, int k) { return k < L.nz - 1 ? k + 1 : 0; }

// a cell's own index and its neighbours' (meaningful where the neighbour exists)
struct Nb {
  int64_t c, xm, xp, ym, yp, zm, zp;
};

// Layouts of a level's x and b.  Natural: (i ny + j) nz + k.
struct Natural {
  __device__ static int64_t at(const MGView &L, int i, int j, int k) { return idx3(L, i, j, k); }
  __device__ static Nb nbrs(const MGView &L, int i, int j, int k) {
    const int64_t c = idx3(L, i, j, k), sx = (int64_t)L.ny * L.nz, sy = L.nz;
    return Nb{c, c - sx, c + sx, c - sy, c + sy, k > 0 ? c - 1 : c + (L.nz - 1), k < L.nz - 1 ? c + 1 : c - (L.nz - 1)};
  }
};

// ColourSplit (the finest level of a multi-level cycle): the red cells
// (i + j + k even) first, then the black ones, each colour at
// (i ny + j) hz + k/2 with hz = ceil(nz / 2).  A colour sweep streams its
// own half and reads the other at unit stride instead of touching every line
// of both at stride 2.  For odd nz one colour has an unused last slot per
// row.  The six neighbours of a cell are the other colour's entries at
// h -+ ny hz, h -+ hz and, along k, h or h -+ 1 by the parity of k.
struct ColourSplit {
  __device__ static int64_t half(const MGView &L) { return (int64_t)L.nx * L.ny * ((L.nz + 1) >> 1); }
  __device__ static int64_t at(const MGView &L, int i, int j, int k) {
    const int hz = (L.nz + 1) >> 1;
    const int64_t h = ((int64_t)i * L.ny + j) * hz + (k >> 1);
    return ((i + j + k) & 1) ? half(L) + h : h;
  }
  __device__ static Nb nbrs(const MGView &L, int i, int j, int k) {
    const int hz = (L.nz + 1) >> 1;
    const int64_t h = ((int64_t)i * L.ny + j) * hz + (k >> 1), ns = half(L);
    const bool black = (i + j + k) & 1;
    const int64_t own = black ? ns + h : h, oth = black ? h : ns + h, sx = (int64_t)L.ny * hz;
    // across the periodic seam (nz even): k = 0's -z neighbour is k = nz - 1
    // at k/2 = hz - 1, and k = nz - 1's +z neighbour k = 0 at k/2 = 0
    const int64_t zm = k > 0 ? oth - ((k & 1) ? 0 : 1) : oth + (hz - 1);
    const int64_t zp = k < L.nz - 1 ? oth + (k & 1) : oth - (hz - 1);
    return Nb{own, oth - sx, oth + sx, oth - hz, oth + hz, zm, zp};
  }
};

template <typename T>
__device__ inline T mg_diag(const MGView &L, int i, int j, int k) {
  return (T)(L.ax[2 * L.nx + i] + L.ay[2 * L.ny + j] + L.az[2 * L.nz + k]);
}

// which of the six neighbours exist
struct Has {
  bool xm, xp, ym, yp, zm, zp;
};
__device__ inline Has has6(const MGView &L, int i, int j, int k) {
  return Has{i > 0, i < L.nx - 1, j > 0, j < L.ny - 1, L.wz || k > 0, L.wz || k < L.nz - 1};
}

// the six neighbour coefficients (an absent neighbour's entry is 0 and in range)
template <typename T>
struct Coef6 {
  T xm, xp, ym, yp, zm, zp;
};
template <typename T>
__device__ inline Coef6<T> coef6(const MGView &L, int i, int j, int k) {
  return Coef6<T>{(T)L.ax[i], (T)L.ax[L.nx + i], (T)L.ay[j], (T)L.ay[L.ny + j], (T)L.az[k], (T)L.az[L.nz + k]};
}

// sum over the neighbours of coefficient * x, in the order -x +x -y +y -z
// +z.  Every load is issued before the sum (an absent neighbour reads the
// cell itself and is left out by a select), so a wave has all of them in
// flight at once instead of one exec-masked load and wait per neighbour.
template <typename T>
__device__ inline T mg_off(const MGView &L, const T *__restrict__ x, int i, int j, int k, const Nb &n) {
  const Has h = has6(L, i, j, k);
  const T vxm = x[h.xm ? n.xm : n.c], vxp = x[h.xp ? n.xp : n.c], vym = x[h.ym ? n.ym : n.c];
  const T vyp = x[h.yp ? n.yp : n.c], vzm = x[h.zm ? n.zm : n.c], vzp = x[h.zp ? n.zp : n.c];
  const Coef6<T> a = coef6<T>(L, i, j, k);
  T s = (T)0;
  s = h.xm ? s + a.xm * vxm : s;
  s = h.xp ? s + a.xp * vxp : s;
  s = h.ym ? s + a.ym * vym : s;
  s = h.yp ? s + a.yp * vyp : s;
  s = h.zm ? s + a.zm * vzm : s;
  s = h.zp ? s + a.zp * vzp : s;
  return s;
}

template <typename T>
__device__ inline double mg_vol(const MGView &L, int i, int j, int k) {
  return L.ax[3 * L.nx + i] * L.ay[3 * L.ny + j] * L.az[3 * L.nz + k];
}

// one colour of a red-black Gauss-Seidel sweep at (i, j, k2), k = 2 k2 +
// parity (colour 0 = red: i + j + k even)
template <typename T, typename Lay>
__device__ inline void rb_at(const MGView &L, int i, int j, int k2, int colour, const T *__restrict__ b,
                             T *__restrict__ x) {
  const int k = 2 * k2 + ((i + j + colour) & 1);
  if (k >= L.nz) return;
  const Nb n = Lay::nbrs(L, i, j, k);
  x[n.c] = (b[n.c] + mg_off<T>(L, x, i, j, k, n)) / mg_diag<T>(L, i, j, k);
}

// parent of fine cell (i, j, k) on the next coarser level
__device__ inline int64_t parent(const MGView &F, const MGView &C, int i, int j, int k) {
  return idx3(C, F.fx == 2 ? i >> 1 : i, F.fy == 2 ? j >> 1 : j, F.fz == 2 ? k >> 1 : k);
}

// the first red sweep after the coarse correction, with the prolongation
// folded in: the black neighbours are read as x + xc[parent].  Red cells are
// overwritten here and black cells by the black sweep that follows, so the
// corrected values are only ever needed as neighbours.
template <typename T, typename Lay>
__device__ inline void rb_prolong_at(const MGView &F, const MGView &C, int i, int j, int k2, const T *__restrict__ b,
                                     T *__restrict__ x, const T *__restrict__ xc) {
  const int k = 2 * k2 + ((i + j) & 1);
  if (k >= F.nz) return;
  const Nb n = Lay::nbrs(F, i, j, k);
  const Has h = has6(F, i, j, k);
  // all loads first (absent neighbours read the cell and its parent), as in mg_off
  const T vxm = x[h.xm ? n.xm : n.c] + xc[parent(F, C, h.xm ? i - 1 : i, j, k)];
  const T vxp = x[h.xp ? n.xp : n.c] + xc[parent(F, C, h.xp ? i + 1 : i, j, k)];
  const T vym = x[h.ym ? n.ym : n.c] + xc[parent(F, C, i, h.ym ? j - 1 : j, k)];
  const T vyp = x[h.yp ? n.yp : n.c] + xc[parent(F, C, i, h.yp ? j + 1 : j, k)];
  const T vzm = x[h.zm ? n.zm : n.c] + xc[parent(F, C, i, j, h.zm ? kdn(F, k) : k)];
  const T vzp = x[h.zp ? n.zp : n.c] + xc[parent(F, C, i, j, h.zp ? kup(F, k) : k)];
  const Coef6<T> a = coef6<T>(F, i, j, k);
  T s = (T)0;
  s = h.xm ? s + a.xm * vxm : s;
  s = h.xp ? s + a.xp * vxp : s;
  s = h.ym ? s + a.ym * vym : s;
  s = h.yp ? s + a.yp * vyp : s;
  s = h.zm ? s + a.zm * vzm : s;
  s = h.zp ? s + a.zp * vzp : s;
  x[n.c] = (b[n.c] + s) / mg_diag<T>(F, i, j, k);
}

// a new right-hand side at cell (i, j, k) and the first red half-sweep from x = 0
template <typename T, typename Lay>
__device__ inline void set_rhs(const MGView &L, int i, int j, int k, T bv, T *__restrict__ b, T *__restrict__ x) {
  const int64_t c = Lay::at(L, i, j, k);
  b[c] = bv;
  x[c] = ((i + j + k) & 1) == 0 ? bv / mg_diag<T>(L, i, j, k) : (T)0;
}

// residual of the children of coarse cell (I, J, K), volume-weighted into
// the coarse right-hand side (which starts its own first red half-sweep);
// the coarse level is always in natural layout
template <typename T, typename Lay>
__device__ inline double child_res(const MGView &F, int i, int j, int k, const T *__restrict__ bf,
                                   const T *__restrict__ xf) {
  const Nb n = Lay::nbrs(F, i, j, k);
  const T r = bf[n.c] - (mg_diag<T>(F, i, j, k) * xf[n.c] - mg_off<T>(F, xf, i, j, k, n));
  return mg_vol<T>(F, i, j, k) * (double)r;
}

template <typename T, typename Lay>
__device__ inline void restrict_at(const MGView &F, const MGView &C, int I, int J, int K, const T *__restrict__ bf,
                                   const T *__restrict__ xf, T *__restrict__ bc, T *__restrict__ xc) {
  double acc = 0.0;
  if (F.fx == 2 && F.fy == 2 && F.fz == 2) {
#pragma unroll
    for (int d = 0; d < 8; ++d)
      acc += child_res<T, Lay>(F, 2 * I + (d >> 2), 2 * J + ((d >> 1) & 1), 2 * K + (d & 1), bf, xf);
  } else {
    for (int di = 0; di < F.fx; ++di)
      for (int dj = 0; dj < F.fy; ++dj)
        for (int dk = 0; dk < F.fz; ++dk)
          acc += child_res<T, Lay>(F, I * F.fx + di, J * F.fy + dj, K * F.fz + dk, bf, xf);
  }
  set_rhs<T, Natural>(C, I, J, K, (T)(acc / mg_vol<T>(C, I, J, K)), bc, xc);
}

// b = r / jp on the finest level (r and jp in natural layout, b and x in Lay)
template <typename T, typename V, typename Lay>
__global__ void __launch_bounds__(256) mg_scale_kernel(MGView L, const V *__restrict__ r, const double *__restrict__ jp,
                                                       T *__restrict__ b, T *__restrict__ x, const KrylovScalars *ks) {
  if (ks && ks->done) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
  if (k >= L.nz || j >= L.ny) return;
  const int64_t c = idx3(L, i, j, k);
  set_rhs<T, Lay>(L, i, j, k, (T)((double)r[c] / jp[c]), b, x);
}

template <typename T, typename V>
__global__ void __launch_bounds__(256) mg_widen_kernel(int64_t n, const T *__restrict__ x, V *__restrict__ out,
                                                       const KrylovScalars *ks) {
  if (ks && ks->done) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = (V)x[t];
}

// 3-D launches (x: k or k2, y: j, z: i): no 64-bit index division per thread
template <typename T, typename Lay>
__global__ void __launch_bounds__(256) mg_rb_kernel(MGView L, int colour, const T *__restrict__ b, T *__restrict__ x,
                                                    const KrylovScalars *ks) {
  if (ks && ks->done) return;
  const int k2 = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
  if (k2 >= (L.nz + 1) / 2 || j >= L.ny) return;
  rb_at<T, Lay>(L, i, j, k2, colour, b, x);
}

// planes of the colour-split finest level that one thread of a sweep walks:
// at 4 the FP64 sweeps hold 72-112 VGPRs (4-7 waves a SIMD; 8 planes: 112-182,
// 2-4 waves), and the 512x256x64 multigrid steps ran 1-1.5% faster than at 8
// (2 planes: 5% slower)
constexpr int MG_PLANES = 4;

enum SplitSweep { SWEEP_PLAIN, SWEEP_PROLONG, SWEEP_OUT };

// One colour of a red-black sweep of the colour-split finest level.  Thread
// (k2, j) walks the planes i0 .. i0 + MG_PLANES - 1 (i0 = MG_PLANES
// blockIdx.z) and updates the cell of the colour at (i, j, k2) of each.  The
// other colour's entry at (i', j, k2), cen(i'), is loaded once: it is the -x
// neighbour of plane i' + 1, the +x neighbour of plane i' - 1 and one of the
// two k neighbours of plane i' (the other is the entry at k2 -+ 1).  The y
// and z coefficients are loaded once per thread (the z ones for both
// parities of k, which alternate from plane to plane), the x ones are
// plane-uniform.  The arithmetic is rb_at's, term for term.
//   SWEEP_PROLONG: the first red half-sweep after the coarse correction: the
//     neighbours are read as x + xc[parent], as in rb_prolong_at.
//   SWEEP_OUT: the last black half-sweep of the finest level (or of a
//     colour-split coarse level, whose natural copy the prolongation above
//     reads); it writes the new black value and the red one (its cen) to
//     out in natural layout, precision V.
template <typename T, typename V, int MODE, typename CLay = Natural>
__global__ void __launch_bounds__(256) mg_split_sweep_kernel(MGView L, MGView C, int colour, const T *__restrict__ b,
                                                             T *__restrict__ x, const T *__restrict__ xc,
                                                             V *__restrict__ out, const KrylovScalars *ks) {
  if (ks && ks->done) return;
  const int k2 = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
  const int hz = (L.nz + 1) >> 1;
  if (k2 >= hz || j >= L.ny) return;
  const int i0 = blockIdx.z * MG_PLANES;
  const int64_t ns = (int64_t)L.nx * L.ny * hz, sx = (int64_t)L.ny * hz;
  T *__restrict__ own = x + (colour ? ns : 0);
  const T *__restrict__ oth = x + (colour ? 0 : ns);
  const T *__restrict__ bown = b + (colour ? ns : 0);
  const int64_t col = (int64_t)j * hz + k2;  // position of (0, j, k2) in a colour
  const bool hym = j > 0, hyp = j < L.ny - 1;
  const T cym = (T)L.ay[j], cyp = (T)L.ay[L.ny + j];
  const double dy = L.ay[2 * L.ny + j];
  T czm[2], czp[2];
  double dz[2];
#pragma unroll
  for (int e = 0; e < 2; ++e) {
    const int k = min(2 * k2 + e, L.nz - 1);  // the clamped entry of an absent cell is never used
    czm[e] = (T)L.az[k];
    czp[e] = (T)L.az[L.nz + k];
    dz[e] = L.az[2 * L.nz + k];
  }
  // the parent's indices along y and z of a neighbour at (j', k'); x is per plane
  auto par = [&](int ip, int jp, int kp) -> int64_t {
    return CLay::at(C, L.fx == 2 ? ip >> 1 : ip, L.fy == 2 ? jp >> 1 : jp, L.fz == 2 ? kp >> 1 : kp);
  };
  T cen[MG_PLANES + 2];
#pragma unroll
  for (int p = 0; p < MG_PLANES + 2; ++p) {
    const int ip = min(max(i0 - 1 + p, 0), L.nx - 1);  // clamped planes are never used
    cen[p] = oth[ip * sx + col];
    if (MODE == SWEEP_PROLONG) {
      // the other colour's cell at (ip, j, k2)
      const int kb = 2 * k2 + ((ip + j + colour + 1) & 1);
      cen[p] += xc[par(ip, j, min(kb, L.nz - 1))];
    }
  }
  // every plane's loads are issued before any result is stored (planes past
  // the grid and the padding slot of odd nz load clamped entries and store
  // nothing), so the wave has them all in flight at once
  T vym[MG_PLANES], vyp[MG_PLANES], side[MG_PLANES], bv[MG_PLANES];
#pragma unroll
  for (int p = 0; p < MG_PLANES; ++p) {
    const int i = min(i0 + p, L.nx - 1);
    const int e = (i + j + colour) & 1;
    const int k = min(2 * k2 + e, L.nz - 1);
    const int64_t h = i * sx + col;
    // the k neighbour that is not cen: k2 - 1 below an even k, k2 + 1 above
    // an odd one; across the periodic seam k2 = hz - 1 below k = 0 and
    // k2 = 0 above k = nz - 1
    const bool hside = L.wz || (e ? k < L.nz - 1 : k > 0);
    const int64_t hs = e ? (k < L.nz - 1 ? h + 1 : h - (hz - 1)) : (k > 0 ? h - 1 : h + (hz - 1));
    vym[p] = oth[hym ? h - hz : h];
    vyp[p] = oth[hyp ? h + hz : h];
    side[p] = oth[hside ? hs : h];
    bv[p] = bown[h];
    if (MODE == SWEEP_PROLONG) {
      vym[p] += xc[par(i, hym ? j - 1 : j, k)];
      vyp[p] += xc[par(i, hyp ? j + 1 : j, k)];
      side[p] += xc[par(i, j, hside ? (e ? kup(L, k) : kdn(L, k)) : k)];
    }
  }
#pragma unroll
  for (int p = 0; p < MG_PLANES; ++p) {
    const int i = i0 + p;
    const int e = (i + j + colour) & 1;
    const int k = 2 * k2 + e;
    if (MODE == SWEEP_OUT && i < L.nx && 2 * k2 + 1 - e < L.nz) out[idx3(L, i, j, 2 * k2 + 1 - e)] = (V)cen[p + 1];
    if (i >= L.nx || k >= L.nz) continue;
    const bool hxm = i > 0, hxp = i < L.nx - 1, hzm = L.wz || k > 0, hzp = L.wz || k < L.nz - 1;
    const T vzm = e ? cen[p + 1] : side[p], vzp = e ? side[p] : cen[p + 1];
    const T cxm = (T)L.ax[i], cxp = (T)L.ax[L.nx + i];
    T s = (T)0;
    s = hxm ? s + cxm * cen[p] : s;
    s = hxp ? s + cxp * cen[p + 2] : s;
    s = hym ? s + cym * vym[p] : s;
    s = hyp ? s + cyp * vyp[p] : s;
    s = hzm ? s + czm[e] * vzm : s;
    s = hzp ? s + czp[e] * vzp : s;
    const T v = (bv[p] + s) / (T)(L.ax[2 * L.nx + i] + dy + dz[e]);
    if (MODE == SWEEP_OUT)
      out[idx3(L, i, j, k)] = (V)v;
    else
      own[i * sx + col] = v;
  }
}

// Restriction from the colour-split finest level when every axis is halved.
// The children of (I, J, K) are the entries at position K of the four (i, j)
// columns 2I..2I+1 x 2J..2J+1, one of each colour per column (k = 2K and
// 2K + 1).  The kernel loads those pairs, the pairs of the eight columns
// beside them and the single entries below (2K - 1) and above (2K + 2) each
// column (40 loads for 8 children, with b), and sums the children's
// residuals in restrict_at's order with child_res's arithmetic.
template <typename T, typename CLay = Natural>
__global__ void __launch_bounds__(256) mg_restrict_split_kernel(MGView F, MGView C, const T *__restrict__ bf,
                                                                const T *__restrict__ xf, T *__restrict__ bc,
                                                                T *__restrict__ xc, const KrylovScalars *ks) {
  if (ks && ks->done) return;
  const int K = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y * blockDim.y + threadIdx.y, I = blockIdx.z;
  if (K >= C.nz || J >= C.ny) return;
  const int hz = F.nz >> 1;  // nz is even where z is halved
  const int64_t ns = (int64_t)F.nx * F.ny * hz;
  // entry at k = 2 KK + m (m = 0, 1) of column (i, j), both clamped into the
  // grid (KK wraps across the periodic seam)
  auto ld = [&](const T *a, int i, int j, int KK, int m) -> T {
    i = min(max(i, 0), F.nx - 1);
    j = min(max(j, 0), F.ny - 1);
    if (F.wz) KK = KK < 0 ? hz - 1 : (KK >= hz ? 0 : KK);
    KK = min(max(KK, 0), hz - 1);
    return a[(((i + j + m) & 1) ? ns : 0) + ((int64_t)i * F.ny + j) * hz + KK];
  };
  // X[a][c] = the pair of column (2I - 1 + a, 2J - 1 + c): the four children's
  // columns (a, c in 1..2) and the eight beside them (no corners)
  T X[4][4][2], B[2][2][2], xkm[2][2], xkp[2][2];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int m = 0; m < 2; ++m)
        X[a][c][m] = ((a == 0 || a == 3) && (c == 0 || c == 3)) ? (T)0 : ld(xf, 2 * I - 1 + a, 2 * J - 1 + c, K, m);
#pragma unroll
  for (int di = 0; di < 2; ++di)
#pragma unroll
    for (int dj = 0; dj < 2; ++dj) {
      const int i = 2 * I + di, j = 2 * J + dj;
      B[di][dj][0] = ld(bf, i, j, K, 0);
      B[di][dj][1] = ld(bf, i, j, K, 1);
      xkm[di][dj] = ld(xf, i, j, K - 1, 1);  // k = 2K - 1
      xkp[di][dj] = ld(xf, i, j, K + 1, 0);  // k = 2K + 2
    }
  double acc = 0.0;
#pragma unroll
  for (int d = 0; d < 8; ++d) {
    const int di = d >> 2, dj = (d >> 1) & 1, m = d & 1;
    const int i = 2 * I + di, j = 2 * J + dj, k = 2 * K + m;
    const Has h = has6(F, i, j, k);
    const Coef6<T> a = coef6<T>(F, i, j, k);
    const T xv = X[di + 1][dj + 1][m];
    const T vzm = m ? X[di + 1][dj + 1][0] : xkm[di][dj], vzp = m ? xkp[di][dj] : X[di + 1][dj + 1][1];
    T s = (T)0;
    s = h.xm ? s + a.xm * X[di][dj + 1][m] : s;
    s = h.xp ? s + a.xp * X[di + 2][dj + 1][m] : s;
    s = h.ym ? s + a.ym * X[di + 1][dj][m] : s;
    s = h.yp ? s + a.yp * X[di + 1][dj + 2][m] : s;
    s = h.zm ? s + a.zm * vzm : s;
    s = h.zp ? s + a.zp * vzp : s;
    const T r = B[di][dj][m] - (mg_diag<T>(F, i, j, k) * xv - s);
    acc += mg_vol<T>(F, i, j, k) * (double)r;
  }
  set_rhs<T, CLay>(C, I, J, K, (T)(acc / mg_vol<T>(C, I, J, K)), bc, xc);
}

template <typename T, typename Lay>
__global__ void __launch_bounds__(256) mg_rb_prolong_kernel(MGView F, MGView C, const T *__restrict__ b,
                                                            T *__restrict__ x, const T *__restrict__ xc,
                                                            const KrylovScalars *ks) {
  if (ks && ks->done) return;
  const int k2 = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
  if (k2 >= (F.nz + 1) / 2 || j >= F.ny) return;
  rb_prolong_at<T, Lay>(F, C, i, j, k2, b, x, xc);
}

template <typename T, typename Lay>
__global__ void __launch_bounds__(256) mg_restrict_kernel(MGView F, MGView C, const T *__restrict__ bf,
                                                          const T *__restrict__ xf, T *__restrict__ bc,
                                                          T *__restrict__ xc, const KrylovScalars *ks) {
  if (ks && ks->done) return;
  const int K = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y * blockDim.y + threadIdx.y, I = blockIdx.z;
  if (K >= C.nz || J >= C.ny) return;
  restrict_at<T, Lay>(F, C, I, J, K, bf, xf, bc, xc);
}

constexpr int MG_SMALL_THREADS = 512;

// dense-inverse entries a thread of the bottom kernel prefetches: with the
// levels in LDS and an inverse of at most DENSE_PF * MG_SMALL_THREADS
// entries (64 cells), the whole inverse is loaded with the first level's x and
// b and staged through the LDS left over, instead of 16+ dependent round
// trips to memory at the coarsest solve
constexpr int DENSE_PF = 8;

template <typename T>
struct SmallArgs {
  int nlev, nu, coarse_sweeps, dense;
  int stage_off, stage_cols;   // dense-inverse staging in LDS: element offset, columns a chunk (0: none)
  int lds_off[MG_MAX_LEVELS];  // with the levels in LDS: element offset of level l's x (its b follows)
  MGView v[MG_MAX_LEVELS];
  T *x[MG_MAX_LEVELS];
  T *b[MG_MAX_LEVELS];
  const T *cinv;  // dense inverse of the coarsest operator, column-major, when dense
};

// fills lds_off and, when the dense inverse is prefetched, stage_off /
// stage_cols; the LDS bytes to launch with (0: the levels do not fit in a
// workgroup's default 64 KiB and stay in global memory)
template <typename T>
size_t small_lds_layout(SmallArgs<T> &S) {
  size_t elems = 0;
  for (int l = 0; l < S.nlev; ++l) {
    S.lds_off[l] = (int)elems;
    elems += 2 * (size_t)S.v[l].n;
  }
  S.stage_off = 0;
  S.stage_cols = 0;
  if (elems * sizeof(T) > 65536) return 0;
  const int64_t n = S.v[S.nlev - 1].n;
  if (S.dense && n * n <= (int64_t)DENSE_PF * MG_SMALL_THREADS) {
    const int64_t cols = std::min<int64_t>(n, (65536 / (int64_t)sizeof(T) - (int64_t)elems) / n);
    if (cols >= 1) {
      S.stage_off = (int)elems;
      S.stage_cols = (int)cols;
      elems += (size_t)(cols * n);
    }
  }
  return elems * sizeof(T);
}

// the bottom of the V-cycle (levels of at most MG_SMALL cells) in one
// workgroup; __syncthreads separates the phases.  When they fit, every
// level's x and b live in LDS for the whole launch: the first level's
// right-hand side and first red half-sweep (written by the launch before)
// are read in at the start and its x is written back at the end, the only
// values the rest of the cycle uses.  Otherwise they stay in global memory
// (L2-resident).  The coarsest level is solved by its dense inverse when it
// is nonsingular, else smoothed coarse_sweeps times.
template <typename T, bool LDS>
__global__ void __launch_bounds__(MG_SMALL_THREADS) mg_small_kernel(SmallArgs<T> S, const KrylovScalars *ks) {
  if (ks && ks->done) return;
  const int nt = MG_SMALL_THREADS;
  extern __shared__ __align__(16) unsigned char small_lds[];
  T *const lds = reinterpret_cast<T *>(small_lds);
  // compile-time choice, so that the LDS accesses compile to LDS instructions
  auto X = [&](int l) -> T * {
    if constexpr (LDS) return lds + S.lds_off[l];
    else return S.x[l];
  };
  auto B = [&](int l) -> T * {
    if constexpr (LDS) return lds + S.lds_off[l] + S.v[l].n;
    else return S.b[l];
  };
  // with the levels in LDS: the first level's x and b (at most MG_SMALL
  // cells, 8 a thread) and, with staging, the dense inverse are loaded
  // before the first LDS store, one round trip to memory for all of them
  T pf[DENSE_PF];
  const int nn = S.stage_cols > 0 ? (int)(S.v[S.nlev - 1].n * S.v[S.nlev - 1].n) : 0;
  if (LDS) {
    constexpr int CP = (int)(MG_SMALL / MG_SMALL_THREADS);
    const int n0 = (int)S.v[0].n;
    T xv[CP], bv[CP];
#pragma unroll
    for (int u = 0; u < CP; ++u) {
      const int t = threadIdx.x + u * nt;
      xv[u] = t < n0 ? S.x[0][t] : (T)0;
      bv[u] = t < n0 ? S.b[0][t] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < DENSE_PF; ++u) {
      const int e = threadIdx.x + u * nt;
      pf[u] = e < nn ? S.cinv[e] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < CP; ++u) {
      const int t = threadIdx.x + u * nt;
      if (t < n0) {
        X(0)[t] = xv[u];
        B(0)[t] = bv[u];
      }
    }
    for (int t = threadIdx.x + CP * nt; t < n0; t += nt) {  // not taken: the levels hold at most MG_SMALL cells
      X(0)[t] = S.x[0][t];
      B(0)[t] = S.b[0][t];
    }
    __syncthreads();
  }
  // nsw sweeps of level l; skip_red0: the first red half-sweep is done;
  // prolong_from >= 0: the first red half-sweep folds in that level's correction
  auto sweep = [&](int l, int nsw, bool skip_red0, int prolong_from) {
    const MGView &L = S.v[l];
    const int half = (L.nz + 1) / 2;
    const int cnt = L.nx * L.ny * half;
    T *xl = X(l);
    const T *bl = B(l);
    for (int s = 0; s < nsw; ++s)
      for (int colour = 0; colour < 2; ++colour) {
        if (s == 0 && colour == 0 && skip_red0) continue;
        const bool fold = s == 0 && colour == 0 && prolong_from >= 0;
        for (int t = threadIdx.x; t < cnt; t += nt) {
          const int k2 = t % half, r = t / half;
          if (fold)
            rb_prolong_at<T, Natural>(L, S.v[prolong_from], r / L.ny, r % L.ny, k2, bl, xl, X(prolong_from));
          else
            rb_at<T, Natural>(L, r / L.ny, r % L.ny, k2, colour, bl, xl);
        }
        __syncthreads();
      }
  };
  const int last = S.nlev - 1;
  for (int l = 0; l < last; ++l) {
    sweep(l, S.nu, true, -1);
    const MGView &C = S.v[l + 1];
    for (int t = threadIdx.x; t < (int)C.n; t += nt) {
      const int K = t % C.nz, r = t / C.nz;
      restrict_at<T, Natural>(S.v[l], C, r / C.ny, r % C.ny, K, B(l), X(l), B(l + 1), X(l + 1));
    }
    __syncthreads();
  }
  if (S.dense && LDS && S.stage_cols > 0) {
    // the prefetched inverse (entry e = q n + t in thread e % nt) through
    // the LDS stage, stage_cols columns at a time; thread t < n sums its row
    // over q = 0 .. n - 1 in order, as below
    const int n = (int)S.v[last].n, sc = S.stage_cols;
    const T *bl = B(last);
    T *xl = X(last);
    T *stage = lds + S.stage_off;
    double acc = 0.0;
    for (int q0 = 0; q0 < n; q0 += sc) {
      const int q1 = min(n, q0 + sc);
#pragma unroll
      for (int u = 0; u < DENSE_PF; ++u) {
        const int e = threadIdx.x + u * nt;
        if (e >= q0 * n && e < q1 * n) stage[e - q0 * n] = pf[u];
      }
      __syncthreads();
      if ((int)threadIdx.x < n)
        for (int q = q0; q < q1; ++q) acc += (double)stage[(q - q0) * n + threadIdx.x] * (double)bl[q];
      __syncthreads();
    }
    if ((int)threadIdx.x < n) xl[threadIdx.x] = (T)acc;
    __syncthreads();
  } else if (S.dense) {
    // consecutive threads read consecutive entries of a column
    const int n = (int)S.v[last].n;
    const T *bl = B(last);
    T *xl = X(last);
    for (int t = threadIdx.x; t < n; t += nt) {
      double acc = 0.0;
      for (int q = 0; q < n; ++q) acc += (double)S.cinv[(int64_t)q * n + t] * (double)bl[q];
      xl[t] = (T)acc;
    }
    __syncthreads();
  } else {
    sweep(last, S.coarse_sweeps, true, -1);
  }
  for (int l = last - 1; l >= 0; --l) sweep(l, S.nu, false, l + 1);
  if (LDS)
    for (int t = threadIdx.x; t < (int)S.v[0].n; t += nt) S.x[0][t] = X(0)[t];
}

// launch shape of a 3-D kernel over (nk, ny, nx) with 256-thread blocks
struct Shape3 {
  dim3 grid, block;
};
Shape3 shape3(int nk, int ny, int nx) {
  int bx = 1;
  while (bx < nk && bx < 64) bx <<= 1;
  const int by = 256 / bx;
  Shape3 s;
  s.block = dim3(bx, by, 1);
  s.grid = dim3((nk + bx - 1) / bx, (ny + by - 1) / by, nx);
  return s;
}

int nblk(int64_t n) { return blocks_for(n, 256, 1 << 30); }

template <typename T>
T *lx(MGLevel &L);
template <>
double *lx<double>(MGLevel &L) { return L.x64; }
template <>
float *lx<float>(MGLevel &L) { return L.x32; }
template <typename T>
T *lxn(MGLevel &L);
template <>
double *lxn<double>(MGLevel &L) { return L.xn64; }
template <>
float *lxn<float>(MGLevel &L) { return L.xn32; }
template <typename T>
T *lb(MGLevel &L);
template <>
double *lb<double>(MGLevel &L) { return L.b64; }
template <>
float *lb<float>(MGLevel &L) { return L.b32; }

// level l's kernels in layout Lay
template <typename T, typename Lay>
struct LevelOps {
  template <typename V>
  static void scale_from(const MGView &v, const V *r, const double *jp, T *b, T *x, const KrylovScalars *ks,
                         hipStream_t st) {
    const Shape3 sh = shape3(v.nz, v.ny, v.nx);
    mg_scale_kernel<T, V, Lay><<<sh.grid, sh.block, 0, st>>>(v, r, jp, b, x, ks);
  }
  static void rb(const MGView &v, int colour, const T *b, T *x, const KrylovScalars *ks, hipStream_t st) {
    const Shape3 sh = shape3((v.nz + 1) / 2, v.ny, v.nx);
    mg_rb_kernel<T, Lay><<<sh.grid, sh.block, 0, st>>>(v, colour, b, x, ks);
  }
  static void rb_prolong(const MGView &v, const MGView &C, const T *b, T *x, const T *xc, const KrylovScalars *ks,
                         hipStream_t st) {
    const Shape3 sh = shape3((v.nz + 1) / 2, v.ny, v.nx);
    mg_rb_prolong_kernel<T, Lay><<<sh.grid, sh.block, 0, st>>>(v, C, b, x, xc, ks);
  }
  static void restrict_to(const MGView &v, const MGView &C, const T *b, const T *x, T *bc, T *xc,
                          const KrylovScalars *ks, hipStream_t st) {
    const Shape3 sh = shape3(C.nz, C.ny, C.nx);
    mg_restrict_kernel<T, Lay><<<sh.grid, sh.block, 0, st>>>(v, C, b, x, bc, xc, ks);
  }
};

// out = V-cycle(L, r / jp): the cycle runs in T, the vectors r and out are V
// (an FP32 cycle inside an FP64 Krylov solve widens its result at the end).
// When the finest level is smoothed by the level kernels (the cycle has more
// than the one-workgroup bottom), its x and b are colour-split and its last
// black half-sweep writes out; otherwise the bottom kernel works on out
// directly (or on x, widened into out).
template <typename T, typename V>
int mg_apply(Ctx &c, const V *r, V *out, const KrylovScalars *ks) {
  MGState &M = *c.mg;
  const int nl = (int)M.lv.size();
  const int ls = M.small_from;
  const bool split = ls > 0;
  T *b0 = lb<T>(M.lv[0]);
  constexpr bool same = std::is_same<T, V>::value;
  T *x0 = (same && !split) ? reinterpret_cast<T *>(out) : lx<T>(M.lv[0]);
  if (!b0 || !x0) return fail(CFD_ERR_STATE, "multigrid buffers of this precision are not allocated");
  auto X = [&](int l) { return l == 0 ? x0 : lx<T>(M.lv[l]); };
  if (split)
    LevelOps<T, ColourSplit>::scale_from(M.lv[0].v, r, c.f[CFD_JP], b0, x0, ks, c.stream);
  else
    LevelOps<T, Natural>::scale_from(M.lv[0].v, r, c.f[CFD_JP], b0, x0, ks, c.stream);
  // nu sweeps of level l: pre-smoothing skips the red half-sweep its
  // right-hand side's producer did; post-smoothing folds the prolongation
  // into its first red half-sweep, and on the finest level ends in the
  // half-sweep that writes out
  auto smooth = [&](int l, bool pre) {
    const MGView &v = M.lv[l].v;
    for (int s = 0; s < M.nu; ++s)
      for (int colour = 0; colour < 2; ++colour) {
        const bool first = s == 0 && colour == 0, last = s == M.nu - 1 && colour == 1;
        if (first && pre) continue;
        if (l < M.split_levels) {
          Shape3 sh = shape3((v.nz + 1) / 2, v.ny, v.nx);
          sh.grid.z = (v.nx + MG_PLANES - 1) / MG_PLANES;
          const MGView &C = M.lv[l + 1].v;
          T *bl = l == 0 ? b0 : lb<T>(M.lv[l]), *xl = X(l);
          if (first && l + 1 < M.split_levels)  // the correction from a split level's natural copy
            mg_split_sweep_kernel<T, T, SWEEP_PROLONG><<<sh.grid, sh.block, 0, c.stream>>>(v, C, 0, bl, xl,
                                                                                          lxn<T>(M.lv[l + 1]),
                                                                                          nullptr, ks);
          else if (last && !pre && l > 0)  // a split level's last half-sweep writes its natural copy
            mg_split_sweep_kernel<T, T, SWEEP_OUT><<<sh.grid, sh.block, 0, c.stream>>>(v, C, 1, bl, xl, nullptr,
                                                                                      lxn<T>(M.lv[l]), ks);
          else if (first)
            mg_split_sweep_kernel<T, T, SWEEP_PROLONG><<<sh.grid, sh.block, 0, c.stream>>>(v, C, 0, bl, xl, X(l + 1),
                                                                                          nullptr, ks);
          else if (last && !pre && l == 0)
            mg_split_sweep_kernel<T, V, SWEEP_OUT><<<sh.grid, sh.block, 0, c.stream>>>(v, C, 1, bl, xl, nullptr, out,
                                                                                      ks);
          else
            mg_split_sweep_kernel<T, T, SWEEP_PLAIN><<<sh.grid, sh.block, 0, c.stream>>>(v, C, colour, bl, xl,
                                                                                        nullptr, nullptr, ks);
        } else {
          using Ops = LevelOps<T, Natural>;
          if (first)
            Ops::rb_prolong(v, M.lv[l + 1].v, lb<T>(M.lv[l]), X(l), X(l + 1), ks, c.stream);
          else
            Ops::rb(v, colour, lb<T>(M.lv[l]), X(l), ks, c.stream);
        }
      }
  };
  for (int l = 0; l < ls; ++l) {
    smooth(l, true);
    const MGView &F = M.lv[l].v;
    const bool fsplit = l < M.split_levels;
    T *bf = l == 0 ? b0 : lb<T>(M.lv[l]);
    if (fsplit && F.fx == 2 && F.fy == 2 && F.fz == 2) {
      const MGView &C = M.lv[l + 1].v;
      const Shape3 sh = shape3(C.nz, C.ny, C.nx);
      if (l + 1 < M.split_levels)  // the coarse right-hand side in its colour-split layout
        mg_restrict_split_kernel<T, ColourSplit><<<sh.grid, sh.block, 0, c.stream>>>(
            F, C, bf, X(l), lb<T>(M.lv[l + 1]), X(l + 1), ks);
      else
        mg_restrict_split_kernel<T><<<sh.grid, sh.block, 0, c.stream>>>(F, C, bf, X(l), lb<T>(M.lv[l + 1]),
                                                                        X(l + 1), ks);
    } else if (fsplit)  // a colour-split level into a natural one, cell by cell
      LevelOps<T, ColourSplit>::restrict_to(F, M.lv[l + 1].v, bf, X(l), lb<T>(M.lv[l + 1]), X(l + 1), ks, c.stream);
    else
      LevelOps<T, Natural>::restrict_to(M.lv[l].v, M.lv[l + 1].v, lb<T>(M.lv[l]), X(l), lb<T>(M.lv[l + 1]),
                                        X(l + 1), ks, c.stream);
  }
  SmallArgs<T> S;
  std::memset(&S, 0, sizeof(S));
  S.nlev = nl - ls;
  S.nu = M.nu;
  S.coarse_sweeps = MG_COARSE_SWEEPS;
  S.dense = M.dense_n > 0 ? 1 : 0;
  S.cinv = std::is_same<T, double>::value ? reinterpret_cast<const T *>(M.cinv64) : reinterpret_cast<const T *>(M.cinv32);
  for (int l = ls; l < nl; ++l) {
    S.v[l - ls] = M.lv[l].v;
    S.x[l - ls] = X(l);
    S.b[l - ls] = lb<T>(M.lv[l]);
  }
  const size_t lds_bytes = small_lds_layout(S);
  if (lds_bytes)
    mg_small_kernel<T, true><<<1, MG_SMALL_THREADS, lds_bytes, c.stream>>>(S, ks);
  else
    mg_small_kernel<T, false><<<1, MG_SMALL_THREADS, 0, c.stream>>>(S, ks);
  for (int l = ls - 1; l >= 0; --l) smooth(l, false);
  if (!same && !split) mg_widen_kernel<T, V><<<nblk(c.g.N), 256, 0, c.stream>>>(c.g.N, x0, out, ks);
  CFD_HIP(hipGetLastError());
  return CFD_OK;
}

// per-axis level arrays from widths w (n) and the centre distances of the
// finest level (ds_lo: centre of cell i-1 to i, for i = 0 the ghost
// distance; ds_hi: centre of i to i+1, for i = n-1 the ghost distance)
// (periodic: the end cells keep their coefficient to the neighbour across
// the seam, whose distances dlo[0] and dhi[n-1] give)
void axis_arrays(const std::vector<double> &w, const std::vector<double> &dlo, const std::vector<double> &dhi,
                 bool dir_lo, bool dir_hi, std::vector<double> &out, bool periodic = false) {
  const size_t n = w.size();
  out.assign(4 * n, 0.0);
  for (size_t i = 0; i < n; ++i) {
    const double xm = 1.0 / w[i] / dlo[i], xp = 1.0 / w[i] / dhi[i];
    out[i] = (i > 0 || periodic) ? xm : 0.0;
    out[n + i] = (i + 1 < n || periodic) ? xp : 0.0;
    out[2 * n + i] = ((i > 0 || dir_lo || periodic) ? xm : 0.0) + ((i + 1 < n || dir_hi || periodic) ? xp : 0.0);
    out[3 * n + i] = w[i];
  }
}

}  // namespace

void mg_release(Ctx &c) {
  if (!c.mg) return;
  for (auto &L : c.mg->lv) {
    mg_free(L.coef);
    mg_free(L.x64);
    mg_free(L.b64);
    mg_free(L.x32);
    mg_free(L.b32);
    mg_free(L.xn64);
    mg_free(L.xn32);
  }
  mg_free(c.mg->cinv64);
  mg_free(c.mg->cinv32);
  delete c.mg;
  c.mg = nullptr;
  mg_free(c.ph64);
  mg_free(c.sh64);
  mg_free(c.ph32);
  mg_free(c.sh32);
}

// builds the hierarchy (once) and the buffers of the requested precision
int mg_ensure(Ctx &c, bool fp32) {
  if (!c.mg) {
    MGState *M = new MGState();
    c.mg = M;
    const int nc[3] = {c.g.nxc, c.g.nyc, c.g.nzc};
    const std::vector<double> *Dh[3] = {&c.hDx, &c.hDy, &c.hDz}, *Dsh[3] = {&c.hDxs, &c.hDys, &c.hDzs};
    std::vector<double> w[3], dlo[3], dhi[3];
    for (int a = 0; a < 3; ++a) {
      for (int q = 0; q < nc[a]; ++q) {
        const int I = q + GC;
        w[a].push_back((*Dh[a])[I]);
        dlo[a].push_back((*Dsh[a])[I - 1]);
        dhi[a].push_back((*Dsh[a])[I]);
      }
    }
    // periodic z: every level wraps in z (the fine level's seam coefficient
    // is the operator's, from the ghost distances Dzs)
    const bool pz = c.prm.periodic_z != 0;
    const bool dl[3] = {c.prm.p_dirichlet[0] != 0, c.prm.p_dirichlet[2] != 0, !pz && c.prm.p_dirichlet[4] != 0};
    const bool dh[3] = {c.prm.p_dirichlet[1] != 0, c.prm.p_dirichlet[3] != 0, !pz && c.prm.p_dirichlet[5] != 0};
    std::vector<double> last_arr[3];
    for (;;) {
      MGLevel L;
      L.v.nx = (int)w[0].size();
      L.v.ny = (int)w[1].size();
      L.v.nz = (int)w[2].size();
      L.v.n = (int64_t)L.v.nx * L.v.ny * L.v.nz;
      L.v.wz = pz ? 1 : 0;
      std::vector<double> arr[3], all;
      for (int a = 0; a < 3; ++a) {
        axis_arrays(w[a], dlo[a], dhi[a], dl[a], dh[a], arr[a], a == 2 && pz);
        all.insert(all.end(), arr[a].begin(), arr[a].end());
        last_arr[a] = arr[a];
      }
      int rc = mg_alloc(&L.coef, all.size());
      if (rc) return rc;
      CFD_HIP(hipMemcpy(L.coef, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice));
      L.v.ax = L.coef;
      L.v.ay = L.coef + 4 * L.v.nx;
      L.v.az = L.coef + 4 * L.v.nx + 4 * L.v.ny;
      int f[3];
      for (int a = 0; a < 3; ++a) f[a] = (w[a].size() % 2 == 0 && w[a].size() >= 4) ? 2 : 1;
      if (pz && (w[2].size() / 2) % 2 != 0) f[2] = 1;  // periodic z stays even
      const bool more = (f[0] > 1 || f[1] > 1 || f[2] > 1) && L.v.n > 64 && (int)M->lv.size() < MG_MAX_LEVELS - 1;
      L.v.fx = more ? f[0] : 1;
      L.v.fy = more ? f[1] : 1;
      L.v.fz = more ? f[2] : 1;
      M->lv.push_back(L);
      if (!more) break;
      // merged widths; centre distances from the widths, mirror ghosts at Dirichlet faces
      for (int a = 0; a < 3; ++a) {
        if (f[a] == 1) continue;
        std::vector<double> wc;
        for (size_t q = 0; q < w[a].size(); q += 2) wc.push_back(w[a][q] + w[a][q + 1]);
        const size_t n = wc.size();
        std::vector<double> lo(n), hi(n);
        const double seam = 0.5 * (wc[n - 1] + wc[0]);  // centre distance across a periodic seam
        for (size_t q = 0; q < n; ++q) {
          lo[q] = q > 0 ? 0.5 * (wc[q - 1] + wc[q]) : (a == 2 && pz ? seam : wc[q]);
          hi[q] = q + 1 < n ? 0.5 * (wc[q] + wc[q + 1]) : (a == 2 && pz ? seam : wc[q]);
        }
        w[a] = wc;
        dlo[a] = lo;
        dhi[a] = hi;
      }
    }
    M->small_from = (int)M->lv.size() - 1;
    while (M->small_from > 0 && M->lv[M->small_from - 1].v.n <= MG_SMALL) M->small_from--;
    // the finest level is colour-split whenever it has level kernels; each
    // next one is too while the level above halves every axis (its split
    // restriction then writes the coarse right-hand side in that layout)
    M->split_levels = M->small_from > 0 ? 1 : 0;
    while (M->split_levels < M->small_from && M->split_levels < MG_SPLIT_MAX) {
      const MGView &P = M->lv[M->split_levels - 1].v;
      if (!(P.fx == 2 && P.fy == 2 && P.fz == 2)) break;
      M->split_levels++;
    }
    // exact coarsest solve: the dense inverse of the coarsest operator when a
    // Dirichlet face makes it nonsingular (all-Neumann keeps the sweeps)
    const MGView &Lc = M->lv.back().v;
    const bool any_dir = dl[0] || dl[1] || dl[2] || dh[0] || dh[1] || dh[2];
    if (any_dir && Lc.n <= 512) {
      const int n = (int)Lc.n, nx = Lc.nx, ny = Lc.ny, nz = Lc.nz;
      std::vector<double> A((size_t)n * n, 0.0), Inv((size_t)n * n, 0.0);
      const std::vector<double> &ax = last_arr[0], &ay = last_arr[1], &az = last_arr[2];
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
          for (int k = 0; k < nz; ++k) {
            const int r = (i * ny + j) * nz + k;
            A[(size_t)r * n + r] = ax[2 * nx + i] + ay[2 * ny + j] + az[2 * nz + k];
            if (i > 0) A[(size_t)r * n + r - ny * nz] -= ax[i];
            if (i < nx - 1) A[(size_t)r * n + r + ny * nz] -= ax[nx + i];
            if (j > 0) A[(size_t)r * n + r - nz] -= ay[j];
            if (j < ny - 1) A[(size_t)r * n + r + nz] -= ay[ny + j];
            // z neighbours, across the seam with periodic z (with nz = 2 both
            // couplings of a cell reach the same neighbour and add)
            if (k > 0 || Lc.wz) A[(size_t)r * n + r - k + (k > 0 ? k - 1 : nz - 1)] -= az[k];
            if (k < nz - 1 || Lc.wz) A[(size_t)r * n + r - k + (k < nz - 1 ? k + 1 : 0)] -= az[nz + k];
          }
      for (int r = 0; r < n; ++r) Inv[(size_t)r * n + r] = 1.0;
      bool ok = true;
      for (int col = 0; col < n && ok; ++col) {  // Gauss-Jordan with partial pivoting
        int piv = col;
        for (int r = col + 1; r < n; ++r)
          if (std::fabs(A[(size_t)r * n + col]) > std::fabs(A[(size_t)piv * n + col])) piv = r;
        if (!(std::fabs(A[(size_t)piv * n + col]) > 0.0)) { ok = false; break; }
        if (piv != col)
          for (int q = 0; q < n; ++q) {
            std::swap(A[(size_t)piv * n + q], A[(size_t)col * n + q]);
            std::swap(Inv[(size_t)piv * n + q], Inv[(size_t)col * n + q]);
          }
        const double d = A[(size_t)col * n + col];
        for (int q = 0; q < n; ++q) { A[(size_t)col * n + q] /= d; Inv[(size_t)col * n + q] /= d; }
        for (int r = 0; r < n; ++r) {
          if (r == col) continue;
          const double f = A[(size_t)r * n + col];
          if (f == 0.0) continue;
          for (int q = 0; q < n; ++q) {
            A[(size_t)r * n + q] -= f * A[(size_t)col * n + q];
            Inv[(size_t)r * n + q] -= f * Inv[(size_t)col * n + q];
          }
        }
      }
      if (ok) {
        // column-major: the kernel's threads (one per row) read a column together
        std::vector<double> InvT((size_t)n * n);
        for (int r = 0; r < n; ++r)
          for (int q = 0; q < n; ++q) InvT[(size_t)q * n + r] = Inv[(size_t)r * n + q];
        Inv.swap(InvT);
        std::vector<float> inv32(Inv.begin(), Inv.end());
        int rc = mg_alloc(&M->cinv64, Inv.size());
        if (!rc) rc = mg_alloc(&M->cinv32, inv32.size());
        if (rc) return rc;
        CFD_HIP(hipMemcpy(M->cinv64, Inv.data(), Inv.size() * sizeof(double), hipMemcpyHostToDevice));
        CFD_HIP(hipMemcpy(M->cinv32, inv32.data(), inv32.size() * sizeof(float), hipMemcpyHostToDevice));
        M->dense_n = n;
      }
    }
    M->nu = c.prm.mg_nu > 0 ? c.prm.mg_nu : 4;
    for (auto &L : M->lv) M->shape += std::to_string(L.v.nx) + "x" + std::to_string(L.v.ny) + "x" +
                                      std::to_string(L.v.nz) + " ";
  }
  MGState &M = *c.mg;
  M.nu = c.prm.mg_nu > 0 ? c.prm.mg_nu : 4;
  for (size_t l = 0; l < M.lv.size(); ++l) {
    MGLevel &L = M.lv[l];
    // colour-split vectors hold nx ny ceil(nz / 2) cells per colour
    const bool sp = (int)l < M.split_levels;
    const size_t n = sp ? (size_t)L.v.nx * L.v.ny * (2 * ((L.v.nz + 1) / 2)) : (size_t)L.v.n;
    const bool nat = sp && l > 0;  // a split level below the finest keeps a natural copy of x
    int rc = CFD_OK;
    if (fp32) {
      if (!L.b32 && (rc = mg_alloc(&L.b32, n))) return rc;
      if (!L.x32 && (rc = mg_alloc(&L.x32, n))) return rc;
      if (nat && !L.xn32 && (rc = mg_alloc(&L.xn32, (size_t)L.v.n))) return rc;
    } else {
      if (!L.b64 && (rc = mg_alloc(&L.b64, n))) return rc;
      if (!L.x64 && (rc = mg_alloc(&L.x64, n))) return rc;
      if (nat && !L.xn64 && (rc = mg_alloc(&L.xn64, (size_t)L.v.n))) return rc;
    }
  }
  const size_t N = (size_t)c.g.N;
  int rc = CFD_OK;
  if (fp32) {
    if ((!c.ph32 && (rc = mg_alloc(&c.ph32, N))) || (!c.sh32 && (rc = mg_alloc(&c.sh32, N)))) return rc;
  } else {
    if ((!c.ph64 && (rc = mg_alloc(&c.ph64, N))) || (!c.sh64 && (rc = mg_alloc(&c.sh64, N)))) return rc;
  }
  return CFD_OK;
}

int mg_apply64(Ctx &c, const double *r, double *out, const KrylovScalars *ks) {
  return mg_apply<double, double>(c, r, out, ks);
}
int mg_apply32(Ctx &c, const float *r, float *out, const KrylovScalars *ks) { return mg_apply<float, float>(c, r, out, ks); }
int mg_apply64_fp32cycle(Ctx &c, const double *r, double *out, const KrylovScalars *ks) {
  return mg_apply<float, double>(c, r, out, ks);
}

const char *mg_shape(const Ctx &c) { return c.mg ? c.mg->shape.c_str() : ""; }

}  // namespace cfd
