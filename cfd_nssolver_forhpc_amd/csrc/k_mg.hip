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
// Dirichlet face keeps its mirror-ghost coefficient).  Restriction is the
// volume-weighted sum of the children's residuals, prolongation adds the
// parent's correction.  Smoother: red-black Gauss-Seidel, nu sweeps before
// and after.  Levels of at most MG_SMALL cells run inside one 1024-thread
// workgroup (one launch for the whole bottom of the cycle); the coarsest is
// smoothed MG_COARSE_SWEEPS times.
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

// per-level device view: ax = cm[nx] cp[nx] dd[nx] w[nx] (and ay, az)
struct MGView {
  int nx, ny, nz, fx, fy, fz;
  int64_t n;
  const double *ax, *ay, *az;
};

struct MGLevel {
  MGView v{};
  double *coef = nullptr;               // ax | ay | az
  double *x64 = nullptr, *b64 = nullptr;
  float *x32 = nullptr, *b32 = nullptr;
};

struct MGState {
  std::vector<MGLevel> lv;
  int small_from = 0;   // first level handled by the one-workgroup kernel
  int nu = 2;
  std::string shape;
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

struct Cell3 {
  int i, j, k;
};

__device__ inline Cell3 dec3(const MGView &L, int64_t t) {
  Cell3 c;
  c.k = (int)(t % L.nz);
  const int64_t r = t / L.nz;
  c.j = (int)(r % L.ny);
  c.i = (int)(r / L.ny);
  return c;
}

__device__ inline int64_t idx3(const MGView &L, int i, int j, int k) { return ((int64_t)i * L.ny + j) * L.nz + k; }

template <typename T>
__device__ inline T mg_diag(const MGView &L, int i, int j, int k) {
  return (T)(L.ax[2 * L.nx + i] + L.ay[2 * L.ny + j] + L.az[2 * L.nz + k]);
}

// sum over the neighbours of coefficient * x
template <typename T>
__device__ inline T mg_off(const MGView &L, const T *__restrict__ x, int i, int j, int k, int64_t c) {
  const int64_t sx = (int64_t)L.ny * L.nz, sy = L.nz;
  T s = (T)0;
  if (i > 0) s += (T)L.ax[i] * x[c - sx];
  if (i < L.nx - 1) s += (T)L.ax[L.nx + i] * x[c + sx];
  if (j > 0) s += (T)L.ay[j] * x[c - sy];
  if (j < L.ny - 1) s += (T)L.ay[L.ny + j] * x[c + sy];
  if (k > 0) s += (T)L.az[k] * x[c - 1];
  if (k < L.nz - 1) s += (T)L.az[L.nz + k] * x[c + 1];
  return s;
}

template <typename T>
__device__ inline double mg_vol(const MGView &L, int i, int j, int k) {
  return L.ax[3 * L.nx + i] * L.ay[3 * L.ny + j] * L.az[3 * L.nz + k];
}

// one colour of a red-black Gauss-Seidel sweep at (i, j, k2), k = 2 k2 +
// parity (colour 0 = red: i + j + k even)
template <typename T>
__device__ inline void rb_at(const MGView &L, int i, int j, int k2, int colour, const T *__restrict__ b,
                             T *__restrict__ x) {
  const int k = 2 * k2 + ((i + j + colour) & 1);
  if (k >= L.nz) return;
  const int64_t c = idx3(L, i, j, k);
  x[c] = (b[c] + mg_off<T>(L, x, i, j, k, c)) / mg_diag<T>(L, i, j, k);
}

// parent of fine cell (i, j, k) on the next coarser level
__device__ inline int64_t parent(const MGView &F, const MGView &C, int i, int j, int k) {
  return idx3(C, F.fx == 2 ? i >> 1 : i, F.fy == 2 ? j >> 1 : j, F.fz == 2 ? k >> 1 : k);
}

// the first red sweep after the coarse correction, with the prolongation
// folded in: the black neighbours are read as x + xc[parent].  Red cells are
// overwritten here and black cells by the black sweep that follows, so the
// corrected values are only ever needed as neighbours.
template <typename T>
__device__ inline void rb_prolong_at(const MGView &F, const MGView &C, int i, int j, int k2, const T *__restrict__ b,
                                     T *__restrict__ x, const T *__restrict__ xc) {
  const int k = 2 * k2 + ((i + j) & 1);
  if (k >= F.nz) return;
  const int64_t c = idx3(F, i, j, k);
  const int64_t sx = (int64_t)F.ny * F.nz, sy = F.nz;
  T s = (T)0;
  if (i > 0) s += (T)F.ax[i] * (x[c - sx] + xc[parent(F, C, i - 1, j, k)]);
  if (i < F.nx - 1) s += (T)F.ax[F.nx + i] * (x[c + sx] + xc[parent(F, C, i + 1, j, k)]);
  if (j > 0) s += (T)F.ay[j] * (x[c - sy] + xc[parent(F, C, i, j - 1, k)]);
  if (j < F.ny - 1) s += (T)F.ay[F.ny + j] * (x[c + sy] + xc[parent(F, C, i, j + 1, k)]);
  if (k > 0) s += (T)F.az[k] * (x[c - 1] + xc[parent(F, C, i, j, k - 1)]);
  if (k < F.nz - 1) s += (T)F.az[F.nz + k] * (x[c + 1] + xc[parent(F, C, i, j, k + 1)]);
  x[c] = (b[c] + s) / mg_diag<T>(F, i, j, k);
}

// a new right-hand side at cell c and the first red half-sweep from x = 0
template <typename T>
__device__ inline void set_rhs(const MGView &L, int i, int j, int k, T bv, T *__restrict__ b, T *__restrict__ x) {
  const int64_t c = idx3(L, i, j, k);
  b[c] = bv;
  x[c] = ((i + j + k) & 1) == 0 ? bv / mg_diag<T>(L, i, j, k) : (T)0;
}

// residual of the children of coarse cell (I, J, K), volume-weighted into
// the coarse right-hand side (which starts its own first red half-sweep)
template <typename T>
__device__ inline double child_res(const MGView &F, int i, int j, int k, const T *__restrict__ bf,
                                   const T *__restrict__ xf) {
  const int64_t c = idx3(F, i, j, k);
  const T r = bf[c] - (mg_diag<T>(F, i, j, k) * xf[c] - mg_off<T>(F, xf, i, j, k, c));
  return mg_vol<T>(F, i, j, k) * (double)r;
}

template <typename T>
__device__ inline void restrict_at(const MGView &F, const MGView &C, int I, int J, int K, const T *__restrict__ bf,
                                   const T *__restrict__ xf, T *__restrict__ bc, T *__restrict__ xc) {
  double acc = 0.0;
  if (F.fx == 2 && F.fy == 2 && F.fz == 2) {
#pragma unroll
    for (int d = 0; d < 8; ++d) acc += child_res<T>(F, 2 * I + (d >> 2), 2 * J + ((d >> 1) & 1), 2 * K + (d & 1), bf, xf);
  } else {
    for (int di = 0; di < F.fx; ++di)
      for (int dj = 0; dj < F.fy; ++dj)
        for (int dk = 0; dk < F.fz; ++dk)
          acc += child_res<T>(F, I * F.fx + di, J * F.fy + dj, K * F.fz + dk, bf, xf);
  }
  set_rhs<T>(C, I, J, K, (T)(acc / mg_vol<T>(C, I, J, K)), bc, xc);
}

template <typename T, typename V>
__global__ void __launch_bounds__(256) mg_scale_kernel(MGView L, const V *__restrict__ r, const double *__restrict__ jp,
                                                       T *__restrict__ b, T *__restrict__ x, const KrylovScalars *ks) {
  if (ks && ks->done) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
  if (k >= L.nz || j >= L.ny) return;
  const int64_t c = idx3(L, i, j, k);
  set_rhs<T>(L, i, j, k, (T)((double)r[c] / jp[c]), b, x);
}

template <typename T, typename V>
__global__ void __launch_bounds__(256) mg_widen_kernel(int64_t n, const T *__restrict__ x, V *__restrict__ out,
                                                       const KrylovScalars *ks) {
  if (ks && ks->done) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = (V)x[t];
}

// 3-D launches (x: k or k2, y: j, z: i): no 64-bit index division per thread
template <typename T>
__global__ void __launch_bounds__(256) mg_rb_kernel(MGView L, int colour, const T *__restrict__ b, T *__restrict__ x,
                                                    const KrylovScalars *ks) {
  if (ks && ks->done) return;
  const int k2 = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
  if (k2 >= (L.nz + 1) / 2 || j >= L.ny) return;
  rb_at<T>(L, i, j, k2, colour, b, x);
}

template <typename T>
__global__ void __launch_bounds__(256) mg_rb_prolong_kernel(MGView F, MGView C, const T *__restrict__ b,
                                                            T *__restrict__ x, const T *__restrict__ xc,
                                                            const KrylovScalars *ks) {
  if (ks && ks->done) return;
  const int k2 = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y, i = blockIdx.z;
  if (k2 >= (F.nz + 1) / 2 || j >= F.ny) return;
  rb_prolong_at<T>(F, C, i, j, k2, b, x, xc);
}

template <typename T>
__global__ void __launch_bounds__(256) mg_restrict_kernel(MGView F, MGView C, const T *__restrict__ bf,
                                                          const T *__restrict__ xf, T *__restrict__ bc,
                                                          T *__restrict__ xc, const KrylovScalars *ks) {
  if (ks && ks->done) return;
  const int K = blockIdx.x * blockDim.x + threadIdx.x, J = blockIdx.y * blockDim.y + threadIdx.y, I = blockIdx.z;
  if (K >= C.nz || J >= C.ny) return;
  restrict_at<T>(F, C, I, J, K, bf, xf, bc, xc);
}

constexpr int MG_SMALL_THREADS = 256;

template <typename T>
struct SmallArgs {
  int nlev, nu, coarse_sweeps, dense;
  MGView v[MG_MAX_LEVELS];
  T *x[MG_MAX_LEVELS];
  T *b[MG_MAX_LEVELS];
  const T *cinv;  // dense inverse of the coarsest operator (row-major), when dense
};

// the bottom of the V-cycle (levels of at most MG_SMALL cells) in one
// workgroup; __syncthreads separates the phases, whose data stays in this
// workgroup (L2-resident).  The first level's right-hand side and first red
// half-sweep were written by the launch before; the coarsest level is solved
// by its dense inverse when it is nonsingular, else smoothed coarse_sweeps
// times.
template <typename T>
__global__ void __launch_bounds__(MG_SMALL_THREADS) mg_small_kernel(SmallArgs<T> S, const KrylovScalars *ks) {
  if (ks && ks->done) return;
  const int nt = MG_SMALL_THREADS;
  // nsw sweeps of level l; skip_red0: the first red half-sweep is done;
  // prolong_from >= 0: the first red half-sweep folds in that level's correction
  auto sweep = [&](int l, int nsw, bool skip_red0, int prolong_from) {
    const MGView &L = S.v[l];
    const int half = (L.nz + 1) / 2;
    const int cnt = L.nx * L.ny * half;
    for (int s = 0; s < nsw; ++s)
      for (int colour = 0; colour < 2; ++colour) {
        if (s == 0 && colour == 0 && skip_red0) continue;
        const bool fold = s == 0 && colour == 0 && prolong_from >= 0;
        for (int t = threadIdx.x; t < cnt; t += nt) {
          const int k2 = t % half, r = t / half;
          if (fold)
            rb_prolong_at<T>(L, S.v[prolong_from], r / L.ny, r % L.ny, k2, S.b[l], S.x[l], S.x[prolong_from]);
          else
            rb_at<T>(L, r / L.ny, r % L.ny, k2, colour, S.b[l], S.x[l]);
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
      restrict_at<T>(S.v[l], C, r / C.ny, r % C.ny, K, S.b[l], S.x[l], S.b[l + 1], S.x[l + 1]);
    }
    __syncthreads();
  }
  if (S.dense) {
    const int n = (int)S.v[last].n;
    const T *bl = S.b[last];
    for (int t = threadIdx.x; t < n; t += nt) {
      const T *row = S.cinv + (int64_t)t * n;
      double acc = 0.0;
      for (int q = 0; q < n; ++q) acc += (double)row[q] * (double)bl[q];
      S.x[last][t] = (T)acc;
    }
    __syncthreads();
  } else {
    sweep(last, S.coarse_sweeps, true, -1);
  }
  for (int l = last - 1; l >= 0; --l) sweep(l, S.nu, false, l + 1);
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
T *lb(MGLevel &L);
template <>
double *lb<double>(MGLevel &L) { return L.b64; }
template <>
float *lb<float>(MGLevel &L) { return L.b32; }

// out = V-cycle(L, r / jp): the cycle runs in T, the vectors r and out are V
// (an FP32 cycle inside an FP64 Krylov solve widens its result at the end)
template <typename T, typename V>
int mg_apply(Ctx &c, const V *r, V *out, const KrylovScalars *ks) {
  MGState &M = *c.mg;
  const int nl = (int)M.lv.size();
  T *b0 = lb<T>(M.lv[0]);
  constexpr bool same = std::is_same<T, V>::value;
  T *x0 = same ? reinterpret_cast<T *>(out) : lx<T>(M.lv[0]);
  if (!b0 || !x0) return fail(CFD_ERR_STATE, "multigrid buffers of this precision are not allocated");
  auto X = [&](int l) { return l == 0 ? x0 : lx<T>(M.lv[l]); };
  {
    const MGView &v = M.lv[0].v;
    const Shape3 sh = shape3(v.nz, v.ny, v.nx);
    mg_scale_kernel<T, V><<<sh.grid, sh.block, 0, c.stream>>>(v, r, c.f[CFD_JP], b0, x0, ks);
  }
  // nu sweeps of level l: pre-smoothing skips the red half-sweep its
  // right-hand side's producer did; post-smoothing folds the prolongation
  // into its first red half-sweep
  auto smooth = [&](int l, bool pre) {
    const MGView &v = M.lv[l].v;
    const Shape3 sh = shape3((v.nz + 1) / 2, v.ny, v.nx);
    for (int s = 0; s < M.nu; ++s)
      for (int colour = 0; colour < 2; ++colour) {
        if (s == 0 && colour == 0) {
          if (pre) continue;
          mg_rb_prolong_kernel<T><<<sh.grid, sh.block, 0, c.stream>>>(v, M.lv[l + 1].v, lb<T>(M.lv[l]), X(l),
                                                                      X(l + 1), ks);
          continue;
        }
        mg_rb_kernel<T><<<sh.grid, sh.block, 0, c.stream>>>(v, colour, lb<T>(M.lv[l]), X(l), ks);
      }
  };
  const int ls = M.small_from;
  for (int l = 0; l < ls; ++l) {
    smooth(l, true);
    const MGView &C = M.lv[l + 1].v;
    const Shape3 sh = shape3(C.nz, C.ny, C.nx);
    mg_restrict_kernel<T><<<sh.grid, sh.block, 0, c.stream>>>(M.lv[l].v, C, lb<T>(M.lv[l]), X(l),
                                                               lb<T>(M.lv[l + 1]), X(l + 1), ks);
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
  mg_small_kernel<T><<<1, MG_SMALL_THREADS, 0, c.stream>>>(S, ks);
  for (int l = ls - 1; l >= 0; --l) smooth(l, false);
  if (!same) mg_widen_kernel<T, V><<<nblk(c.g.N), 256, 0, c.stream>>>(c.g.N, x0, out, ks);
  CFD_HIP(hipGetLastError());
  return CFD_OK;
}

// per-axis level arrays from widths w (n) and the centre distances of the
// finest level (ds_lo: centre of cell i-1 to i, for i = 0 the ghost
// distance; ds_hi: centre of i to i+1, for i = n-1 the ghost distance)
void axis_arrays(const std::vector<double> &w, const std::vector<double> &dlo, const std::vector<double> &dhi,
                 bool dir_lo, bool dir_hi, std::vector<double> &out) {
  const size_t n = w.size();
  out.assign(4 * n, 0.0);
  for (size_t i = 0; i < n; ++i) {
    const double xm = 1.0 / w[i] / dlo[i], xp = 1.0 / w[i] / dhi[i];
    out[i] = i > 0 ? xm : 0.0;
    out[n + i] = i + 1 < n ? xp : 0.0;
    out[2 * n + i] = ((i > 0 || dir_lo) ? xm : 0.0) + ((i + 1 < n || dir_hi) ? xp : 0.0);
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
    const bool dl[3] = {c.prm.p_dirichlet[0] != 0, c.prm.p_dirichlet[2] != 0, c.prm.p_dirichlet[4] != 0};
    const bool dh[3] = {c.prm.p_dirichlet[1] != 0, c.prm.p_dirichlet[3] != 0, c.prm.p_dirichlet[5] != 0};
    std::vector<double> last_arr[3];
    for (;;) {
      MGLevel L;
      L.v.nx = (int)w[0].size();
      L.v.ny = (int)w[1].size();
      L.v.nz = (int)w[2].size();
      L.v.n = (int64_t)L.v.nx * L.v.ny * L.v.nz;
      std::vector<double> arr[3], all;
      for (int a = 0; a < 3; ++a) {
        axis_arrays(w[a], dlo[a], dhi[a], dl[a], dh[a], arr[a]);
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
        for (size_t q = 0; q < n; ++q) {
          lo[q] = q > 0 ? 0.5 * (wc[q - 1] + wc[q]) : wc[q];
          hi[q] = q + 1 < n ? 0.5 * (wc[q] + wc[q + 1]) : wc[q];
        }
        w[a] = wc;
        dlo[a] = lo;
        dhi[a] = hi;
      }
    }
    M->small_from = (int)M->lv.size() - 1;
    while (M->small_from > 0 && M->lv[M->small_from - 1].v.n <= MG_SMALL) M->small_from--;
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
            if (i > 0) A[(size_t)r * n + r - ny * nz] = -ax[i];
            if (i < nx - 1) A[(size_t)r * n + r + ny * nz] = -ax[nx + i];
            if (j > 0) A[(size_t)r * n + r - nz] = -ay[j];
            if (j < ny - 1) A[(size_t)r * n + r + nz] = -ay[ny + j];
            if (k > 0) A[(size_t)r * n + r - 1] = -az[k];
            if (k < nz - 1) A[(size_t)r * n + r + 1] = -az[nz + k];
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
        std::vector<float> inv32(Inv.begin(), Inv.end());
        int rc = mg_alloc(&M->cinv64, Inv.size());
        if (!rc) rc = mg_alloc(&M->cinv32, inv32.size());
        if (rc) return rc;
        CFD_HIP(hipMemcpy(M->cinv64, Inv.data(), Inv.size() * sizeof(double), hipMemcpyHostToDevice));
        CFD_HIP(hipMemcpy(M->cinv32, inv32.data(), inv32.size() * sizeof(float), hipMemcpyHostToDevice));
        M->dense_n = n;
      }
    }
    M->nu = c.prm.mg_nu > 0 ? c.prm.mg_nu : 2;
    for (auto &L : M->lv) M->shape += std::to_string(L.v.nx) + "x" + std::to_string(L.v.ny) + "x" +
                                      std::to_string(L.v.nz) + " ";
  }
  MGState &M = *c.mg;
  M.nu = c.prm.mg_nu > 0 ? c.prm.mg_nu : 2;
  for (size_t l = 0; l < M.lv.size(); ++l) {
    MGLevel &L = M.lv[l];
    int rc = CFD_OK;
    if (fp32) {
      if (!L.b32 && (rc = mg_alloc(&L.b32, L.v.n))) return rc;
      if (!L.x32 && (rc = mg_alloc(&L.x32, L.v.n))) return rc;  // level 0 too: the FP64 solve's FP32 cycle
    } else {
      if (!L.b64 && (rc = mg_alloc(&L.b64, L.v.n))) return rc;
      if (l > 0 && !L.x64 && (rc = mg_alloc(&L.x64, L.v.n))) return rc;
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
