// The LDS-tiled QUICK predictor's phases (csrc/pred_tiled.h) run on the host,
// one thread at a time between the kernel's barriers, against the oracle's
// predictor (orc_predictor_les / orc_predictor_bl).  Built with contraction
// off, the kernel's arithmetic (every face gradient, QUICK weight, flux and
// sum with the cell-by-cell form's operands in its order) must give the
// oracle's bits in every cell; the LDS starts as NaN, so a read of a slot no
// phase wrote shows up as a mismatch.
//
// Usage: pred_tiled_host <path of liboracle.so>; exit 0 when every case is
// bit-identical.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <random>
#include <vector>

#include "../../cfd_nssolver_forhpc_amd/csrc/pred_tiled.h"

using namespace cfd;

typedef void (*pred_les_t)(const cfd_grid_desc *, int, double, double, const double *, const double *, const double *,
                           const double *, double *, double *, double *);
typedef void (*pred_bl_t)(const cfd_grid_desc *, const cfd_params *, const double *, const double *, const double *,
                          const double *, double *, double *, double *);
typedef int (*axis_tanh_t)(int, double, double, double, double, double, double *, double *, double *);
typedef void (*axis_uni_t)(int, double, double *, double *, double *);

template <int TK>
static void emulate(const DevGrid &g, const PtShape &s, const PtArgs &a) {
  const int nblk = s.njt * s.nkt * s.nit;
  std::vector<double> lds(pt_lds_doubles(s));
  std::vector<PtCol> col(PT_THREADS);
  typedef double Regs[3][PtGeom<TK>::NPT];
  std::vector<Regs> regs(PT_THREADS);
  for (int b = 0; b < nblk; ++b) {
    std::fill(lds.begin(), lds.end(), std::nan(""));
    for (auto &c : col) std::memset(&c, 0xff, sizeof c);  // NaN: unwritten column values show up
    const PtLds L = pt_lds<TK>(lds.data());
    const PtBlock B = pt_block(g, s, b);
    for (int t = 0; t < PT_THREADS; ++t) {
      pt_fetch<TK>(g, B, a.u, a.v, a.w, B.ib, t, regs[t]);
      if (pt_active<TK>(g, B, t)) pt_col_init(g, a.u, a.v, a.w, B.ib, B.j0 + t / TK, B.k0 + t % TK, col[t]);
      pt_tables<TK>(g, B, L, t);
      pt_store<TK>(L, B.ib, t, regs[t]);
    }
    for (int i = B.ib; i < B.ie; ++i) {
      const bool more = i + 1 < B.ie;
      for (int t = 0; t < PT_THREADS; ++t)
        if (more) pt_fetch<TK>(g, B, a.u, a.v, a.w, i + 1, t, regs[t]);
      for (int t = 0; t < PT_THREADS; ++t) pt_grads<TK>(g, B, L, i, t);
      for (int t = 0; t < PT_THREADS; ++t)
        if (more) pt_store<TK>(L, i + 1, t, regs[t]);
      for (int t = 0; t < PT_THREADS; ++t) {
        const bool act = pt_active<TK>(g, B, t);
        const int j = B.j0 + t / TK, k = B.k0 + t % TK;
        double nxt[3];
        if (act && more) pt_col_fetch(g, a.u, a.v, a.w, i, j, k, nxt);
        if (act) pt_cell<TK>(g, B, L, a, col[t], i, t);
        if (act && more) pt_col_shift(B, L, i, nxt, col[t]);
      }
    }
  }
}

static int run_case(void *lib, int nxc, int nyc, int nzc, bool stretched, bool periodic, bool les, int src_mode) {
  auto pred_les = (pred_les_t)dlsym(lib, "orc_predictor_les");
  auto pred_bl = (pred_bl_t)dlsym(lib, "orc_predictor_bl");
  auto axis_tanh = (axis_tanh_t)dlsym(lib, "orc_axis_tanh");
  auto axis_uni = (axis_uni_t)dlsym(lib, "orc_axis_uniform");
  const int nx = nxc + 4, ny = nyc + 4, nz = nzc + 4;
  std::vector<double> X(nx + 1), Y(ny + 1), Z(nz + 1), Dx(nx), Dy(ny), Dz(nz), Dxs(nx), Dys(ny), Dzs(nz);
  if (stretched) {
    if (axis_tanh(nxc, 20.0, 5.0, 20.0 / nxc * 0.5, 2.0, 1.05, X.data(), Dx.data(), Dxs.data())) return 2;
    if (axis_tanh(nyc, 10.0, 5.0, 10.0 / nyc * 0.5, 2.0, 1.05, Y.data(), Dy.data(), Dys.data())) return 2;
  } else {
    axis_uni(nxc, 20.0, X.data(), Dx.data(), Dxs.data());
    axis_uni(nyc, 10.0, Y.data(), Dy.data(), Dys.data());
  }
  axis_uni(nzc, 2.5, Z.data(), Dz.data(), Dzs.data());
  cfd_grid_desc gd{nxc, nyc, nzc, X.data(), Y.data(), Z.data(), Dx.data(), Dy.data(), Dz.data(),
                   Dxs.data(), Dys.data(), Dzs.data(), 20.0, 10.0, 2.5};
  DevGrid g{};
  g.nx = nx; g.ny = ny; g.nz = nz; g.nxc = nxc; g.nyc = nyc; g.nzc = nzc;
  g.G = (int64_t)nx * ny * nz; g.N = (int64_t)nxc * nyc * nzc;
  g.X = X.data(); g.Y = Y.data(); g.Z = Z.data(); g.Dx = Dx.data(); g.Dy = Dy.data(); g.Dz = Dz.data();
  g.Dxs = Dxs.data(); g.Dys = Dys.data(); g.Dzs = Dzs.data();
  std::mt19937_64 rng(nxc * 1000003 + nyc * 1009 + nzc);
  std::normal_distribution<double> nd;
  std::vector<double> u(g.G), v(g.G), w(g.G), vis(les ? g.N : 0), src(3 * g.G);
  for (auto &x : u) x = 1.0 + 0.3 * nd(rng);
  for (auto &x : v) x = 0.3 * nd(rng);
  for (auto &x : w) x = 0.3 * nd(rng);
  for (auto &x : vis) x = 1e-3 * std::fabs(nd(rng));
  for (auto &x : src) x = nd(rng);
  const double Re = 200.0, dt = 5e-4, nu = 1.0 / Re;
  const double sentinel = 12345.678;
  std::vector<double> ou(g.G, sentinel), ov(g.G, sentinel), ow(g.G, sentinel);
  if (periodic) {
    cfd_params P{};
    P.Re = Re; P.dt = dt; P.periodic_z = 1; P.convection = CFD_CONV_QUICK;
    pred_bl(&gd, &P, les ? vis.data() : nullptr, u.data(), v.data(), w.data(), ou.data(), ov.data(), ow.data());
  } else {
    pred_les(&gd, CFD_CONV_QUICK, nu, dt, les ? vis.data() : nullptr, u.data(), v.data(), w.data(), ou.data(),
             ov.data(), ow.data());
  }
  // the emulated kernel
  std::vector<double> u1(g.G, sentinel), v1(g.G, sentinel), w1(g.G, sentinel);
  std::vector<double> pu(g.G, sentinel), pv(g.G, sentinel), pw(g.G, sentinel);
  PtArgs a{};
  a.nu = nu; a.dt = dt; a.u = u.data(); a.v = v.data(); a.w = w.data();
  a.u1 = u1.data(); a.v1 = v1.data(); a.w1 = w1.data();
  a.src = src.data(); a.src_mode = src_mode;
  a.pu = src_mode ? pu.data() : nullptr; a.pv = src_mode ? pv.data() : nullptr; a.pw = src_mode ? pw.data() : nullptr;
  a.vis = les ? vis.data() : nullptr;
  a.kw_end = periodic ? nz - GC : nz - GC - 1;
  const PtShape s = pt_shape(g);
  switch (s.tk) {
    case 4: emulate<4>(g, s, a); break;
    case 8: emulate<8>(g, s, a); break;
    case 16: emulate<16>(g, s, a); break;
    case 32: emulate<32>(g, s, a); break;
    default: emulate<64>(g, s, a); break;
  }
  const int nblk = s.njt * s.nkt * s.nit;
  // compare: the oracle writes the predictor's positions (sentinel elsewhere);
  // with src_mode 1 the kernel adds src[t] after keeping the prediction
  int64_t bad = 0, written = 0;
  double maxrel = 0.0;
  const double *O[3] = {ou.data(), ov.data(), ow.data()}, *K[3] = {u1.data(), v1.data(), w1.data()};
  const double *P3[3] = {pu.data(), pv.data(), pw.data()};
  for (int f = 0; f < 3; ++f)
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const int64_t ic = icel(g, i, j, k);
          const double o = O[f][ic];
          double kv = K[f][ic];
          if (src_mode == 1 && kv != sentinel) {
            kv = P3[f][ic];  // the kept prediction (before the source)
            const int64_t t = icelcal(g, i, j, k);
            if (K[f][ic] != kv + src[(size_t)f * g.N + t]) ++bad;
          }
          if (o == sentinel && kv == sentinel) continue;
          ++written;
          if (o == sentinel || kv == sentinel) {
            ++bad;
            continue;
          }
          if (o != kv) {
            ++bad;
            maxrel = std::fmax(maxrel, std::fabs(o - kv) / std::fmax(1e-300, std::fabs(o)));
          }
        }
  std::printf("%4d x %4d x %4d stretched %d periodic %d les %d src %d  tk %2d tj %3d blocks %6d lds %6zu B  written %9lld  mismatches %lld (max rel %.3g)\n",
              nxc, nyc, nzc, stretched, periodic, les, src_mode, s.tk, s.tj, nblk, pt_lds_doubles(s) * 8,
              (long long)written, (long long)bad, maxrel);
  return bad ? 1 : 0;
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::printf("usage: %s <liboracle.so>\n", argv[0]);
    return 2;
  }
  void *lib = dlopen(argv[1], RTLD_NOW);
  if (!lib) {
    std::printf("oracle: %s\n", dlerror());
    return 2;
  }
  int rc = 0;
  const int cases[][3] = {{40, 20, 6}, {64, 32, 16}, {128, 64, 4}, {33, 17, 40}, {24, 12, 70}, {17, 9, 2},
                          {48, 24, 12}, {20, 300, 3}, {64, 48, 64}, {40, 24, 32}, {70, 33, 64}};
  for (auto &c : cases)
    for (int st = 0; st < 2; ++st)
      for (int per = 0; per < 2; ++per) {
        if (per && (c[2] & 1)) continue;
        rc |= run_case(lib, c[0], c[1], c[2], st, per, (c[0] + per) % 2 == 0, 0);
      }
  rc |= run_case(lib, 40, 20, 6, true, false, true, 1);
  rc |= run_case(lib, 33, 17, 40, false, true, false, 1);
  std::printf(rc ? "FAIL\n" : "all cases bit-identical to the oracle\n");
  return rc;
}
