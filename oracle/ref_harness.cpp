// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Drives the reference solver's own header-only functions, compiled from the
// sources where they lie under /root/reference (see oracle/Makefile), at the
// reference's hard-coded native case (190x210x40 SML grid, z-cylinder at
// (6.5, 6), R = 0.5, BC_Cavity_A table, all-Neumann pressure, Re 40, dt 0.002)
// and dumps raw arrays for the parity tests:
//   grid.bin        X,Y,Z (n+1 each) then Dx,Dy,Dz,Dxs,Dys,Dzs (n each)
//   eta.bin         DFIB_CylinderZ result (G doubles)
//   ell_col.bin     createPressureMatrix ELL columns, slot-major int32 [7][N]
//   ell_val.bin     ELL values, slot-major [7][N]
//   jp.bin          gA.jp after createPressureMatrix (N)
//   spe_jp.bin      SPE_matrix1::set_Jp_pc (N)
//   csr_*.bin       createPressureMatrixCSR: ptr (N+1), idx, val (int32/fp64), its jp (N)
//   step_XXX_*.bin  state after each timestep (T0 u,v,w, p, X, f)
//   fn_*.bin        inputs/outputs of each hot-path function in pass 1 of step FN_STEP
//   stats.txt       per pass: step pass iters error cD cL, per step: maxdiff
//
// The time loop below calls the reference functions in the order of
// runSeqOmp_predictionMethod (inc/run/RUN_CPU_PC_METHOD.hpp:287-627); it
// restates only the sequencing, never the arithmetic.
//
// usage: ref_harness OUTDIR NSTEPS TOL OP(ell|spe) FN_STEP [DUMP_EVERY [SRC_AT_CELL]]
// (ref_harness_lud / ref_harness_ud: the same with the LUD / UD convection scheme;
//  ref_harness_les: with the Smagorinsky model, TERBULENCE_SMAGORINSKY)
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <numeric>
#include <random>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>
#include <omp.h>
#include <dirent.h>
#include <sys/stat.h>
#include <unistd.h>

// ELL_matrix keeps its (column, value) slots private; expose them to dump the
// slot layout without touching the reference sources.
#define private public
#include "controlPanel.hpp"
#if defined(HARNESS_CONV_LUD) || defined(HARNESS_CONV_UD)
// controlPanel.hpp:125 fixes the convection scheme to QUICK; these builds
// select another of the reference's own schemes (:127, :129) before
// run/general.hpp compiles ConvectionDifussion
#undef CONVECTION_DIFUSSION_QUICK
#if defined(HARNESS_CONV_LUD)
#define CONVECTION_DIFUSSION_LUD
#else
#define CONVECTION_DIFUSSION_UD
#endif
#endif
#include "run/general.hpp"
#undef private

static std::string g_out;

template <typename T>
static void dump(const std::string &name, const std::vector<T> &v) {
  std::ofstream f(g_out + "/" + name, std::ios::binary);
  f.write(reinterpret_cast<const char *>(v.data()), sizeof(T) * v.size());
}

static void dump_vel(const std::string &prefix, const velocity &V) {
  dump(prefix + "_u.bin", V.u);
  dump(prefix + "_v.bin", V.v);
  dump(prefix + "_w.bin", V.w);
}

// Seeded inputs shared with the oracle's tests (oracle/cfd_oracle.c orc_noise):
// value = base + amp * (2 u - 1), u from splitmix64(seed ^ (tag << 56) ^ index).
static uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static void fill_noise(std::vector<double> &v, uint64_t seed, int tag, double base, double amp) {
  for (size_t m = 0; m < v.size(); ++m) {
    const uint64_t z = splitmix64(seed ^ ((uint64_t)tag << 56) ^ (uint64_t)m);
    const double u01 = (double)(z >> 11) * 0x1.0p-53;
    volatile double nz = amp * (2.0 * u01 - 1.0);  // rounded before the add, never fused
    v[m] = base + nz;
  }
}

// usage: ref_harness OUTDIR rules -- bicgstabRe2's failure handling
// (bicgstabRe2.hpp:91-189) driven through the reference's own ELL_matrix and
// solver on a 200-row tridiagonal system (diagonal 2 + 0.01 (i mod 7), -1 off
// the diagonal, entries set in the order i-1, i, i+1):
//   nan:     solve b1 from x = 0, solve b2, then b1 with a NaN at row 7: every
//            entry of x diverges, and after iters_max the reset rule leaves
//            px2, the initial guess of the solve before (x1);
//   restart: ten solves at tol 1e-10, then from x = 0 a right-hand side of
//            1e-12 noise, whose |r . r0| is below 1e-7 tol from the start, so
//            the shadow residual restarts every iteration; the same solve in
//            a fresh solver (timestep_ 1) does not restart.
static int rules_main() {
  const int n = 200;
  const uint64_t seed = 20240611;
  ELL_type A;
  A.resize(n, 3);
  for (int i = 0; i < n; ++i) {
    if (i > 0) A.set(i, i - 1, -1.0);
    A.set(i, i, 2.0 + 0.01 * (i % 7));
    if (i < n - 1) A.set(i, i + 1, -1.0);
  }
  std::ofstream s(g_out + "/rules.txt");
  s << std::setprecision(17);
  auto line = [&](const char *key, const std::vector<double> &v) {
    s << key;
    for (double a : v) s << " " << a;
    s << "\n";
  };
  {
    solver::bicgstabRe2<ELL_type> sol(A);
    sol.setTolerance(1e-10);
    std::vector<double> b1(n), b2(n), x(n, 0.0);
    fill_noise(b1, seed, 40, 0.0, 1.0);
    fill_noise(b2, seed, 41, 0.0, 1.0);
    auto [it1, e1] = sol(b1, x);
    const std::vector<double> x1 = x;
    auto [it2, e2] = sol(b2, x);
    std::vector<double> b3 = b1;
    b3[7] = std::nan("");
    auto [it3, e3] = sol(b3, x);
    s << "nan " << it1 << " " << e1 << " " << it2 << " " << e2 << " " << it3 << " " << e3 << " "
      << (x == x1 ? 1 : 0) << "\n";
    line("nan_x1", x1);
    line("nan_x3", x);
  }
  {
    solver::bicgstabRe2<ELL_type> sol(A);
    sol.setTolerance(1e-10);
    std::vector<double> b(n), x(n, 0.0);
    for (int k = 0; k < 10; ++k) {
      fill_noise(b, seed, 50 + k, 0.0, 1.0);
      sol(b, x);
    }
    fill_noise(b, seed, 61, 0.0, 1e-12);
    std::fill(x.begin(), x.end(), 0.0);
    auto [itr, er] = sol(b, x);
    const std::vector<double> xr = x;
    solver::bicgstabRe2<ELL_type> fresh(A);
    fresh.setTolerance(1e-10);
    std::fill(x.begin(), x.end(), 0.0);
    auto [itn, en] = fresh(b, x);
    s << "restart " << itr << " " << er << " " << itn << " " << en << " " << sol.numRestarts_ << " "
      << fresh.numRestarts_ << "\n";
    line("restart_x", xr);
    line("norestart_x", x);
  }
  return 0;
}

// usage: ref_harness OUTDIR bench COL VAL N NIT -- the CPU baseline's
// iteration: the reference's own bicgstabRe2 (bicgstabRe2.hpp:91-189) on its
// own ELL_matrix, loaded from a slot-major [7][N] operator (int32 columns,
// FP64 values; the 512x256x64 workload's operator is written by the bench),
// right-hand side splitmix64 noise in [-1, 1), x = 0, run for exactly NIT
// iterations (iters_max_ = NIT, tolerance 1e-300).  Prints one line:
//   ref_bench n N nit I seconds S ms_per_iteration M threads T relres R
// The grid sizes of the reference's grid class are compile-time constants
// (structureGrid.hpp:44-96), so the workload's operator cannot be assembled by
// the reference itself at this size; its solver is size-generic.
static int bench_main(const char *colf, const char *valf, long n, int nit) {
  std::vector<int32_t> col((size_t)7 * n);
  std::vector<double> val((size_t)7 * n);
  {
    std::ifstream fc(colf, std::ios::binary), fv(valf, std::ios::binary);
    fc.read(reinterpret_cast<char *>(col.data()), sizeof(int32_t) * col.size());
    fv.read(reinterpret_cast<char *>(val.data()), sizeof(double) * val.size());
    if (!fc || !fv) {
      std::cerr << "ref_bench: cannot read the operator files\n";
      return 2;
    }
  }
  ELL_type A;
  A.resize(n, 7);
  for (long r = 0; r < n; ++r)
    for (int q = 0; q < 7; ++q) {
      const double v = val[(size_t)q * n + r];
      if (v != 0.0) A.set(r, col[(size_t)q * n + r], v);
    }
  std::vector<double>().swap(val);
  std::vector<int32_t>().swap(col);
  std::vector<double> b(n), x(n, 0.0);
  fill_noise(b, 20240611, 70, 0.0, 1.0);
  solver::bicgstabRe2<ELL_type> sol(A);
  sol.setTolerance(1e-300);
  sol.iters_max_ = nit;
  const auto t0 = std::chrono::steady_clock::now();
  auto [it, err] = sol(b, x);
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cout << std::setprecision(9) << "ref_bench n " << n << " nit " << it << " seconds " << sec
            << " ms_per_iteration " << 1e3 * sec / std::max(it, 1) << " threads " << omp_get_max_threads()
            << " relres " << err << "\n";
  return 0;
}

// usage: ref_harness OUTDIR synth -- every hot-path function once on seeded inputs
static int synth_main() {
  simuClass simu;
  grid gA;
  simu.set_time_max(40);
  simu.set_dt(0.002);
  simu.set_Re(40.0);
  simu.Locality = 1;
  DfibArray Dfib;
  pressure t1;
  velocity T0, T1, T3;
  shareMenory ShareM;
  MxClass Mx;
  calDomain dom;
  std::vector<int> grid_size{gA.nx, gA.ny, gA.nz};
  std::vector<int> dims{1, 1, 1};
  dom.initTable(grid_size, dims);
  dom.initLocal(0, 0, 0);
  resize_variable(gA, t1, T0, T1, T3, Dfib);
  generateGride(simu, ShareM, dom, gA);
  const size_t N = gA.iceltotCal;
  createPressureMatrix(Mx, simu, dom, gA);
  std::fill(Dfib.eta.begin(), Dfib.eta.end(), 0.0);
  DFIB_CylinderZ(Dfib, dom, gA);
  const uint64_t seed = 20240611ULL;
  // predictor: T0 = (1, 0, 0) + 0.3 noise, T1 = 7 (sentinel for unwritten positions)
  fill_noise(T0.u, seed, 1, 1.0, 0.3);
  fill_noise(T0.v, seed, 2, 0.0, 0.3);
  fill_noise(T0.w, seed, 3, 0.0, 0.3);
  T1.u.assign(T1.u.size(), 7.0);
  T1.v.assign(T1.v.size(), 7.0);
  T1.w.assign(T1.w.size(), 7.0);
#ifdef TERBULENCE_SMAGORINSKY
  // the eddy viscosity of the seeded T0 (RUN_CPU_PC_METHOD.hpp:298-300)
  SmagorinskyModel(ShareM, simu, T0, t1, dom, gA);
  dump("syn_les_viseff.bin", T0.Viseff);
#endif
  ConvectionDifussion(simu, T0, T1, dom, gA);
  dump_vel("syn_pred_T1", T1);
  // source add with the driver's indexing (RUN_CPU_PC_METHOD.hpp:309-316)
  std::vector<double> src(3 * N);
  fill_noise(src, seed, 4, 0.0, 1e-3);
  for (int c = 0; c < 3; c++)
    for (auto i = dom.i_begin; i < dom.i_endof; ++i)
      for (auto j = dom.j_begin; j < dom.j_endof; ++j)
        for (auto k = dom.k_begin; k < dom.k_endof; ++k)
          T1.U(c)[gA.icelCal(i, j, k)] += src[gA.icelCal(i, j, k) + gA.iceltotCal * c];
  dump_vel("syn_srcadd_T1", T1);
  // RHS from the T1 just made
  createBMatrix_omp(T1, Mx, simu, dom, gA);
  dump("syn_rhs_b.bin", Mx.matB);
  // SpMV of seeded x, ELL and SPE
  std::vector<double> x(N), y(N);
  fill_noise(x, seed, 5, 0.0, 1.0);
  Mx.matA_ell.multiply_omp(x, y);
  dump("syn_spmv_ell.bin", y);
  Mx.matA_spe.resize(gA.nxCal, gA.nyCal, gA.nzCal);
  Mx.matA_spe.setupPressure(gA.gC, gA.Dx, gA.Dy, gA.Dz, gA.Dxs, gA.Dys, gA.Dzs);
  std::vector<double> spe_jp = Mx.matA_spe.set_Jp_pc();
  Mx.matA_spe.multiply_omp(x, y);
  dump("syn_spmv_spe.bin", y);
  {
    // CSR operator (grid::createPressureMatrixCSR, structureGrid.hpp:630-728),
    // CSR_matrix::multiply_omp (CSR_sparseMatrix.hpp:278-292); it rewrites gA.jp
    const std::vector<double> keep_jp = gA.jp;
    mat::CSR_matrix<double> A(gA.createPressureMatrixCSR());
    A.multiply_omp(x, y);
    dump("syn_spmv_csr.bin", y);
    gA.jp = keep_jp;
  }
  // DFIB update: p seeded, T3 = 9 sentinel
  fill_noise(t1.p, seed, 6, 0.0, 1.0);
  T3.u.assign(T3.u.size(), 9.0);
  T3.v.assign(T3.v.size(), 9.0);
  T3.w.assign(T3.w.size(), 9.0);
  std::fill(Dfib.f.begin(), Dfib.f.end(), 0.0);
  update_UandF_omp(Dfib, simu, t1, T1, T3, dom, gA);
  dump_vel("syn_upd_T3", T3);
  dump("syn_upd_f.bin", Dfib.f);
  {
    virtualF_Int(Dfib, dom, gA, 0);
    virtualF_Int(Dfib, dom, gA, 1);
    std::ofstream s(g_out + "/syn_scalars.txt");
    s << std::setprecision(17) << "cd " << -2.0 * Dfib.ValSum[0] / gA.lz << "\ncl " << -2.0 * Dfib.ValSum[1] / gA.lz
      << "\n";
    auto [ud, vd, wd] = CheckSteadyStatebyMaxVal_omp(simu, ShareM, T0, T3, dom, gA);
    s << "maxdiff " << ud << " " << vd << " " << wd << "\n";
    // BiCGSTAB on the ELL operator from x = 0, reference tolerance 1e-3, on a
    // consistent right-hand side b = A xt (the all-Neumann operator is singular)
    std::vector<double> b(N), xs(N, 0.0), xt(N);
    fill_noise(xt, seed, 7, 0.0, 1.0);
    Mx.matA_ell.multiply_omp(xt, b);
    solver::bicgstabRe2<ELL_type> sol(Mx.matA_ell);
    sol.setTolerance(1e-3);
    auto [it, err] = sol(b, xs);
    s << "bicg_ell " << it << " " << err << "\n";
    dump("syn_bicg_x.bin", xs);
  }
  // BC main on seeded T0 and p, then copy into seeded T1
  fill_noise(T0.u, seed, 8, 1.0, 0.3);
  fill_noise(T0.v, seed, 9, 0.0, 0.3);
  fill_noise(T0.w, seed, 10, 0.0, 0.3);
  fill_noise(T1.u, seed, 11, 1.0, 0.3);
  fill_noise(T1.v, seed, 12, 0.0, 0.3);
  fill_noise(T1.w, seed, 13, 0.0, 0.3);
  BC_staggered_main(dom, T0, t1, gA);
  BC_staggered_copy(dom, T0, T1, gA);
  dump_vel("syn_bc_T0", T0);
  dump_vel("syn_bc_T1", T1);
  dump("syn_bc_p.bin", t1.p);
  // SorPipeLine_seq: 12 lexicographic sweeps from p = 0 on the seeded T1
  // (simu.loop == 1 so that the coefficients are computed)
  {
    SORcoefficient Sor;
    fill_noise(T1.u, seed, 14, 1.0, 0.05);
    fill_noise(T1.v, seed, 15, 0.0, 0.05);
    fill_noise(T1.w, seed, 16, 0.0, 0.05);
    t1.init_p(0.0);
    simu.loop = 1;
    simu.p_criteria = 0.0;
    simu.p_sor_iter_max = 12;
    SorPipeLine_seq(Sor, simu, T1, t1, dom, gA);
    dump("syn_sor_p.bin", t1.p);
    std::ofstream s(g_out + "/syn_scalars.txt", std::ios::app);
    s << std::setprecision(17) << "sor_seq " << simu.iters << " " << simu.error << "\n";
  }
  return 0;
}

// usage: ref_harness OUTDIR io -- the output writers and the Plot3D reader on
// seeded fields of the native grid, run inside OUTDIR (the writers use the
// relative paths mx_out/ and Information/):
//   mx_out/P3D.x, mx_out/P3D01.q   OutputPlot3D_Xfile, OutputPLOT3D_Qfile
//                                  (p tag 20, T3 tags 21-23, eta of the cylinder,
//                                  Re 40, time 0.25, currentFile 1)
//   read_{p,u,v,w,eta}.bin, read_header.txt   QfileRead of that .q file
//   cdcl_delay.dat      Delay_IO_CD_CL, delayIO 3, loops 1..7
//   cdcl_nodelay.dat    noDelay_IO_CD_CL, loops 1..4
//   (f of loop L: tag 30 + L, amplitude 1; time advanced by dt = 0.002 per loop)
static int io_main() {
  simuClass simu;
  grid gA;
  simu.set_time_max(40);
  simu.set_dt(0.002);
  simu.set_Re(40.0);
  simu.Locality = 1;
  simu.DfibMethod = "DFIB_Cylinder-Z";
  DfibArray Dfib;
  pressure t1;
  velocity T0, T1, T3;
  shareMenory ShareM;
  calDomain dom;
  std::vector<int> grid_size{gA.nx, gA.ny, gA.nz};
  std::vector<int> dims{1, 1, 1};
  dom.initTable(grid_size, dims);
  dom.initLocal(0, 0, 0);
  resize_variable(gA, t1, T0, T1, T3, Dfib);
  generateGride(simu, ShareM, dom, gA);
  std::fill(Dfib.eta.begin(), Dfib.eta.end(), 0.0);
  DFIB_CylinderZ(Dfib, dom, gA);
  const uint64_t seed = 20240611ULL;
  fill_noise(t1.p, seed, 20, 0.0, 1.0);
  fill_noise(T3.u, seed, 21, 1.0, 0.3);
  fill_noise(T3.v, seed, 22, 0.0, 0.3);
  fill_noise(T3.w, seed, 23, 0.0, 0.3);
  if (chdir(g_out.c_str()) != 0) return 1;
  mkdir("mx_out", 0755);
  mkdir("Information", 0755);
  OutputPlot3D_Xfile(simu, gA);
  simu.currentFile = 1;
  simu.set_time(0.25);
  OutputPLOT3D_Qfile(Dfib, simu, t1, T3, gA);
  {
    auto [nb, mach, alpha, reyn, time, p, u, v, w, eta] = QfileRead(gA, "mx_out/P3D01.q");
    g_out = ".";
    dump("read_p.bin", p);
    dump("read_u.bin", u);
    dump("read_v.bin", v);
    dump("read_w.bin", w);
    dump("read_eta.bin", eta);
    std::ofstream h("read_header.txt");
    h << std::setprecision(9) << nb << " " << mach << " " << alpha << " " << reyn << " " << time << "\n";
  }
  simu.delayIO = 3;
  simu.set_time(0.0);
  for (simu.loop = 1; simu.loop <= 7; simu.finishloop()) {
    fill_noise(Dfib.f, seed, 30 + (int)simu.loop, 0.0, 1.0);
    Delay_IO_CD_CL(simu, dom, Dfib, gA);
  }
  if (rename("Information/Time_cDcL.dat", "cdcl_delay.dat") != 0) return 1;
  simu.set_time(0.0);
  for (simu.loop = 1; simu.loop <= 4; simu.finishloop()) {
    fill_noise(Dfib.f, seed, 30 + (int)simu.loop, 0.0, 1.0);
    noDelay_IO_CD_CL(simu, dom, Dfib, gA);
  }
  if (rename("Information/Time_cDcL.dat", "cdcl_nodelay.dat") != 0) return 1;
  return 0;
}

int main(int argc, char **argv) {
  if (argc >= 3 && std::string(argv[2]) == "synth") {
    g_out = argv[1];
    mkdir(g_out.c_str(), 0755);
    return synth_main();
  }
  if (argc >= 3 && std::string(argv[2]) == "io") {
    g_out = argv[1];
    mkdir(g_out.c_str(), 0755);
    return io_main();
  }
  if (argc >= 7 && std::string(argv[2]) == "bench")
    return bench_main(argv[3], argv[4], std::atol(argv[5]), std::atoi(argv[6]));
  if (argc >= 3 && std::string(argv[2]) == "rules") {
    g_out = argv[1];
    mkdir(g_out.c_str(), 0755);
    return rules_main();
  }
  if (argc < 6) {
    std::cerr << "usage: ref_harness OUTDIR NSTEPS TOL OP(ell|spe) FN_STEP [DUMP_EVERY [SRC_AT_CELL]]\n";
    return 2;
  }
  // state dumps every DUMP_EVERY steps (default 1; 0: statistics only)
  const int dump_every = argc >= 7 ? std::atoi(argv[6]) : 1;
  // source term index: 0 (default) adds source_termF[icelCal] at
  // T1[icelCal] as the reference's driver does (RUN_CPU_PC_METHOD.hpp:309-316);
  // 1 adds it at the cell it belongs to, T1[icel]
  const int src_at_cell = argc >= 8 ? std::atoi(argv[7]) : 0;
  g_out = argv[1];
  const int nsteps = std::atoi(argv[2]);
  const double tol = std::atof(argv[3]);
  const bool use_spe = std::string(argv[4]) == "spe";
  const int fn_step = std::atoi(argv[5]);
  mkdir(g_out.c_str(), 0755);

  // src/CFD_MX.cpp:28-57
  simuClass simu;
  grid gA;
  simu.set_time_max(40);
  simu.set_dt(0.002);
  simu.set_Re(40.0);
  simu.init_fileStep(1);
  simu.DfibMethod = "DFIB_Cylinder-Z";
  simu.Dfib_boolT = true;
  simu.Locality = 1;
  simu.p_criteria = tol;

  // RUN_CPU_PC_METHOD.hpp:30-119
  DfibArray Dfib;
  pressure t1;
  velocity T0, T1, T3;
  shareMenory ShareM;
  MxClass Mx;
  calDomain dom;
  std::vector<int> grid_size{gA.nx, gA.ny, gA.nz};
  std::vector<int> dims{1, 1, 1};
  dom.initTable(grid_size, dims);
  dom.initLocal(0, 0, 0);
  resize_variable(gA, t1, T0, T1, T3, Dfib);
  T0.iniU_omp(1.0, 0.0, 0.0);
  T1.iniU_omp(1.0, 0.0, 0.0);
  T3.iniU_omp(1.0, 0.0, 0.0);
  t1.init_p(0.0);
  generateGride(simu, ShareM, dom, gA);
  T1.u = T3.u = T0.u;
  T1.v = T3.v = T0.v;
  T1.w = T3.w = T0.w;
  BC_staggered_main(dom, T0, t1, gA);
  BC_staggered_copy(dom, T0, T1, gA);

  {
    std::vector<double> g;
    for (auto *a : {&gA.X, &gA.Y, &gA.Z, &gA.Dx, &gA.Dy, &gA.Dz, &gA.Dxs, &gA.Dys, &gA.Dzs})
      g.insert(g.end(), a->begin(), a->end());
    dump("grid.bin", g);
    dump("init_T0_u.bin", T0.u);
  }

  const size_t N = gA.iceltotCal;
  Mx.X_result.resize(N, 0.0);

  // ELL operator (RUN_CPU_PC_METHOD.hpp:137-148)
  createPressureMatrix(Mx, simu, dom, gA);
  {
    std::vector<int> col(7 * N);
    std::vector<double> val(7 * N);
    for (size_t r = 0; r < N; ++r)
      for (int s = 0; s < 7; ++s) {
        auto &pr = Mx.matA_ell.Idx_val_[r * 7 + s];
        col[s * N + r] = pr.first;
        val[s * N + r] = pr.second;
      }
    dump("ell_col.bin", col);
    dump("ell_val.bin", val);
    dump("jp.bin", gA.jp);
  }
  std::vector<double> ell_jp = gA.jp;
  {
    // CSR operator (RUN_CPU_PC_METHOD.hpp:156, structureGrid.hpp:630-728)
    auto [ptr, idx, val] = gA.createPressureMatrixCSR();
    dump("csr_ptr.bin", ptr);
    dump("csr_idx.bin", idx);
    dump("csr_val.bin", val);
    dump("csr_jp.bin", gA.jp);
    gA.jp = ell_jp;
  }
  // SPE operator (RUN_CPU_PC_METHOD.hpp:167-175)
  Mx.matA_spe.resize(gA.nxCal, gA.nyCal, gA.nzCal);
  Mx.matA_spe.setupPressure(gA.gC, gA.Dx, gA.Dy, gA.Dz, gA.Dxs, gA.Dys, gA.Dzs);
  std::vector<double> spe_jp = Mx.matA_spe.set_Jp_pc();
  dump("spe_jp.bin", spe_jp);
  gA.jp = use_spe ? spe_jp : ell_jp;

  solver::bicgstabRe2<ELL_type> ellSolver(Mx.matA_ell);
  solver::bicgstabRe2<SPE_type> speSolver(Mx.matA_spe);
  ellSolver.setTolerance(simu.p_criteria);
  speSolver.setTolerance(simu.p_criteria);
  auto solve = [&](std::vector<double> &b, std::vector<double> &x) {
    return use_spe ? speSolver(b, x) : ellSolver(b, x);
  };

  // RUN_CPU_PC_METHOD.hpp:258-264
  std::fill(Dfib.eta.begin(), Dfib.eta.end(), 0.0);
  DFIB_CylinderZ(Dfib, dom, gA);
  dump("eta.bin", Dfib.eta);
  auto source_termF = Dfib.f;

  std::ofstream stats(g_out + "/stats.txt");
  stats << std::setprecision(17);

  auto drag = [&](int dir) {
    virtualF_Int(Dfib, dom, gA, dir);
    return -2.0 * Dfib.ValSum[dir] / gA.lz;
  };

  for (simu.loop = 1; (int)simu.loop <= nsteps; simu.finishloop()) {
    const bool fn = (int)simu.loop == fn_step;
    char tag[64];
    std::fill(source_termF.begin(), source_termF.end(), double());
    for (int pass = 0; pass < 4; ++pass) {
      const bool fnp = fn && pass == 1;
      if (fnp) { dump_vel("fn_pred_in_T0", T0); dump_vel("fn_pred_in_T1", T1); }
#ifdef TERBULENCE_SMAGORINSKY
      SmagorinskyModel(ShareM, simu, T0, t1, dom, gA);  // RUN_CPU_PC_METHOD.hpp:298-300, 449-451
#endif
      ConvectionDifussion(simu, T0, T1, dom, gA);
      if (fnp) { dump_vel("fn_pred_out_T1", T1); dump("fn_src.bin", source_termF); }
      for (int c = 0; c < 3; c++)
        for (auto i = dom.i_begin; i < dom.i_endof; ++i)
          for (auto j = dom.j_begin; j < dom.j_endof; ++j)
            for (auto k = dom.k_begin; k < dom.k_endof; ++k)
              T1.U(c)[src_at_cell ? gA.icel(i, j, k) : gA.icelCal(i, j, k)] +=
                  source_termF[gA.icelCal(i, j, k) + gA.iceltotCal * c];
      if (fnp) dump_vel("fn_srcadd_out_T1", T1);
      int iters = -1;
      double err = 0.0;
      if (simu.loop > 1) {
        createBMatrix_omp(T1, Mx, simu, dom, gA);
        if (fnp) { dump("fn_rhs_out_b.bin", Mx.matB); dump("fn_solve_in_x.bin", Mx.X_result); }
        std::tie(iters, err) = solve(Mx.matB, Mx.X_result);
        if (fnp) {
          dump("fn_solve_out_x.bin", Mx.X_result);
          std::vector<double> ax(N), ax2(N);
          Mx.matA_ell.multiply_omp(Mx.X_result, ax);
          Mx.matA_spe.multiply_omp(Mx.X_result, ax2);
          dump("fn_spmv_ell.bin", ax);
          dump("fn_spmv_spe.bin", ax2);
          stats << "fn_solve " << iters << " " << err << "\n";
        }
        Pressure_transform_X_result(t1, Mx, dom, gA);
      }
      if (fnp) dump("fn_update_in_p.bin", t1.p);
      update_UandF_omp(Dfib, simu, t1, T1, T3, dom, gA);
      if (fnp) { dump_vel("fn_update_out_T3", T3); dump("fn_update_out_f.bin", Dfib.f); }
      if (pass < 3)
        for (int c = 0; c < 3; c++)
          for (auto i = dom.i_begin; i < dom.i_endof; ++i)
            for (auto j = dom.j_begin; j < dom.j_endof; ++j)
              for (auto k = dom.k_begin; k < dom.k_endof; ++k)
                source_termF[gA.icelCal(i, j, k) + gA.iceltotCal * c] +=
                    simu.dt * Dfib.f[gA.icelCal(i, j, k) + gA.iceltotCal * c];
      stats << "pass " << simu.loop << " " << pass << " " << iters << " " << err << " " << drag(0) << " "
            << drag(1) << "\n";
    }
    auto [ud, vd, wd] = CheckSteadyStatebyMaxVal_omp(simu, ShareM, T0, T3, dom, gA);
    stats << "maxdiff " << simu.loop << " " << ud << " " << vd << " " << wd << "\n";
    T0.copy_omp(T3);
    if (fn) { dump_vel("fn_bc_in_T0", T0); dump("fn_bc_in_p.bin", t1.p); dump_vel("fn_bccopy_in_T1", T1); }
    BC_staggered_main(dom, T0, t1, gA);
    BC_staggered_copy(dom, T0, T1, gA);
    if (fn) { dump_vel("fn_bc_out_T0", T0); dump("fn_bc_out_p.bin", t1.p); dump_vel("fn_bccopy_out_T1", T1); }
    if (dump_every > 0 && (int)simu.loop % dump_every == 0) {
      std::snprintf(tag, sizeof(tag), "step_%03d", (int)simu.loop);
      dump_vel(std::string(tag) + "_T0", T0);
      dump_vel(std::string(tag) + "_T1", T1);
      dump(std::string(tag) + "_p.bin", t1.p);
      dump(std::string(tag) + "_X.bin", Mx.X_result);
      dump(std::string(tag) + "_f.bin", Dfib.f);
    }
    stats.flush();
  }
  return 0;
}
