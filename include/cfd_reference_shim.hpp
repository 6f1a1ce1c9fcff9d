// cfd_reference_shim.hpp -- header-only C++17 mirror of the reference
// solver's hot-path interfaces on top of the C-ABI in cfd_capi.h.
//
// The types and functions keep the reference's names and argument lists so
// that a driver written against the reference (inc/run/RUN_CPU_PC_METHOD.hpp)
// ports by changing its namespace: simuClass, grid, velocity, pressure,
// DfibArray, MxClass, calDomain, ConvectionDifussion, createBMatrix_omp,
// createPressureMatrix, update_UandF_omp, Pressure_transform_X_result,
// BC_staggered_main, BC_staggered_copy, DFIB_CylinderZ, virtualF_Int,
// CheckSteadyStatebyMaxVal_omp, Delay_IO_CD_CL, noDelay_IO_CD_CL,
// OutputPlot3D_Xfile, OutputPLOT3D_Qfile, QfileRead,
// grid::createPressureMatrixCSR, mat::ELL_matrix, mat::CSR_matrix,
// solver::bicgstabRe2.  Fields stay in
// caller-owned std::vector<double> in the reference layout; each call
// uploads its inputs, runs on the GPU and downloads its outputs, so it is a
// migration path, not the fast path.  The fast path keeps the state on the
// device: grid::step() runs one whole timestep (cfd_step) without host traffic.
//
// Errors: a status CFD_ERR_ARG throws std::invalid_argument (where the
// reference throws it), every other failure std::runtime_error.
#pragma once

#include <cmath>
#include <cstring>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "cfd_capi.h"

namespace cfdgpu {

inline void check(int status, const char *where) {
  if (status == CFD_OK) return;
  const std::string msg = std::string(where) + ": " + cfd_last_error();
  if (status == CFD_ERR_ARG) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

// inc/simuClass.hpp:7-296 (the members the hot path and its output read or write)
struct simuClass {
  double Re = 0.0, nu = 0.0, dt = 0.0, time = 0.0, time_max = 0.0;
  size_t loop = 1, loop_max = 0;
  int delayIO = 100;        // Delay_IO_CD_CL ring-buffer length
  size_t currentFile = 0;   // Plot3D solution file number
  std::string zone = "ELL (BICG)";  // ZONE(): the pressure solver's name
  std::shared_ptr<cfd_cdcl_log> cdcl;  // the cd / cl ring buffers and their file
  int iters = 0;
  double error = 0.0;
  double p_criteria = 1.e-3;
  double p_sor_omega = 1.8;
  size_t p_iterMax = 5000;
  std::string DfibMethod = "OFF";
  bool set_Re(double re) { Re = re; nu = 1.0 / re; return true; }
  void set_time_max(double t) { time_max = t; }
  bool set_dt(double t) { dt = t; loop_max = (size_t)(time_max / dt); return true; }
  bool get_finishloop() const { return loop < loop_max || time < time_max; }
  bool finishloop() { time += dt; loop++; return true; }
  std::string ZONE() const { return zone; }
  // the Time_cDcL.dat log of this run, opened on first use
  void cdcl_record(const char *path, int mode, std::pair<double, double> cdl) {
    if (!cdcl) {
      cfd_cdcl_log *h = nullptr;
      check(cfd_cdcl_open(path, delayIO, zone.c_str(), mode, &h), "cfd_cdcl_open");
      cdcl.reset(h, [](cfd_cdcl_log *q) { cfd_cdcl_close(q); });
    }
    check(cfd_cdcl_record(cdcl.get(), (int)loop, time, cdl.first, cdl.second), "cfd_cdcl_record");
  }
};

// simuClass::ZONE (simuClass.hpp:203-236) for a parameter set
inline std::string ZONE(const cfd_params &p) {
  if (p.solver == CFD_SOLVER_RBSOR) return "SOR";
  if (p.op == CFD_OP_CSR) return "CSR (BICG)";
  if (p.op == CFD_OP_SPE) return "SPE (BICG)";
  return "ELL (BICG)";
}

// inc/physicalVariables.hpp:7-169
struct velocity {
  std::vector<double> u, v, w;
  std::vector<double> &U(int d) {
    if (d < 0 || d > 2) throw std::runtime_error("out of dimensions!");
    return d == 0 ? u : (d == 1 ? v : w);
  }
};
struct pressure { std::vector<double> p; };
struct DfibArray { std::vector<double> f, eta; };
struct MxClass { std::vector<double> matB, X_result; };  // inc/pressure/PressureClass.hpp:24-45

// inc/domain.hpp:5-42 (one sub-domain covering the interior)
struct calDomain {
  int i_begin = 2, j_begin = 2, k_begin = 2, i_endof = 0, j_endof = 0, k_endof = 0;
};

// The grid owns the device context (one HIP stream and the device copies of
// the fields); its geometry follows inc/grid/structureGrid.hpp.
class grid {
 public:
  int nxCal, nyCal, nzCal, gC = 2, nx, ny, nz;
  size_t iceltot, iceltotCal;
  double lx, ly, lz;
  std::vector<double> X, Y, Z, Dx, Dy, Dz, Dxs, Dys, Dzs;
  cfd_params prm{};

  // uniform grid (grid::init_uniform) over lx x ly x lz
  grid(int nxc, int nyc, int nzc, double Lx, double Ly, double Lz, const cfd_params &p, int device = 0)
      : nxCal(nxc), nyCal(nyc), nzCal(nzc), lx(Lx), ly(Ly), lz(Lz), prm(p) {
    init_sizes();
    check(cfd_axis_uniform(nxCal, lx, X.data(), Dx.data(), Dxs.data()), "cfd_axis_uniform");
    check(cfd_axis_uniform(nyCal, ly, Y.data(), Dy.data(), Dys.data()), "cfd_axis_uniform");
    check(cfd_axis_uniform(nzCal, lz, Z.data(), Dz.data(), Dzs.data()), "cfd_axis_uniform");
    create(device);
  }
  // the reference's compiled-in 'non_uniform-SML' grid (structureGrid.hpp:388-405)
  static grid native(const cfd_params &p, int device = 0) { return grid(p, device); }

  grid(const grid &) = delete;
  grid &operator=(const grid &) = delete;
  grid(grid &&o) noexcept { *this = std::move(o); }
  grid &operator=(grid &&o) noexcept {
    if (this != &o) {
      release();
      nxCal = o.nxCal; nyCal = o.nyCal; nzCal = o.nzCal; gC = o.gC; nx = o.nx; ny = o.ny; nz = o.nz;
      iceltot = o.iceltot; iceltotCal = o.iceltotCal; lx = o.lx; ly = o.ly; lz = o.lz;
      X = std::move(o.X); Y = std::move(o.Y); Z = std::move(o.Z); Dx = std::move(o.Dx); Dy = std::move(o.Dy);
      Dz = std::move(o.Dz); Dxs = std::move(o.Dxs); Dys = std::move(o.Dys); Dzs = std::move(o.Dzs);
      prm = o.prm; ctx_ = o.ctx_; o.ctx_ = nullptr;
    }
    return *this;
  }
  ~grid() { release(); }

  size_t icel(int i, int j, int k) const { return ((size_t)i * ny + j) * nz + k; }
  size_t icelCal(int i, int j, int k) const {
    return ((size_t)(i - gC) * nyCal + (j - gC)) * nzCal + (k - gC);
  }
  cfd_ctx *ctx() const { return ctx_; }
  // the C-ABI view of the metrics (host pointers into this grid)
  cfd_grid_desc desc() const {
    return cfd_grid_desc{nxCal,     nyCal,     nzCal,     X.data(), Y.data(), Z.data(), Dx.data(), Dy.data(),
                         Dz.data(), Dxs.data(), Dys.data(), Dzs.data(), lx,     ly,       lz};
  }
  calDomain domain() const {
    calDomain d;
    d.i_endof = nx - gC; d.j_endof = ny - gC; d.k_endof = nz - gC;
    return d;
  }

  // device <-> host of one array by field id
  void put(int field, const std::vector<double> &h) const {
    check(cfd_upload(ctx_, field, h.data(), h.size()), "cfd_upload");
  }
  void get(int field, std::vector<double> &h) const {
    size_t G = 0, N = 0;
    check(cfd_sizes(ctx_, &G, &N), "cfd_sizes");
    h.resize(field <= CFD_ETA ? G : ((field == CFD_F || field == CFD_SRC) ? 3 * N : N));
    check(cfd_download(ctx_, field, h.data(), h.size()), "cfd_download");
  }
  void put(int f0, const velocity &V) const { put(f0, V.u); put(f0 + 1, V.v); put(f0 + 2, V.w); }
  void get(int f0, velocity &V) const { get(f0, V.u); get(f0 + 1, V.v); get(f0 + 2, V.w); }

  // structureGrid.hpp:630-728: assembles the CSR operator on the device and
  // returns host copies of its row pointers, columns and values (the tuple
  // the reference returns); jp follows it when prm.op == CFD_OP_CSR
  std::tuple<std::vector<int>, std::vector<int>, std::vector<double>> createPressureMatrixCSR() {
    check(cfd_csr_assemble(ctx_), "cfd_csr_assemble");
    size_t nnz = 0;
    check(cfd_csr_nnz(ctx_, &nnz), "cfd_csr_nnz");
    std::vector<int> ptr(iceltotCal + 1), idx(nnz);
    std::vector<double> val(nnz);
    check(cfd_csr_download(ctx_, ptr.data(), idx.data(), val.data(), iceltotCal, nnz), "cfd_csr_download");
    return {std::move(ptr), std::move(idx), std::move(val)};
  }

  // device-resident timestep (runSeqOmp_predictionMethod's loop body)
  cfd_step_stats step() {
    cfd_step_stats st;
    check(cfd_step(ctx_, &st), "cfd_step");
    return st;
  }

 private:
  cfd_ctx *ctx_ = nullptr;

  explicit grid(const cfd_params &p, int device) : nxCal(190), nyCal(210), nzCal(40), lx(17.0), ly(12.0), lz(3.0), prm(p) {
    init_sizes();
    check(cfd_axis_sml(nxCal, lx, 0.3 * lx, 100, -4, X.data(), Dx.data(), Dxs.data()), "cfd_axis_sml");
    check(cfd_axis_sml(nyCal, ly, 0.3 * ly, 100, 0, Y.data(), Dy.data(), Dys.data()), "cfd_axis_sml");
    check(cfd_axis_sml(nzCal, lz, 0.5 * lz, 0, 0, Z.data(), Dz.data(), Dzs.data()), "cfd_axis_sml");
    create(device);
  }
  void init_sizes() {
    nx = nxCal + 2 * gC; ny = nyCal + 2 * gC; nz = nzCal + 2 * gC;
    iceltot = (size_t)nx * ny * nz;
    iceltotCal = (size_t)nxCal * nyCal * nzCal;
    X.assign(nx + 1, 0.0); Y.assign(ny + 1, 0.0); Z.assign(nz + 1, 0.0);
    Dx.assign(nx, 0.0); Dy.assign(ny, 0.0); Dz.assign(nz, 0.0);
    Dxs.assign(nx, 0.0); Dys.assign(ny, 0.0); Dzs.assign(nz, 0.0);
  }
  void create(int device) {
    const cfd_grid_desc d = desc();
    check(cfd_create(&d, &prm, device, &ctx_), "cfd_create");
  }
  void release() {
    if (ctx_) cfd_destroy(ctx_);
    ctx_ = nullptr;
  }
};

// inc/resize_and_simuSetting.hpp:5-31
inline void resize_variable(grid &g, pressure &t1, velocity &T0, velocity &T1, velocity &T3, DfibArray &Dfib) {
  Dfib.eta.assign(g.iceltot, 0.0);
  Dfib.f.assign(3 * g.iceltotCal, 0.0);
  t1.p.assign(g.iceltot, 0.0);
  for (velocity *V : {&T0, &T1, &T3}) {
    V->u.assign(g.iceltot, 0.0); V->v.assign(g.iceltot, 0.0); V->w.assign(g.iceltot, 0.0);
  }
}

// inc/source/convection_and_difussion.hpp:168-446
inline bool ConvectionDifussion(simuClass &, velocity &oldVel, velocity &curtVel, const calDomain &, grid &g) {
  g.put(CFD_T0_U, oldVel);
  g.put(CFD_T1_U, curtVel);
  check(cfd_predictor(g.ctx()), "cfd_predictor");
  g.get(CFD_T1_U, curtVel);
  return true;
}

// inc/matrix/b_matrix.hpp:51-102 (JACOBI_PC): needs gA.jp from the assembly
inline void createBMatrix_omp(velocity &T1, MxClass &Mx, simuClass &, calDomain &, grid &g) {
  g.put(CFD_T1_U, T1);
  check(cfd_build_rhs(g.ctx()), "cfd_build_rhs");
  g.get(CFD_RHS, Mx.matB);
}

// inc/matrix/PressureMatrix.hpp:5-117
inline bool createPressureMatrix(MxClass &, simuClass &, calDomain &, grid &g) {
  check(cfd_ell_assemble(g.ctx()), "cfd_ell_assemble");
  return true;
}

// inc/dfib/setEta.hpp:136-204
inline void DFIB_CylinderZ(DfibArray &Dfib, calDomain &, grid &g) {
  check(cfd_set_eta_cylinder_z(g.ctx()), "cfd_set_eta_cylinder_z");
  g.get(CFD_ETA, Dfib.eta);
}

// inc/dfib/updateUandF.hpp:95-192
inline void update_UandF_omp(DfibArray &Dfib, const simuClass &, const pressure &t1, const velocity &oldVel,
                             velocity &curtVel, const calDomain &, grid &g) {
  g.put(CFD_ETA, Dfib.eta);
  g.put(CFD_P, t1.p);
  g.put(CFD_T1_U, oldVel);
  g.put(CFD_T3_U, curtVel);
  g.put(CFD_F, Dfib.f);
  check(cfd_dfib_update(g.ctx()), "cfd_dfib_update");
  g.get(CFD_T3_U, curtVel);
  g.get(CFD_F, Dfib.f);
}

// inc/pressure/Transform.hpp:8-24
inline void Pressure_transform_X_result(pressure &t1, MxClass &Mx, calDomain &, grid &g) {
  g.put(CFD_P, t1.p);
  g.put(CFD_X, Mx.X_result);
  check(cfd_pressure_scatter(g.ctx()), "cfd_pressure_scatter");
  g.get(CFD_P, t1.p);
}

// inc/BoundaryCondition/Boundary_Condition.hpp:37-175
inline bool BC_staggered_main(calDomain &, velocity &vel, pressure &pre, grid &g) {
  g.put(CFD_T0_U, vel);
  g.put(CFD_P, pre.p);
  check(cfd_bc_main(g.ctx()), "cfd_bc_main");
  g.get(CFD_T0_U, vel);
  g.get(CFD_P, pre.p);
  return true;
}

// inc/BoundaryCondition/updateBC.hpp:6-109
inline bool BC_staggered_copy(const calDomain &, const velocity &oldVel, velocity &currtVel, grid &g) {
  g.put(CFD_T0_U, oldVel);
  g.put(CFD_T1_U, currtVel);
  check(cfd_bc_copy(g.ctx()), "cfd_bc_copy");
  g.get(CFD_T1_U, currtVel);
  return true;
}

// inc/dfib/virtualForceIntergrator.hpp:5-37 (both directions at once): cD, cL
inline std::pair<double, double> virtualF_Int(DfibArray &Dfib, calDomain &, grid &g) {
  g.put(CFD_F, Dfib.f);
  double cd = 0.0, cl = 0.0;
  check(cfd_drag(g.ctx(), &cd, &cl), "cfd_drag");
  return {cd, cl};
}

// inc/analyses/CheckSteadystate.hpp:165-219
inline std::tuple<double, double, double> CheckSteadyStatebyMaxVal_omp(simuClass &, velocity &T0, velocity &T3,
                                                                       calDomain &, grid &g) {
  g.put(CFD_T0_U, T0);
  g.put(CFD_T3_U, T3);
  double m[3];
  check(cfd_maxdiff(g.ctx(), m), "cfd_maxdiff");
  return {m[0], m[1], m[2]};
}

// inc/dfib/virtualForceIntergrator.hpp:105-177: cD, cL of f into simuClass's
// delayIO ring buffer, flushed to Information/Time_cDcL.dat every delayIO steps
// (as written: each flush repeats the current line delayIO times)
inline std::pair<double, double> Delay_IO_CD_CL(simuClass &simu, calDomain &Lo, DfibArray &Dfib, grid &g) {
  auto cdl = virtualF_Int(Dfib, Lo, g);
  simu.cdcl_record("Information/Time_cDcL.dat", CFD_CDCL_REFERENCE, cdl);
  return cdl;
}

// inc/dfib/virtualForceIntergrator.hpp:44-99: one line per step
inline std::pair<double, double> noDelay_IO_CD_CL(simuClass &simu, calDomain &Lo, DfibArray &Dfib, grid &g) {
  auto cdl = virtualF_Int(Dfib, Lo, g);
  std::cout << "[cD, cL] : " << cdl.first << ", " << cdl.second << std::endl;
  simu.cdcl_record("Information/Time_cDcL.dat", CFD_CDCL_EVERY_STEP, cdl);
  return cdl;
}

// inc/io/Plot3D/xfileWrite.hpp:6-61
inline int OutputPlot3D_Xfile(simuClass &, grid &g) {
  const cfd_grid_desc d = g.desc();
  check(cfd_plot3d_write_x(&d, "mx_out/P3D.x"), "cfd_plot3d_write_x");
  return 0;
}

// inc/io/Plot3D/qfile.hpp:7-104: mx_out/P3D<currentFile, two digits at least>.q
inline int OutputPLOT3D_Qfile(DfibArray &Dfib, simuClass &simu, pressure &pres, velocity &currt, grid &g) {
  const size_t io = simu.currentFile;
  const std::string name = std::string("mx_out/P3D") + (io < 10 ? "0" : "") + std::to_string(io) + ".q";
  std::cout << "(Output Plaot3D):" << io << std::endl;
  g.put(CFD_P, pres.p);
  g.put(CFD_T3_U, currt);
  g.put(CFD_ETA, Dfib.eta);
  check(cfd_plot3d_write_q(g.ctx(), name.c_str(), CFD_T3_U, simu.time), "cfd_plot3d_write_q");
  return 0;
}

// inc/io/Plot3D/qfileRead.hpp:6-100: (Nblock, mach, alpha, reyn, time, p, u, v, w, eta)
inline std::tuple<int, float, float, float, float, std::vector<double>, std::vector<double>, std::vector<double>,
                  std::vector<double>, std::vector<double>>
QfileRead(grid &g, const std::string &filename) {
  std::vector<double> p(g.iceltot), u(g.iceltot), v(g.iceltot), w(g.iceltot), eta(g.iceltot);
  int nb = 0;
  float h[4] = {0, 0, 0, 0};
  const cfd_grid_desc d = g.desc();
  check(cfd_plot3d_read_q(filename.c_str(), &d, &nb, h, p.data(), u.data(), v.data(), w.data(), eta.data()),
        "cfd_plot3d_read_q");
  return {nb, h[0], h[1], h[2], h[3], std::move(p), std::move(u), std::move(v), std::move(w), std::move(eta)};
}

namespace mat {
// inc/matrix/ELL_sparseMatrix.hpp:35-139: the assembled operator of a grid
class ELL_matrix {
 public:
  static constexpr int op = CFD_OP_ELL;
  explicit ELL_matrix(grid &g) : g_(&g) {}
  int row() const { return (int)g_->iceltotCal; }
  // r = A x (multiply_omp, :477-492)
  bool multiply_omp(std::vector<double> &x, std::vector<double> &r) {
    g_->put(CFD_VEC_A, x);
    check(cfd_spmv(g_->ctx(), CFD_OP_ELL, CFD_FP64, CFD_VEC_A, CFD_VEC_B), "cfd_spmv");
    g_->get(CFD_VEC_B, r);
    return true;
  }
  grid &owner() { return *g_; }

 private:
  grid *g_;
};

// inc/matrix/CSR_sparseMatrix.hpp:42-292: the CSR operator of a grid, made on
// the device by grid::createPressureMatrixCSR
class CSR_matrix {
 public:
  static constexpr int op = CFD_OP_CSR;
  explicit CSR_matrix(grid &g) : g_(&g) {}
  // set (:42-120): takes the tuple grid::createPressureMatrixCSR returned;
  // the device already holds that operator, so this checks the shapes
  void set(const std::tuple<std::vector<int>, std::vector<int>, std::vector<double>> &m) {
    size_t nnz = 0;
    check(cfd_csr_nnz(g_->ctx(), &nnz), "cfd_csr_nnz");
    if (std::get<0>(m).size() != g_->iceltotCal + 1 || std::get<1>(m).size() != nnz ||
        std::get<2>(m).size() != nnz)
      throw std::invalid_argument("CSR_matrix::set: not the operator of this grid");
  }
  int row() const { return (int)g_->iceltotCal; }
  // r = A x (multiply_omp, :278-292)
  bool multiply_omp(std::vector<double> &x, std::vector<double> &r) {
    g_->put(CFD_VEC_A, x);
    check(cfd_spmv(g_->ctx(), CFD_OP_CSR, CFD_FP64, CFD_VEC_A, CFD_VEC_B), "cfd_spmv");
    g_->get(CFD_VEC_B, r);
    return true;
  }
  grid &owner() { return *g_; }

 private:
  grid *g_;
};
}  // namespace mat

namespace solver {
// inc/solver/bicgstabRe2.hpp:13-189; the solver state (px1, px2, solve count)
// lives with the grid's device context
template <typename matrixT>
class bicgstabRe2 {
 public:
  explicit bicgstabRe2(matrixT &A) : A_(&A) {}
  void setTolerance(double tol) { tol_ = tol; }
  std::pair<int, double> operator()(const std::vector<double> &rhs, std::vector<double> &x) {
    grid &g = A_->owner();
    // the device solve runs on the operator the grid's parameters name
    if (g.prm.op != matrixT::op) throw std::invalid_argument("bicgstabRe2: cfd_params.op names another operator");
    g.put(CFD_RHS, rhs);
    g.put(CFD_X, x);
    int it = 0;
    double rel = 0.0;
    check(cfd_bicgstab(g.ctx(), tol_, iters_max_, &it, &rel), "cfd_bicgstab");
    g.get(CFD_X, x);
    return {it, rel};
  }

 private:
  matrixT *A_;
  double tol_ = 1e-3;
  int iters_max_ = 5000;
};
}  // namespace solver

}  // namespace cfdgpu
