// cfd_reference_shim.hpp -- header-only C++17 mirror of the reference
// solver's hot-path interfaces on top of the C-ABI in cfd_capi.h.
//
// The types and functions keep the reference's names and argument lists so
// that a driver written against the reference (inc/run/RUN_CPU_PC_METHOD.hpp)
// ports by changing its namespace: simuClass, grid, velocity, pressure,
// DfibArray, MxClass, shareMenory, SORcoefficient, calDomain,
// ConvectionDifussion, createBMatrix_omp, createPressureMatrix,
// update_UandF_omp, Pressure_transform_X_result, BC_staggered_main,
// BC_staggered_copy, DFIB_CylinderZ, virtualF_Int,
// CheckSteadyStatebyMaxVal_omp / _seq, SorPipeLine_omp / _seq,
// Delay_IO_CD_CL, noDelay_IO_CD_CL, OutputPlot3D_Xfile, OutputPLOT3D_Qfile,
// QfileRead, grid::createPressureMatrixCSR, mat::ELL_matrix<T>,
// mat::CSR_matrix, solver::bicgstabRe2.  Fields stay in
// caller-owned std::vector<double> in the reference layout; each call
// uploads its inputs, runs on the GPU and downloads its outputs, so it is a
// migration path, not the fast path.  The fast path keeps the state on the
// device: grid::step() runs one whole timestep (cfd_step) without host traffic.
//
// Errors: a status CFD_ERR_ARG throws std::invalid_argument (where the
// reference throws it), every other failure std::runtime_error.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
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
  size_t p_sor_iter_max = 5000;  // simuClass.hpp:80
  bool firstSOR = true;
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
  if (p.solver == CFD_SOLVER_RBSOR || p.solver == CFD_SOLVER_SOR_PIPELINE) return "SOR";
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
// inc/physicalVariables.hpp:159-169
struct DfibArray {
  std::vector<double> f, eta, Val_sumz, Val_sumz_sumy, ValSum;
  double cylinderDimension = 0.0;
  std::vector<double> cylinderCenter;
};
// inc/physicalVariables.hpp:172-185
struct shareMenory {
  double pChangeMax = 0.0, Out = 0.0, uDif_Max = 0.0, vDif_Max = 0.0, wDif_Max = 0.0;
};
// inc/pressure/PressureClass.hpp:48-50.  SorPipeLine forms its face
// coefficients per cell on the device, so cf stays empty here.
struct SORcoefficient { std::vector<double> cf; };

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
      prm = o.prm; ctx_ = o.ctx_; o.ctx_ = nullptr; ell_token_ = o.ell_token_;
    }
    return *this;
  }
  ~grid() { release(); }

  size_t icel(int i, int j, int k) const { return ((size_t)i * ny + j) * nz + k; }
  size_t icelCal(int i, int j, int k) const {
    return ((size_t)(i - gC) * nyCal + (j - gC)) * nzCal + (k - gC);
  }
  cfd_ctx *ctx() const { return ctx_; }
  // which host ELL matrix the device operator currently holds (0: none, or
  // one assembled on the device by createPressureMatrix and not yet named)
  uint64_t ell_token() const { return ell_token_; }
  void set_ell_token(uint64_t t) { ell_token_ = t; }
  static uint64_t new_token() {
    static uint64_t next = 0;
    return ++next;
  }
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
  uint64_t ell_token_ = 0;

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

namespace mat {
// inc/matrix/ELL_sparseMatrix.hpp:30-377: an ELL matrix of rows x cols with
// at most max_idx_size entries a row.  The host copy keeps the reference's
// set() semantics (an existing column is overwritten, a zero value removes
// the entry, entries stay in insertion order, empty slots hold (cols - 1,
// 0)); the device holds the operator of the grid the matrix is bound to
// (createPressureMatrix binds it to the operator it assembled on the device;
// bind() lets a matrix built with set() use a grid's context).
// multiply_omp runs on the device (cfd_ell_upload when the device holds
// another matrix, then cfd_spmv): rows must equal the grid's interior cell
// count and max_idx_size must be at most 7, the device layout's slots.
template <typename T>
class ELL_matrix {
 public:
  using CSR_type = std::tuple<std::vector<int>, std::vector<int>, std::vector<T>>;
  static constexpr int op = CFD_OP_ELL;

  ELL_matrix() {}
  ELL_matrix(int rows, int cols, int max_idx_size) { construct(rows, cols, max_idx_size); }
  ELL_matrix(int n, int max_idx_size) { construct(n, n, max_idx_size); }

  void resize(int rows, int cols, int max_idx_size) { construct(rows, cols, max_idx_size); }
  void resize(int n, int max_idx_size) { construct(n, n, max_idx_size); }
  int row() const { return rows_; }
  int col() const { return cols_; }
  int MaxRow() const { return max_idx_size_; }

  // ELL_matrix::set(row, col, val) (ELL_sparseMatrix.hpp:246-309)
  ELL_matrix &set(int row, int col, T val) {
    if (row < 0 || col < 0) throw std::invalid_argument("Matrix dimensions cannot be negative.");
    if (col >= cols_) throw std::invalid_argument("col >= cols_");
    if (row >= rows_) throw std::invalid_argument("row >= rows_");
    ensure_host();
    token_ = 0;  // the host copy now differs from any device copy
    const size_t pt = (size_t)row * max_idx_size_;
    int &n = count_[row];
    for (int i = 0; i < n; ++i) {
      if (idx_[pt + i] != col) continue;
      if (val == T()) {  // remove the entry, shift the rest forward
        for (int j = i; j + 1 < n; ++j) {
          idx_[pt + j] = idx_[pt + j + 1];
          val_[pt + j] = val_[pt + j + 1];
        }
        --n;
        idx_[pt + n] = cols_ - 1;
        val_[pt + n] = T();
      } else {
        val_[pt + i] = val;
      }
      return *this;
    }
    if (val == T()) return *this;
    if (n >= max_idx_size_) throw std::invalid_argument("[ELL::set] -> curr_idx_size[row] >max_idx_size_");
    idx_[pt + n] = col;
    val_[pt + n] = val;
    ++n;
    return *this;
  }
  // ELL_matrix::set(CSR_type) (ELL_sparseMatrix.hpp:220-244): the rows of a
  // CSR triple, in its entry order
  ELL_matrix &set(const CSR_type &c) {
    const auto &ptr = std::get<0>(c);
    const auto &idx = std::get<1>(c);
    const auto &val = std::get<2>(c);
    if (ptr.empty()) throw std::invalid_argument("ELL_matrix::set: empty CSR row pointers");
    int width = 0;
    for (size_t r = 0; r + 1 < ptr.size(); ++r) width = std::max(width, ptr[r + 1] - ptr[r]);
    construct((int)ptr.size() - 1, (int)ptr.size() - 1, width);
    for (size_t r = 0; r + 1 < ptr.size(); ++r)
      for (int e = ptr[r]; e < ptr[r + 1]; ++e) set((int)r, idx[e], val[e]);
    return *this;
  }
  // the stored value at (row, col), 0 when absent
  T at(int row, int col) {
    ensure_host();
    const size_t pt = (size_t)row * max_idx_size_;
    for (int i = 0; i < count_.at(row); ++i)
      if (idx_[pt + i] == col) return val_[pt + i];
    return T();
  }

  // the device context this matrix runs on
  void bind(grid &g) { g_ = &g; }
  // createPressureMatrix: the operator the device assembled for grid g
  void bind_assembled(grid &g) {
    g_ = &g;
    rows_ = cols_ = (int)g.iceltotCal;
    max_idx_size_ = 7;
    host_ok_ = false;
    token_ = grid::new_token();
    g.set_ell_token(token_);
  }

  // r = A x on the device (multiply_omp, ELL_sparseMatrix.hpp:477-492)
  bool multiply_omp(std::vector<T> &x, std::vector<T> &r) {
    static_assert(std::is_same<T, double>::value, "the device operator holds FP64 values");
    to_device();
    g_->put(CFD_VEC_A, x);
    check(cfd_spmv(g_->ctx(), CFD_OP_ELL, CFD_FP64, CFD_VEC_A, CFD_VEC_B), "cfd_spmv");
    g_->get(CFD_VEC_B, r);
    return true;
  }
  std::vector<T> multiply_omp(std::vector<T> &x) {
    std::vector<T> r;
    multiply_omp(x, r);
    return r;
  }

  // get_CSR (ELL_sparseMatrix.hpp:351-377): each row's entries sorted by
  // column (sort_Idx), then row pointers, columns and values
  CSR_type get_CSR() {
    ensure_host();
    std::vector<int> ptr((size_t)rows_ + 1, 0), idx;
    std::vector<T> val;
    for (int r = 0; r < rows_; ++r) {
      const size_t pt = (size_t)r * max_idx_size_;
      std::vector<std::pair<int, T>> e;
      for (int i = 0; i < count_[r]; ++i) e.emplace_back(idx_[pt + i], val_[pt + i]);
      std::sort(e.begin(), e.end());
      for (const auto &q : e) {
        idx.push_back(q.first);
        val.push_back(q.second);
      }
      ptr[(size_t)r + 1] = (int)idx.size();
    }
    return {std::move(ptr), std::move(idx), std::move(val)};
  }

  grid &owner() {
    if (!g_) throw std::invalid_argument("ELL_matrix: not bound to a grid");
    return *g_;
  }
  // make the device hold this matrix (solver::bicgstabRe2 calls it)
  void to_device() {
    grid &g = owner();
    if (token_ != 0 && g.ell_token() == token_) return;
    ensure_host();
    if ((size_t)rows_ != g.iceltotCal || cols_ != rows_)
      throw std::invalid_argument("ELL_matrix: rows must equal the grid's interior cell count");
    if (max_idx_size_ > 7) throw std::invalid_argument("ELL_matrix: the device layout holds 7 entries a row");
    const size_t n = (size_t)rows_;
    std::vector<int32_t> col(7 * n, (int32_t)(cols_ - 1));
    std::vector<double> val(7 * n, 0.0);
    for (size_t r = 0; r < n; ++r)
      for (int s = 0; s < max_idx_size_; ++s) {
        col[(size_t)s * n + r] = idx_[r * max_idx_size_ + s];
        val[(size_t)s * n + r] = (double)val_[r * max_idx_size_ + s];
      }
    check(cfd_ell_upload(g.ctx(), col.data(), val.data(), n, 7), "cfd_ell_upload");
    if (token_ == 0) token_ = grid::new_token();
    g.set_ell_token(token_);
  }

 private:
  int rows_ = 0, cols_ = 0, max_idx_size_ = 0;
  std::vector<int> idx_, count_;
  std::vector<T> val_;
  grid *g_ = nullptr;
  uint64_t token_ = 0;   // the device copy of this content (grid::ell_token)
  bool host_ok_ = true;  // false: the content exists only on the device

  void construct(int rows, int cols, int max_idx_size) {
    if (rows < 0 || cols < 0) throw std::invalid_argument("Matrix dimensions cannot be negative.");
    rows_ = rows;
    cols_ = cols;
    max_idx_size_ = max_idx_size;
    idx_.assign((size_t)rows * max_idx_size, cols - 1);
    val_.assign((size_t)rows * max_idx_size, T());
    count_.assign((size_t)rows, 0);
    host_ok_ = true;
    token_ = 0;
  }
  // the host copy of a device-assembled operator (cfd_ell_download); its
  // padding slots are the trailing (cols - 1, 0) entries
  void ensure_host() {
    if (host_ok_) return;
    grid &g = owner();
    if (g.ell_token() != token_) throw std::runtime_error("ELL_matrix: the device no longer holds this operator");
    const size_t n = (size_t)rows_;
    std::vector<int32_t> col(7 * n);
    std::vector<double> val(7 * n);
    check(cfd_ell_download(g.ctx(), col.data(), val.data(), n, 7), "cfd_ell_download");
    idx_.assign(7 * n, cols_ - 1);
    val_.assign(7 * n, T());
    count_.assign(n, 0);
    for (size_t r = 0; r < n; ++r) {
      int last = 0;
      for (int s = 0; s < 7; ++s) {
        idx_[r * 7 + s] = col[(size_t)s * n + r];
        val_[r * 7 + s] = (T)val[(size_t)s * n + r];
        if (!(col[(size_t)s * n + r] == cols_ - 1 && val[(size_t)s * n + r] == 0.0)) last = s + 1;
      }
      count_[r] = last;
    }
    host_ok_ = true;
  }
};
}  // namespace mat

using ELL_type = mat::ELL_matrix<double>;

// inc/pressure/PressureClass.hpp:24-45
struct MxClass {
  std::vector<double> matB, X_result;
  mat::ELL_matrix<double> matA_ell;
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

// inc/matrix/PressureMatrix.hpp:5-117: the operator is assembled on the
// device; Mx.matA_ell refers to it (its host copy is fetched on demand)
inline bool createPressureMatrix(MxClass &Mx, simuClass &, calDomain &, grid &g) {
  check(cfd_ell_assemble(g.ctx()), "cfd_ell_assemble");
  Mx.matA_ell.bind_assembled(g);
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

// inc/dfib/virtualForceIntergrator.hpp:5-37: Dfib.ValSum[whichDirection] =
// sum over the interior of f[whichDirection] times the cell volume
inline int virtualF_Int(DfibArray &Dfib, calDomain &, grid &g, const int whichDirection) {
  if (Dfib.ValSum.size() < 3) Dfib.ValSum.resize(3);
  g.put(CFD_F, Dfib.f);
  double sum = 0.0;
  check(cfd_force_sum(g.ctx(), whichDirection, &sum), "cfd_force_sum");
  Dfib.ValSum[(size_t)whichDirection] = sum;
  return 0;
}

// cD = -2 sum f_x dV / lz and cL from f_y, the driver's formulas
// (RUN_CPU_PC_METHOD.hpp:403-435), from one device reduction
inline std::pair<double, double> drag_lift(DfibArray &Dfib, grid &g) {
  g.put(CFD_F, Dfib.f);
  double cd = 0.0, cl = 0.0;
  check(cfd_drag(g.ctx(), &cd, &cl), "cfd_drag");
  return {cd, cl};
}

// inc/analyses/CheckSteadystate.hpp:164-226 (the _seq variant, :104-160, is
// the same maximum)
inline std::tuple<double, double, double> CheckSteadyStatebyMaxVal_omp(simuClass &, shareMenory &, velocity &T0,
                                                                       velocity &T3, calDomain &, grid &g) {
  g.put(CFD_T0_U, T0);
  g.put(CFD_T3_U, T3);
  double m[3];
  check(cfd_maxdiff(g.ctx(), m), "cfd_maxdiff");
  return {m[0], m[1], m[2]};
}
inline std::tuple<double, double, double> CheckSteadyStatebyMaxVal_seq(simuClass &simu, shareMenory &ShareM,
                                                                       velocity &T0, velocity &T3, calDomain &Lo,
                                                                       grid &g) {
  return CheckSteadyStatebyMaxVal_omp(simu, ShareM, T0, T3, Lo, g);
}

// inc/pressure/solver/SorPipeLine.seq.hpp:4-122: the volume-weighted SOR on
// t1.p in place, omega = simu.p_sor_omega, stop on max |dp| <= simu.p_criteria
// or after simu.p_sor_iter_max sweeps; simu.iters = sweeps - 1, simu.error =
// the last sweep's max |dp|.  The device sweeps in red-black order.
inline bool SorPipeLine_seq(SORcoefficient &, simuClass &simu, velocity &T1, pressure &t1, calDomain &, grid &g) {
  g.put(CFD_T1_U, T1);
  g.put(CFD_P, t1.p);
  int it = 0;
  double err = 0.0;
  check(cfd_sor_pipeline(g.ctx(), simu.p_sor_omega, simu.p_criteria, (int)simu.p_sor_iter_max, &it, &err),
        "cfd_sor_pipeline");
  g.get(CFD_P, t1.p);
  simu.iters = it;
  simu.error = err;
  return true;
}
// inc/pressure/solver/SorPipeLine.omp.hpp:7-118: the same solve (the
// reference's OpenMP sweep races on p and stops on sqrt(sum |dp| / |rhs|);
// the device sweep is race-free and stops as SorPipeLine_seq does)
inline bool SorPipeLine_omp(SORcoefficient &Sor, simuClass &simu, velocity &T1, pressure &t1, calDomain &Lo,
                            grid &g) {
  return SorPipeLine_seq(Sor, simu, T1, t1, Lo, g);
}

// inc/dfib/virtualForceIntergrator.hpp:105-177: cD, cL of f into simuClass's
// delayIO ring buffer, flushed to Information/Time_cDcL.dat every delayIO steps
// (as written: each flush repeats the current line delayIO times)
inline std::pair<double, double> Delay_IO_CD_CL(simuClass &simu, calDomain &, DfibArray &Dfib, grid &g) {
  auto cdl = drag_lift(Dfib, g);
  simu.cdcl_record("Information/Time_cDcL.dat", CFD_CDCL_REFERENCE, cdl);
  return cdl;
}

// inc/dfib/virtualForceIntergrator.hpp:44-99: one line per step
inline std::pair<double, double> noDelay_IO_CD_CL(simuClass &simu, calDomain &, DfibArray &Dfib, grid &g) {
  auto cdl = drag_lift(Dfib, g);
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
// inc/solver/bicgstabRe2.hpp:13-189.  Like the reference it keeps its own
// copy of the matrix (an ELL_matrix is re-uploaded when the device holds
// another one); the solver state (px1, px2, solve count) lives with the
// grid's device context
template <typename matrixT>
class bicgstabRe2 {
 public:
  explicit bicgstabRe2(const matrixT &A) : A_(A) {}
  void setTolerance(double tol) { tol_ = tol; }
  std::pair<int, double> operator()(const std::vector<double> &rhs, std::vector<double> &x) {
    grid &g = A_.owner();
    // the device solve runs on the operator the grid's parameters name
    if (g.prm.op != matrixT::op) throw std::invalid_argument("bicgstabRe2: cfd_params.op names another operator");
    if constexpr (matrixT::op == CFD_OP_ELL) A_.to_device();
    g.put(CFD_RHS, rhs);
    g.put(CFD_X, x);
    int it = 0;
    double rel = 0.0;
    check(cfd_bicgstab(g.ctx(), tol_, iters_max_, &it, &rel), "cfd_bicgstab");
    g.get(CFD_X, x);
    return {it, rel};
  }

 private:
  matrixT A_;
  double tol_ = 1e-3;
  int iters_max_ = 5000;
};
}  // namespace solver

}  // namespace cfdgpu
