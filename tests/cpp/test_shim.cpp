// The reference's prediction-correction loop (inc/run/RUN_CPU_PC_METHOD.hpp
// :287-627) written against cfd_reference_shim.hpp -- host vectors, one
// reference function per call -- must reproduce the device-resident fused
// timestep (grid::step, cfd_step) on the same inputs.
#include <sys/stat.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "cfd_reference_shim.hpp"

using namespace cfdgpu;

static double rel_l2(const std::vector<double> &a, const std::vector<double> &b) {
  double num = 0.0, den = 0.0;
  for (size_t i = 0; i < a.size(); ++i) {
    num += (a[i] - b[i]) * (a[i] - b[i]);
    den += b[i] * b[i];
  }
  return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

int main() {
  cfd_params prm;
  check(cfd_params_default(&prm), "cfd_params_default");
  // cylinder in a channel: inflow x-, outflow x+ (p = 0), slip y/z walls
  const double dir[6][3] = {{1, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  const double num[6][3] = {{0, 0, 0}, {1, 1, 1}, {1, 0, 1}, {1, 0, 1}, {1, 1, 0}, {1, 1, 0}};
  std::memcpy(prm.bc_dir, dir, sizeof(dir));
  std::memcpy(prm.bc_num, num, sizeof(num));
  for (int f = 0; f < 6; ++f) prm.p_dirichlet[f] = f == 1;
  prm.Re = 100.0;
  prm.dt = 0.002;
  prm.tol = 1e-12;
  prm.ib_cx = 2.0;
  prm.ib_cy = 2.0;
  prm.ib_r = 0.5;

  bool vf_ok = true;
  // ---- reference-shaped driver, host vectors ----
  grid gA(32, 16, 8, 8.0, 4.0, 2.0, prm);
  simuClass simu;
  simu.set_Re(prm.Re);
  simu.set_time_max(3 * prm.dt);
  simu.set_dt(prm.dt);
  simu.p_criteria = prm.tol;
  DfibArray Dfib;
  pressure t1;
  velocity T0, T1, T3;
  MxClass Mx;
  calDomain dom = gA.domain();
  resize_variable(gA, t1, T0, T1, T3, Dfib);
  std::fill(T0.u.begin(), T0.u.end(), 1.0);
  T1 = T0;
  T3 = T0;
  BC_staggered_main(dom, T0, t1, gA);
  BC_staggered_copy(dom, T0, T1, gA);
  shareMenory ShareM;
  Mx.X_result.assign(gA.iceltotCal, 0.0);
  createPressureMatrix(Mx, simu, dom, gA);
  solver::bicgstabRe2<ELL_type> pSolver(Mx.matA_ell);  // RUN_CPU_PC_METHOD.hpp:137
  pSolver.setTolerance(simu.p_criteria);
  DFIB_CylinderZ(Dfib, dom, gA);
  auto source_termF = Dfib.f;
  const size_t N = gA.iceltotCal;
  for (simu.loop = 1; simu.get_finishloop(); simu.finishloop()) {
    std::fill(source_termF.begin(), source_termF.end(), 0.0);
    for (int pass = 0; pass < 4; ++pass) {
      ConvectionDifussion(simu, T0, T1, dom, gA);
      if (pass > 0)  // (pass 0 adds zeros)
        for (int c = 0; c < 3; ++c)
          for (int i = dom.i_begin; i < dom.i_endof; ++i)
            for (int j = dom.j_begin; j < dom.j_endof; ++j)
              for (int k = dom.k_begin; k < dom.k_endof; ++k)
                T1.U(c)[gA.icel(i, j, k)] += source_termF[gA.icelCal(i, j, k) + N * c];
      if (simu.loop > 1) {
        createBMatrix_omp(T1, Mx, simu, dom, gA);
        std::tie(simu.iters, simu.error) = pSolver(Mx.matB, Mx.X_result);
        Pressure_transform_X_result(t1, Mx, dom, gA);
      }
      update_UandF_omp(Dfib, simu, t1, T1, T3, dom, gA);
      if (pass < 3)
        for (size_t m = 0; m < 3 * N; ++m) source_termF[m] += simu.dt * Dfib.f[m];
    }
    auto [uDif, vDif, wDif] = CheckSteadyStatebyMaxVal_omp(simu, ShareM, T0, T3, dom, gA);  // :581
    (void)vDif; (void)wDif;
    T0 = T3;
    BC_staggered_main(dom, T0, t1, gA);
    BC_staggered_copy(dom, T0, T1, gA);
    virtualF_Int(Dfib, dom, gA, 0);
    virtualF_Int(Dfib, dom, gA, 1);
    const double cD = -2.0 * Dfib.ValSum[0] / gA.lz;
    const auto cdl = drag_lift(Dfib, gA);
    vf_ok = vf_ok && std::fabs(cD - cdl.first) <= 1e-12 * std::fabs(cdl.first) &&
            std::fabs(-2.0 * Dfib.ValSum[1] / gA.lz - cdl.second) <= 1e-12 * std::fabs(cdl.second) + 1e-300;
    std::printf("loop %zu iters %d err %.3e cD %.6f maxdiff u %.3e\n", simu.loop, simu.iters, simu.error, cD, uDif);
  }

  // ---- device-resident fused timestep ----
  prm.source_index_quirk = 0;
  grid gB(32, 16, 8, 8.0, 4.0, 2.0, prm);
  check(cfd_init_state(gB.ctx(), 0.0, 0), "cfd_init_state");
  check(cfd_ell_assemble(gB.ctx()), "cfd_ell_assemble");
  check(cfd_set_eta_cylinder_z(gB.ctx()), "cfd_set_eta");
  for (int s = 0; s < 3; ++s) gB.step();
  velocity V;
  gB.get(CFD_T0_U, V);
  std::vector<double> x;
  gB.get(CFD_X, x);
  // w carries only roundoff in this z-uniform flow, so its error is measured
  // against |u|; X is a solve to 1e-12 relative residual on an operator with
  // condition number ~1e3, so two solves may differ by ~1e-9
  double dw = 0.0, nu2 = 0.0;
  for (size_t m = 0; m < V.w.size(); ++m) {
    dw += (V.w[m] - T0.w[m]) * (V.w[m] - T0.w[m]);
    nu2 += T0.u[m] * T0.u[m];
  }
  const double ew = std::sqrt(dw / nu2);
  const double eu = rel_l2(V.u, T0.u), ev = rel_l2(V.v, T0.v), ex = rel_l2(x, Mx.X_result);
  std::printf("shim vs fused step: u %.2e v %.2e w %.2e X %.2e\n", eu, ev, ew, ex);
  // the CSR operator through the reference-shaped types: the tuple that
  // createPressureMatrixCSR returns, applied on the host with the FMA chain
  // of CSR_matrix::multiply_omp, equals the device product bit for bit
  auto csr = gA.createPressureMatrixCSR();
  mat::CSR_matrix Ac(gA);
  Ac.set(csr);
  const auto &cptr = std::get<0>(csr);
  const auto &cidx = std::get<1>(csr);
  const auto &cval = std::get<2>(csr);
  std::vector<double> xs(N), yd, yh(N);
  for (size_t r = 0; r < N; ++r) xs[r] = std::sin(0.37 * (double)r);
  Ac.multiply_omp(xs, yd);
  for (size_t r = 0; r < N; ++r) {
    double acc = 0.0;
    for (int e = cptr[r]; e < cptr[r + 1]; ++e) acc = std::fma(cval[e], xs[cidx[e]], acc);
    yh[r] = acc;
  }
  const bool csr_ok = cptr.size() == N + 1 && (size_t)cptr.back() == cidx.size() && yd == yh;
  // a solver over the CSR type on a grid configured for the ELL operator is refused
  bool op_threw = false;
  try {
    solver::bicgstabRe2<mat::CSR_matrix> sc(Ac);
    sc(Mx.matB, xs);
  } catch (const std::invalid_argument &) {
    op_threw = true;
  }
  std::printf("CSR: nnz %zu product %s, solver on another operator %s\n", cidx.size(), csr_ok ? "exact" : "DIFFERS",
              op_threw ? "refused" : "NOT refused");
  // output through the reference-shaped calls, in a scratch directory: the
  // grid and solution files, QfileRead of the latter, and the drag history
  char tmpl[] = "/tmp/cfd_shim_io_XXXXXX";
  bool io_ok = mkdtemp(tmpl) != nullptr && chdir(tmpl) == 0 && mkdir("mx_out", 0755) == 0 &&
               mkdir("Information", 0755) == 0;
  if (io_ok) {
    OutputPlot3D_Xfile(simu, gA);
    simu.currentFile = 3;
    OutputPLOT3D_Qfile(Dfib, simu, t1, T3, gA);
    auto [nb, mach, alpha, reyn, qtime, qp, qu, qv, qw, qeta] = QfileRead(gA, "mx_out/P3D03.q");
    (void)mach; (void)alpha; (void)qu; (void)qv; (void)qw;
    io_ok = nb == 1 && reyn == (float)prm.Re && qtime == (float)simu.time;
    for (int i = dom.i_begin; i < dom.i_endof && io_ok; ++i)
      for (int j = dom.j_begin; j < dom.j_endof; ++j)
        for (int k = dom.k_begin; k < dom.k_endof; ++k) {
          const size_t m = gA.icel(i, j, k);
          io_ok = io_ok && qp[m] == (double)(float)t1.p[m] && qeta[m] == (double)(float)Dfib.eta[m];
        }
    simu.delayIO = 2;
    simu.zone = ZONE(prm);
    for (simu.loop = 1; simu.loop <= 4; ++simu.loop) Delay_IO_CD_CL(simu, dom, Dfib, gA);
    std::ifstream h("Information/Time_cDcL.dat");
    int lines = 0;
    for (std::string l; std::getline(h, l);) ++lines;
    io_ok = io_ok && lines == 3 + 4;  // header, then two flushes of two lines
  }
  std::printf("output files: %s\n", io_ok ? "ok" : "FAILED");
  // ---- P_SOLVER_SOR: SorPipeLine_omp in the loop (RUN_CPU_PC_METHOD.hpp:333)
  //      against the fused device step with the same solver ----
  bool sor_ok = true;
  {
    cfd_params ps = prm;
    ps.solver = CFD_SOLVER_SOR_PIPELINE;
    ps.tol = 1e-11;
    ps.maxit = 200000;
    grid gS(32, 16, 8, 8.0, 4.0, 2.0, ps);
    simuClass su;
    su.set_Re(ps.Re);
    su.set_time_max(3 * ps.dt);
    su.set_dt(ps.dt);
    su.p_criteria = ps.tol;
    su.p_sor_iter_max = 200000;
    SORcoefficient Sor;
    DfibArray Df;
    pressure ts;
    velocity S0, S1, S3;
    resize_variable(gS, ts, S0, S1, S3, Df);
    std::fill(S0.u.begin(), S0.u.end(), 1.0);
    S1 = S0;
    S3 = S0;
    calDomain ds = gS.domain();
    BC_staggered_main(ds, S0, ts, gS);
    BC_staggered_copy(ds, S0, S1, gS);
    DFIB_CylinderZ(Df, ds, gS);
    auto src = Df.f;
    std::vector<int> sweeps;
    for (su.loop = 1; su.get_finishloop(); su.finishloop()) {
      std::fill(src.begin(), src.end(), 0.0);
      for (int pass = 0; pass < 4; ++pass) {
        ConvectionDifussion(su, S0, S1, ds, gS);
        if (pass > 0)
          for (int c = 0; c < 3; ++c)
            for (int i = ds.i_begin; i < ds.i_endof; ++i)
              for (int j = ds.j_begin; j < ds.j_endof; ++j)
                for (int k = ds.k_begin; k < ds.k_endof; ++k)
                  S1.U(c)[gS.icel(i, j, k)] += src[gS.icelCal(i, j, k) + N * c];
        if (su.loop > 1) {
          SorPipeLine_omp(Sor, su, S1, ts, ds, gS);
          sweeps.push_back(su.iters);
          sor_ok = sor_ok && su.error <= su.p_criteria;
        }
        update_UandF_omp(Df, su, ts, S1, S3, ds, gS);
        if (pass < 3)
          for (size_t m = 0; m < 3 * N; ++m) src[m] += su.dt * Df.f[m];
      }
      S0 = S3;
      BC_staggered_main(ds, S0, ts, gS);
      BC_staggered_copy(ds, S0, S1, gS);
    }
    ps.source_index_quirk = 0;
    grid gF(32, 16, 8, 8.0, 4.0, 2.0, ps);
    check(cfd_init_state(gF.ctx(), 0.0, 0), "cfd_init_state");
    check(cfd_set_eta_cylinder_z(gF.ctx()), "cfd_set_eta");
    std::vector<int> fused;
    for (int s = 0; s < 3; ++s) {
      cfd_step_stats st = gF.step();
      for (int q = 0; q < st.n_passes; ++q)
        if (st.iters[q] >= 0) fused.push_back(st.iters[q]);
    }
    velocity F;
    gF.get(CFD_T0_U, F);
    pressure fp;
    gF.get(CFD_P, fp.p);
    const double e_u = rel_l2(F.u, S0.u), e_p = rel_l2(fp.p, ts.p);
    // the host-side source accumulation may round differently from the
    // device's (contracted) one, so a pass may stop one sweep apart
    bool close = sweeps.size() == fused.size() && !sweeps.empty();
    for (size_t q = 0; close && q < sweeps.size(); ++q) close = std::abs(sweeps[q] - fused[q]) <= 1;
    sor_ok = sor_ok && e_u < 1e-10 && e_p < 1e-8 && close;
    std::printf("SorPipeLine loop vs fused SOR step: u %.2e p %.2e, %zu passes, sweeps %d..%d %s\n", e_u, e_p,
                sweeps.size(), sweeps.empty() ? 0 : sweeps.front(), sweeps.empty() ? 0 : sweeps.back(),
                close ? "within 1" : "DIFFER");
  }

  // ---- mat::ELL_matrix<double>: the reference container on the device ----
  bool ell_ok = true;
  {
    cfd_params pe = prm;
    grid gE(10, 5, 4, 1.0, 1.0, 1.0, pe);  // 200 interior cells
    const int n = 200;
    ELL_type E(n, 3);
    for (int r = 0; r < n; ++r) {
      if (r > 0) E.set(r, r - 1, -1.0);
      E.set(r, r, 1.0);  // overwritten below
      if (r < n - 1) E.set(r, r + 1, -1.0);
      E.set(r, r, 2.0 + 0.01 * (r % 7));
    }
    E.set(5, 4, 0.0);  // removes the entry and shifts the row
    bool threw = false;
    try {
      E.set(7, 9, 1.0);  // a fourth entry in a row of three
    } catch (const std::invalid_argument &) {
      threw = true;
    }
    E.bind(gE);
    std::vector<double> xe(n), ye, yh(n);
    for (int r = 0; r < n; ++r) xe[r] = std::cos(0.11 * r);
    E.multiply_omp(xe, ye);
    auto csr = E.get_CSR();
    const auto &cp = std::get<0>(csr);
    const auto &ci = std::get<1>(csr);
    const auto &cv = std::get<2>(csr);
    // host product in each row's insertion order: the device's FMA chain
    for (int r = 0; r < n; ++r) {
      double acc = 0.0;
      std::vector<std::pair<int, double>> row;
      if (r > 0 && r != 5) row.push_back({r - 1, -1.0});
      row.push_back({r, 2.0 + 0.01 * (r % 7)});
      if (r < n - 1) row.push_back({r + 1, -1.0});
      for (auto &q : row) acc = std::fma(q.second, xe[q.first], acc);
      for (size_t s = row.size(); s < 3; ++s) acc = std::fma(0.0, xe[n - 1], acc);
      yh[r] = acc;
    }
    ell_ok = threw && ye == yh && cp.back() == 3 * n - 3 && ci[cp[5]] == 5 && cv[cp[5]] == 2.0 + 0.01 * 5;
    // the solver keeps its own copy: after the device received another
    // operator, its solve uploads this one again
    solver::bicgstabRe2<ELL_type> sE(E);
    sE.setTolerance(1e-12);
    gE.prm.op = CFD_OP_ELL;
    check(cfd_ell_assemble(gE.ctx()), "cfd_ell_assemble");  // the device now holds the grid's operator
    gE.set_ell_token(0);
    std::vector<double> bE = yh, xE(n, 0.0), rE;
    auto [itE, relE] = sE(bE, xE);
    E.multiply_omp(xE, rE);
    double err = 0.0, nb = 0.0;
    for (int r = 0; r < n; ++r) {
      err += (rE[r] - bE[r]) * (rE[r] - bE[r]);
      nb += bE[r] * bE[r];
    }
    ell_ok = ell_ok && relE < 1e-12 && std::sqrt(err / nb) < 1e-11 && itE > 0;
    // the device-assembled pressure operator through the host container
    auto pc = Mx.matA_ell.get_CSR();
    std::vector<int32_t> dc(7 * N);
    std::vector<double> dv(7 * N);
    check(cfd_ell_download(gA.ctx(), dc.data(), dv.data(), N, 7), "cfd_ell_download");
    size_t nnz = 0;
    for (size_t q = 0; q < 7 * N; ++q) nnz += !(dc[q] == (int32_t)N - 1 && dv[q] == 0.0);
    bool sorted = true;
    for (size_t r = 0; r < N; ++r)
      for (int e = std::get<0>(pc)[r] + 1; e < std::get<0>(pc)[r + 1]; ++e)
        sorted = sorted && std::get<1>(pc)[e - 1] < std::get<1>(pc)[e];
    ell_ok = ell_ok && (size_t)std::get<0>(pc).back() == nnz && sorted && Mx.matA_ell.at(0, 0) == dv[0];
    std::printf("ELL_matrix<double>: set / remove / overflow %s, product %s, solve relres %.1e, get_CSR nnz %zu\n",
                threw ? "ok" : "FAILED", ye == yh ? "exact" : "DIFFERS", relE, nnz);
  }

  // an out-of-range argument surfaces as the reference's exception type
  bool threw = false;
  try {
    check(cfd_axis_sml(20, 1.0, 0.3, 7, 0, x.data(), x.data(), x.data()), "cfd_axis_sml");
  } catch (const std::invalid_argument &) {
    threw = true;
  }
  std::printf("virtualF_Int per direction vs cfd_drag: %s\n", vf_ok ? "ok" : "FAILED");
  const bool ok = eu < 1e-11 && ev < 1e-11 && ew < 1e-11 && ex < 1e-8 && threw && csr_ok && op_threw && io_ok &&
                  vf_ok && sor_ok && ell_ok;
  std::printf("%s\n", ok ? "PASS" : "FAIL");
  return ok ? 0 : 1;
}
