"""SOR on the GPU against the oracle.

SorPipeLine (P_SOLVER_SOR): the volume-weighted update on the ghosted
pressure with the reference's max |dp| stop, in red-black order; every cell's
arithmetic is the oracle's red-black restatement, so the pressure, the sweep
count and the error are bit for bit the oracle's.  Red-black SOR on the CSR
operator likewise.  Timesteps with either solver follow the oracle's
trajectory within 1e-10.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from helpers import rel_l2
from cfd_nssolver_forhpc_amd import capi, configs
from cfd_nssolver_forhpc_amd.solver import Simulation, uniform_grid
from test_gpu_parity import sml_small

pytestmark = pytest.mark.gpu


def sor_case(seed):
    grid = sml_small()
    g = O.Grid(grid)
    rng = np.random.default_rng(seed)
    u1 = [1.0 + 0.05 * rng.standard_normal(g.G), 0.05 * rng.standard_normal(g.G), 0.05 * rng.standard_normal(g.G)]
    p = rng.standard_normal(g.G)
    return grid, g, u1, p


@pytest.mark.parametrize("crit,itmax", [(0.0, 40), (1e-9, 100000)])
def test_sor_pipeline_matches_oracle(require_gpu, crit, itmax):
    grid, g, u1, p0 = sor_case(3)
    prm = configs.cylinder_params(100.0, 0.002, 2.0, 1.0, 0.5, crit)
    prm.solver = capi.CFD_SOLVER_SOR_PIPELINE
    sim = Simulation(grid, prm)
    sim.init_state()
    for c in range(3):
        sim.upload(capi.T1_U + c, u1[c])
    sim.upload(capi.P, p0)
    it, err = sim.SorPipeLine(omega=1.8, crit=crit, itmax=itmax)
    p = p0.copy()
    oit = C.c_int()
    oerr = O.lib().orc_sor_rb_vw(g.ref, 0.002, 1.8, crit, itmax, *u1, p, C.byref(oit))
    print(f"SorPipeLine crit={crit}: sweeps-1 {it} (oracle {oit.value}), error {err:.3e}", flush=True)
    assert it == oit.value and err == oerr
    assert np.array_equal(sim.download(capi.P), p)
    if crit > 0.0:
        assert err <= crit and it < itmax
    else:
        assert it == itmax - 1
    sim.close()


def test_rbsor_csr_matches_oracle(require_gpu):
    grid = uniform_grid((24, 16, 8), (6.0, 4.0, 2.0))
    prm = configs.cylinder_params(100.0, 0.002, 2.0, 2.0, 1.0, 1e-10)
    prm.solver, prm.op = capi.CFD_SOLVER_RBSOR, capi.CFD_OP_CSR
    sim = Simulation(grid, prm)
    sim.init_state()
    sim.createPressureMatrixCSR()
    ptr, col, val = sim.csr()
    g = O.Grid(grid)
    b = np.random.default_rng(4).standard_normal(g.N)
    sim.upload(capi.RHS, b)
    sim.upload(capi.X, np.zeros(g.N))
    it, rel = sim.rbsor(omega=1.8, tol=1e-300, maxit=40, check_every=10)
    ox = np.zeros(g.N)
    oit = C.c_int()
    orel = O.lib().orc_rbsor_csr(g.N, g.ref, ptr, col, val, b, ox, 1.8, 1e-300, 40, 10, C.byref(oit))
    assert it == oit.value == 40
    assert np.array_equal(sim.download(capi.X), ox)
    assert abs(rel - orel) <= 1e-12 * orel
    sim.close()


def test_steps_sor_pipeline_match_oracle(require_gpu):
    """P_SOLVER_SOR timesteps: the cylinder channel (Dirichlet outflow
    pressure, so the ghost pressures pin the level), p_criteria 1e-11."""
    w = configs.small_cylinder(ncal=(32, 16, 8), lengths=(8.0, 4.0, 2.0), quirk=0, tol=1e-11)
    w.params.solver = capi.CFD_SOLVER_SOR_PIPELINE
    w.params.maxit = 200000
    sim = Simulation(w.grid, w.params)
    sim.setup(noise=w.noise, seed=w.seed)
    osim = O.Sim(O.Grid(w.grid), w.params)
    osim.init_state(w.noise, w.seed)
    osim.setup()
    for k in range(3):
        st = sim.step()
        ost = osim.step()
        print(f"SOR step {k + 1}: sweeps {list(st.iters[:st.n_passes])} oracle {list(ost.iters[:ost.n_passes])}",
              flush=True)
        for c in range(3):
            e = rel_l2(sim.download(capi.T0_U + c), osim.field(capi.T0_U + c))
            assert e < 1e-10, (k, c, e)
        assert rel_l2(sim.download(capi.P), osim.field(capi.P)) < 1e-9
        for q in range(st.n_passes):
            assert (st.iters[q] < 0) == (ost.iters[q] < 0)
            if st.iters[q] >= 0:
                assert abs(st.iters[q] - ost.iters[q]) <= max(2, ost.iters[q] // 50)
                assert st.relres[q] <= 1e-11
            assert abs(st.cd[q] - ost.cd[q]) <= 1e-9 * max(1.0, abs(ost.cd[q]))
    sim.close()


def test_steps_rbsor_csr_match_oracle(require_gpu):
    """Red-black SOR on the CSR operator inside the timestep (closed cavity,
    the CSR operator's all-Neumann closure; means removed)."""
    grid = sml_small()
    p = configs.cylinder_params(100.0, 0.002, 1.5, 1.0, 0.5, 1e-11)
    p.set_bc(capi.CAVITY_A_DIR, capi.CAVITY_A_NUM)
    for f in range(6):
        p.p_dirichlet[f] = 0
    p.u_init[0] = 0.0
    p.solver, p.op = capi.CFD_SOLVER_RBSOR, capi.CFD_OP_CSR
    p.maxit, p.sor_check = 200000, 10
    sim = Simulation(grid, p)
    sim.setup(noise=1e-3, seed=3)
    osim = O.Sim(O.Grid(grid), p)
    osim.init_state(1e-3, 3)
    osim.setup()
    demean = lambda a: a - a.mean()
    for k in range(3):
        st = sim.step()
        ost = osim.step()
        for c in range(3):
            e = rel_l2(sim.download(capi.T0_U + c), osim.field(capi.T0_U + c))
            assert e < 1e-10, (k, c, e)
        assert rel_l2(demean(sim.download(capi.X)), demean(osim.field(capi.X))) < 1e-8
        for q in range(st.n_passes):
            assert (st.iters[q] < 0) == (ost.iters[q] < 0)
            if st.iters[q] >= 0:
                assert st.relres[q] < 1e-11
    sim.close()
