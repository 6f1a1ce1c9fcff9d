"""HIP library against the CPU oracle, function by function and step by step.

Bars: bit-exact for the ELL assembly, the IB mask, the SpMVs, red-black
SOR (same FMA chains on both sides), the source add and accumulation, the
right-hand side, the DFIB update and the BC (the same operations in the same
order, with contraction off on both sides); 1e-13 relative L2 for the
predictor (FMA contraction of a different compiler); 1e-10 relative L2 for
pressure solves and whole timesteps (the FP64 bar of the north star).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from helpers import bc_tables, pd6, random_state, rel_l2
from cfd_nssolver_forhpc_amd import capi, configs
from cfd_nssolver_forhpc_amd.solver import Simulation, native_grid, uniform_grid, axis_sml, axis_tanh
from cfd_nssolver_forhpc_amd.capi import GridArrays

pytestmark = pytest.mark.gpu


def sml_small():
    ax = axis_sml(30, 8.0, 2.4, 12, -2)
    ay = axis_sml(22, 4.0, 1.2, 10, 0)
    az = axis_sml(8, 2.0, 1.0, 0, 0)
    return GridArrays((30, 22, 8), ax[0], ay[0], az[0], ax[1], ay[1], az[1], ax[2], ay[2], az[2], 8.0, 4.0, 2.0)


GRIDS = {
    "uniform": lambda: uniform_grid((24, 16, 8), (6.0, 4.0, 2.0)),
    "sml": sml_small,
}


def make_sim(grid, quirk=1, pd=(0, 0, 0, 0, 0, 0), op=capi.CFD_OP_ELL, tol=1e-12, inner=3):
    p = configs.cylinder_params(100.0, 0.002, grid.lx / 3.0, grid.ly / 2.0, 1.0, tol)
    p.source_index_quirk = quirk
    for f in range(6):
        p.p_dirichlet[f] = pd[f]
    p.op = op
    p.inner_loops = inner
    return Simulation(grid, p), p


@pytest.mark.parametrize("gname", list(GRIDS))
def test_eta_and_ell_bit_exact(require_gpu, gname):
    grid = GRIDS[gname]()
    sim, p = make_sim(grid, pd=(0, 1, 0, 0, 0, 0))
    sim.init_state()
    sim.DFIB_CylinderZ()
    sim.createPressureMatrix()
    sim.setupPressure()
    g = O.Grid(grid)
    eta = np.zeros(g.G)
    O.lib().orc_eta_cylinder_z(g.ref, p.ib_cx, p.ib_cy, p.ib_r, p.ib_nsub, eta)
    assert np.array_equal(sim.download(capi.ETA), eta)
    assert (eta > 0).sum() > 0 and (eta == 1.0).sum() > 0
    col, val = sim.ell()
    ocol = np.zeros(7 * g.N, np.int32)
    oval = np.zeros(7 * g.N)
    ojp = np.zeros(g.N)
    assert O.lib().orc_ell_assemble(g.ref, pd6(p), ocol, oval, ojp) == 0
    assert np.array_equal(col, ocol)
    assert np.array_equal(val, oval)
    assert np.array_equal(sim.download(capi.JP), ojp)


def test_native_grid_ell_and_eta(require_gpu):
    """The reference's own case; the oracle side is pinned to the compiled reference."""
    grid = native_grid()
    p = capi.CfdParams()
    from cfd_nssolver_forhpc_amd.solver import default_params
    p = default_params()
    sim = Simulation(grid, p)
    sim.init_state()
    sim.DFIB_CylinderZ()
    sim.createPressureMatrix()
    g = O.Grid(grid)
    eta = np.zeros(g.G)
    O.lib().orc_eta_cylinder_z(g.ref, 6.5, 6.0, 0.5, 100, eta)
    assert np.array_equal(sim.download(capi.ETA), eta)
    col, val = sim.ell()
    ocol = np.zeros(7 * g.N, np.int32)
    oval = np.zeros(7 * g.N)
    ojp = np.zeros(g.N)
    O.lib().orc_ell_assemble(g.ref, pd6(p), ocol, oval, ojp)
    assert np.array_equal(col, ocol) and np.array_equal(val, oval)


@pytest.mark.parametrize("gname", list(GRIDS))
def test_spmv_bit_exact(require_gpu, gname):
    grid = GRIDS[gname]()
    sim, p = make_sim(grid)
    sim.init_state()
    sim.createPressureMatrix()
    sim.setupPressure()
    g = O.Grid(grid)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(g.N)
    sim.upload(capi.X, x)
    col, val = sim.ell()
    y = np.zeros(g.N)
    O.lib().orc_spmv_ell(g.N, 7, col, val, x, y)
    sim.spmv(capi.X, capi.VEC_A, op=capi.CFD_OP_ELL)
    assert np.array_equal(sim.download(capi.VEC_A), y)
    coef = np.zeros(2 * (g.nxc + g.nyc + g.nzc))
    jp = np.zeros(g.N)
    O.lib().orc_spe_setup(g.ref, pd6(p), coef, jp)
    O.lib().orc_spmv_spe(g.ref, coef, jp, x, y)
    sim.spmv(capi.X, capi.VEC_B, op=capi.CFD_OP_SPE)
    assert np.array_equal(sim.download(capi.VEC_B), y)


@pytest.mark.parametrize("gname", list(GRIDS))
def test_velocity_phases(require_gpu, gname):
    grid = GRIDS[gname]()
    sim, p = make_sim(grid, quirk=1)
    sim.init_state()
    sim.DFIB_CylinderZ()
    sim.createPressureMatrix()
    g = O.Grid(grid)
    T0 = random_state(grid, 11)
    T1 = random_state(grid, 12)
    for c in range(3):
        sim.upload(capi.T0_U + c, T0[c])
        sim.upload(capi.T1_U + c, T1[c])
    # predictor
    sim.ConvectionDifussion()
    o1 = [t.copy() for t in T1]
    O.lib().orc_predictor(g.ref, 1.0 / p.Re, p.dt, *T0, *o1)
    for c in range(3):
        assert rel_l2(sim.download(capi.T1_U + c), o1[c]) < 1e-13
    # source add with the reference's index mapping
    rng = np.random.default_rng(5)
    src = 1e-3 * rng.standard_normal(3 * g.N)
    sim.upload(capi.SRC, src)
    for c in range(3):
        sim.upload(capi.T1_U + c, o1[c])
    sim.add_source()
    O.lib().orc_source_add(g.ref, 1, src, *o1)
    for c in range(3):
        assert np.array_equal(sim.download(capi.T1_U + c), o1[c])
    # RHS
    sim.createBMatrix()
    b = np.zeros(g.N)
    O.lib().orc_rhs(g.ref, p.dt, sim.download(capi.JP), *o1, b)
    assert np.array_equal(sim.download(capi.RHS), b)  # the same operations in the same order
    # DFIB update from a given pressure
    pres = rng.standard_normal(g.G)
    sim.upload(capi.P, pres)
    eta = sim.download(capi.ETA)
    T3 = random_state(grid, 13)
    for c in range(3):
        sim.upload(capi.T3_U + c, T3[c])
    f0 = rng.standard_normal(3 * g.N)
    sim.upload(capi.F, f0)
    sim.update_UandF()
    f = f0.copy()
    O.lib().orc_dfib_update(g.ref, p.dt, eta, pres, *o1, *T3, f)
    for c in range(3):
        assert np.array_equal(sim.download(capi.T3_U + c), T3[c])
    assert np.array_equal(sim.download(capi.F), f)
    # drag
    cd, cl = sim.drag()
    ocd, ocl = C.c_double(), C.c_double()
    O.lib().orc_drag(g.ref, f, C.byref(ocd), C.byref(ocl))
    assert abs(cd - ocd.value) <= 1e-12 * max(1.0, abs(ocd.value))
    assert abs(cl - ocl.value) <= 1e-12 * max(1.0, abs(ocl.value))
    # source accumulate
    sim.upload(capi.SRC, src)
    sim.accumulate_source()
    s2 = src.copy()
    O.lib().orc_source_accumulate(g.ref, p.dt, f, s2)
    assert np.array_equal(sim.download(capi.SRC), s2)
    # BC main + copy
    Dir, Num = bc_tables(p)
    sim.BC_staggered_main()
    T0o = [t.copy() for t in T0]
    O.lib().orc_bc_main(g.ref, C.byref(Dir), C.byref(Num), pd6(p), *T0o, pres)
    for c in range(3):
        assert np.array_equal(sim.download(capi.T0_U + c), T0o[c])
    assert np.array_equal(sim.download(capi.P), pres)
    sim.BC_staggered_copy()
    O.lib().orc_bc_copy(g.ref, *T0o, *o1)
    for c in range(3):
        assert np.array_equal(sim.download(capi.T1_U + c), o1[c])
    # max difference
    md = sim.CheckSteadyStatebyMaxVal()
    om = (C.c_double * 3)()
    O.lib().orc_maxdiff(g.ref, *T0o, *T3, C.byref(om))
    assert np.allclose(md, list(om), rtol=0, atol=0)


@pytest.mark.parametrize("gname", list(GRIDS))
@pytest.mark.parametrize("scheme", [capi.CFD_CONV_LUD, capi.CFD_CONV_UD])
def test_predictor_schemes_match_oracle(require_gpu, gname, scheme):
    """The LUD / UD convection schemes on the device against the oracle's,
    at the predictor's bar (FMA contraction of a different compiler)."""
    grid = GRIDS[gname]()
    sim, p = make_sim(grid, quirk=1)
    p.convection = scheme
    sim = Simulation(grid, p)
    sim.init_state()
    g = O.Grid(grid)
    T0 = random_state(grid, 21)
    T1 = random_state(grid, 22)
    for c in range(3):
        sim.upload(capi.T0_U + c, T0[c])
        sim.upload(capi.T1_U + c, T1[c])
    sim.ConvectionDifussion()
    o1 = [t.copy() for t in T1]
    O.lib().orc_predictor_ex(g.ref, scheme, 1.0 / p.Re, p.dt, *T0, *o1)
    for c in range(3):
        assert rel_l2(sim.download(capi.T1_U + c), o1[c]) < 1e-13


@pytest.mark.parametrize("scheme", [capi.CFD_CONV_LUD, capi.CFD_CONV_UD])
def test_steps_with_schemes_match_oracle(require_gpu, scheme):
    grid = GRIDS["sml"]()
    sim, p = make_sim(grid, quirk=0, pd=(0, 1, 0, 0, 0, 0), tol=1e-13, inner=3)
    p.convection = scheme
    sim = Simulation(grid, p)
    sim.setup(noise=1e-3, seed=3)
    osim = O.Sim(O.Grid(grid), p)
    osim.init_state(1e-3, 3)
    osim.setup()
    for k in range(3):
        st = sim.step()
        ost = osim.step()
        for c in range(3):
            e = rel_l2(sim.download(capi.T0_U + c), osim.field(capi.T0_U + c))
            assert e < 1e-10, (k, c, e)
        for q in range(st.n_passes):
            assert abs(st.cd[q] - ost.cd[q]) <= 1e-9 * max(1.0, abs(ost.cd[q]))


@pytest.mark.parametrize("gname", list(GRIDS))
def test_predictor_with_smagorinsky_matches_oracle(require_gpu, gname):
    grid = GRIDS[gname]()
    sim, p = make_sim(grid, quirk=1)
    p.les = 1
    sim = Simulation(grid, p)
    sim.init_state()
    g = O.Grid(grid)
    T0 = random_state(grid, 31)
    T1 = random_state(grid, 32)
    for c in range(3):
        sim.upload(capi.T0_U + c, T0[c])
        sim.upload(capi.T1_U + c, T1[c])
    sim.ConvectionDifussion()
    vis = np.zeros(g.N)
    O.lib().orc_smagorinsky(g.ref, *T0, vis)
    assert vis.max() > 1e-3 * (1.0 / p.Re)  # the model is active on these fields
    o1 = [t.copy() for t in T1]
    O.lib().orc_predictor_les(g.ref, capi.CFD_CONV_QUICK, 1.0 / p.Re, p.dt, vis.ctypes.data, *T0, *o1)
    for c in range(3):
        assert rel_l2(sim.download(capi.T1_U + c), o1[c]) < 1e-13


def stretched_grid(ncal):
    """Stretched x / y axes (uniform band, growth 1.05 outside), uniform z."""
    nx, ny, nz = ncal
    ax = axis_tanh(nx, 20.0, 6.0, 20.0 / nx * 0.5, 2.0, 1.05)
    ay = axis_tanh(ny, 10.0, 5.0, 10.0 / ny * 0.5, 2.0, 1.05)
    az = uniform_grid((nz, nz, nz), (2.5, 2.5, 2.5))
    return GridArrays(tuple(ncal), ax[0], ay[0], az.Z, ax[1], ay[1], az.Dz, ax[2], ay[2], az.Dzs, 20.0, 10.0, 2.5)


@pytest.mark.parametrize("ncal", [(20, 70, 4), (17, 33, 6), (37, 19, 16), (23, 11, 32), (35, 9, 64), (19, 13, 70)])
def test_tiled_quick_predictor_every_patch_shape(require_gpu, ncal):
    """The QUICK predictor runs on LDS patches of tj x tk (j, k) columns
    (tk = 4 ... 64 from the grid's depth: the extents here select each, 70
    cells take five 16-deep patches) marching over 16 i-planes. None of these
    extents fills its patches or plane ranges, the x / y axes are stretched,
    z is closed or periodic, the eddy viscosity is off or on: against the
    oracle at the predictor's bar."""
    grid = stretched_grid(ncal)
    g = O.Grid(grid)
    T0 = random_state(grid, 41)
    for periodic in (0, 1):
        if periodic and ncal[2] % 2:
            continue
        for les in (0, 1):
            sim, p = make_sim(grid, quirk=0)  # periodic z needs the source at its own cell
            p.periodic_z, p.les = periodic, les
            sim = Simulation(grid, p)
            sim.init_state()
            T1 = random_state(grid, 42)
            for c in range(3):
                sim.upload(capi.T0_U + c, T0[c])
                sim.upload(capi.T1_U + c, T1[c])
            sim.ConvectionDifussion()
            vis = np.zeros(g.N)
            if les:
                O.lib().orc_smagorinsky(g.ref, *T0, vis)
            o1 = [t.copy() for t in T1]
            O.lib().orc_predictor_bl(g.ref, C.byref(p), vis.ctypes.data if les else None, *T0, *o1)
            for c in range(3):
                out = sim.download(capi.T1_U + c)
                assert rel_l2(out, o1[c]) < 1e-13, (ncal, periodic, les, c)
                assert np.array_equal(out == T1[c], o1[c] == T1[c]), (ncal, periodic, les, c)  # the same cells written


def test_steps_with_smagorinsky_match_oracle(require_gpu):
    grid = GRIDS["sml"]()
    sim, p = make_sim(grid, quirk=0, pd=(0, 1, 0, 0, 0, 0), tol=1e-13, inner=3)
    p.les = 1
    sim = Simulation(grid, p)
    sim.setup(noise=1e-3, seed=3)
    osim = O.Sim(O.Grid(grid), p)
    osim.init_state(1e-3, 3)
    osim.setup()
    for k in range(3):
        st = sim.step()
        ost = osim.step()
        for c in range(3):
            e = rel_l2(sim.download(capi.T0_U + c), osim.field(capi.T0_U + c))
            assert e < 1e-10, (k, c, e)
        for q in range(st.n_passes):
            assert abs(st.cd[q] - ost.cd[q]) <= 1e-9 * max(1.0, abs(ost.cd[q]))


def oracle_csr(g):
    ptr, col, val, jp = np.zeros(g.N + 1, np.int32), np.zeros(7 * g.N, np.int32), np.zeros(7 * g.N), np.zeros(g.N)
    nnz = O.lib().orc_csr_assemble(g.ref, ptr, col, val, jp)
    return ptr, col[:nnz].copy(), val[:nnz].copy(), jp


@pytest.mark.parametrize("gname", list(GRIDS))
def test_csr_assembly_and_spmv_bit_exact(require_gpu, gname):
    """createPressureMatrixCSR and CSR_matrix::multiply_omp on the device:
    row pointers, columns, values, jp (which becomes gA.jp with op = CSR)
    and A x identical to the oracle's; the pressure-Dirichlet face of the
    table does not change the CSR builder's Neumann closure."""
    grid = GRIDS[gname]()
    sim, p = make_sim(grid, pd=(0, 1, 0, 0, 0, 0), op=capi.CFD_OP_CSR)
    sim.init_state()
    sim.createPressureMatrix()
    sim.createPressureMatrixCSR()
    g = O.Grid(grid)
    optr, ocol, oval, ojp = oracle_csr(g)
    ptr, col, val = sim.csr()
    assert np.array_equal(ptr, optr) and np.array_equal(col, ocol) and np.array_equal(val, oval)
    assert np.array_equal(sim.download(capi.JP), ojp)
    x = np.random.default_rng(5).standard_normal(g.N)
    sim.upload(capi.X, x)
    y = np.zeros(g.N)
    O.lib().orc_spmv_csr(g.N, optr, ocol, oval, x, y)
    sim.spmv(capi.X, capi.VEC_A, op=capi.CFD_OP_CSR)
    assert np.array_equal(sim.download(capi.VEC_A), y)


def test_native_grid_csr_against_reference_hashes(require_gpu):
    """The device CSR operator of the reference's native case hashes to the
    compiled reference's own dump, and so does its A x on the seeded x."""
    import hashlib
    import json
    from pathlib import Path
    from cfd_nssolver_forhpc_amd.solver import default_params
    fx = json.loads((Path(__file__).parent / "golden" / "native_fixtures.json").read_text())
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    p = default_params()
    p.op = capi.CFD_OP_CSR
    sim = Simulation(native_grid(), p)
    sim.init_state()
    sim.createPressureMatrixCSR()
    ptr, col, val = sim.csr()
    assert sha(ptr) == fx["setup"]["csr_ptr.bin"]
    assert sha(col) == fx["setup"]["csr_idx.bin"]
    assert sha(val) == fx["setup"]["csr_val.bin"]
    assert sha(sim.download(capi.JP)) == fx["setup"]["csr_jp.bin"]
    sim.upload(capi.X, O.noise(sim.N, fx["seed"], 5, 0.0, 1.0))
    sim.spmv(capi.X, capi.VEC_A, op=capi.CFD_OP_CSR)
    assert sha(sim.download(capi.VEC_A)) == fx["synth"]["syn_spmv_csr.bin"]["sha256"]


@pytest.mark.parametrize("op", [capi.CFD_OP_ELL, capi.CFD_OP_SPE, capi.CFD_OP_CSR])
def test_bicgstab_matches_oracle(require_gpu, op):
    grid = GRIDS["sml"]()
    sim, p = make_sim(grid, pd=(0, 1, 0, 0, 0, 0), op=op)
    sim.init_state()
    sim.createPressureMatrix()
    sim.setupPressure()
    if op == capi.CFD_OP_CSR:
        sim.createPressureMatrixCSR()
    g = O.Grid(grid)
    rng = np.random.default_rng(9)
    b = rng.standard_normal(g.N)
    if op == capi.CFD_OP_CSR:
        # Neumann on every face: a consistent b = A x_t keeps the singular system solvable
        cptr, ccol, cval, _ = oracle_csr(g)
        xt = b - b.mean()
        O.lib().orc_spmv_csr(g.N, cptr, ccol, cval, xt, b)
    # two solves that each stop below a relative residual tol may differ by
    # up to kappa * tol in x; kappa is ~1e3 here (~1.5e3 for the all-Neumann
    # CSR system), so 1e-13 keeps them within ~1e-10 of each other whatever
    # order either side sums its dot products in (1e-12 allows 2.5e-10)
    tol = 1e-13
    sim.upload(capi.RHS, b)
    sim.upload(capi.X, np.zeros(g.N))
    it, rel = sim.bicgstab(tol=tol, maxit=5000)
    x = sim.download(capi.X)
    col, val = sim.ell()
    coef = np.zeros(2 * (g.nxc + g.nyc + g.nzc))
    jp = np.zeros(g.N)
    O.lib().orc_spe_setup(g.ref, pd6(p), coef, jp)
    if op == capi.CFD_OP_CSR:
        s = O.lib().orc_bicg_create_csr(g.N, cptr, ccol, cval)
    else:
        s = O.lib().orc_bicg_create(g.N, op, col, val, g.ref, coef.ctypes.data, jp.ctypes.data)
    O.lib().orc_bicg_set_tolerance(s, tol)
    ox = np.zeros(g.N)
    oit = C.c_int()
    orel = O.lib().orc_bicg_solve(s, b, ox, C.byref(oit))
    O.lib().orc_bicg_destroy(s)
    assert rel < tol and orel < tol
    assert abs(it - oit.value) <= max(3, oit.value // 10)
    if op == capi.CFD_OP_CSR:
        # the singular system's constant mode is fixed only by rounding
        x, ox = x - x.mean(), ox - ox.mean()
    assert rel_l2(x, ox) < 1e-10


def test_rbsor_matches_oracle(require_gpu):
    grid = GRIDS["uniform"]()
    sim, p = make_sim(grid, pd=(0, 1, 0, 0, 0, 0))
    sim.init_state()
    sim.createPressureMatrix()
    g = O.Grid(grid)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(g.N)
    sim.upload(capi.RHS, b)
    sim.upload(capi.X, np.zeros(g.N))
    it, rel = sim.rbsor(omega=1.8, tol=1e-300, maxit=40, check_every=10)
    col, val = sim.ell()
    ox = np.zeros(g.N)
    oit = C.c_int()
    orel = O.lib().orc_rbsor(g.N, g.ref, col, val, b, ox, 1.8, 1e-300, 40, 10, C.byref(oit))
    assert it == oit.value == 40
    assert np.array_equal(sim.download(capi.X), ox)
    assert abs(rel - orel) <= 1e-12 * orel


def test_steps_csr_match_oracle(require_gpu):
    """P_SOLVER_BICG_CSR timesteps.  The CSR operator is Neumann on every
    face, so the case is the reference's closed cavity (lid-driven, from
    rest, source added at its own cell): a singular system whose right-hand
    side stays consistent.  Its constant mode is fixed only by rounding, so
    X and p are compared without their means, at 1e-9."""
    grid = GRIDS["sml"]()
    sim, p = make_sim(grid, quirk=0, op=capi.CFD_OP_CSR, tol=1e-12, inner=3)
    p.set_bc(capi.CAVITY_A_DIR, capi.CAVITY_A_NUM)
    p.u_init[0] = 0.0
    sim = Simulation(grid, p)
    sim.setup(noise=1e-3, seed=3)
    osim = O.Sim(O.Grid(grid), p)
    osim.init_state(1e-3, 3)
    osim.setup()
    assert np.array_equal(sim.download(capi.JP), osim.field(capi.JP))
    demean = lambda a: a - a.mean()
    for k in range(4):
        st = sim.step()
        ost = osim.step()
        for c in range(3):
            e = rel_l2(sim.download(capi.T0_U + c), osim.field(capi.T0_U + c))
            assert e < 1e-10, (k, c, e)
        assert rel_l2(demean(sim.download(capi.X)), demean(osim.field(capi.X))) < 1e-9
        assert rel_l2(sim.download(capi.F), osim.field(capi.F)) < 1e-9
        for q in range(st.n_passes):
            assert (st.iters[q] < 0) == (ost.iters[q] < 0)
            assert st.relres[q] < 1e-12 or st.iters[q] < 0
            assert abs(st.cd[q] - ost.cd[q]) <= 1e-9 * max(1.0, abs(ost.cd[q]))


@pytest.mark.parametrize("quirk,inner,op", [(1, 3, capi.CFD_OP_ELL), (0, 3, capi.CFD_OP_ELL),
                                            (1, 0, capi.CFD_OP_ELL), (1, 3, capi.CFD_OP_SPE)])
def test_steps_match_oracle(require_gpu, quirk, inner, op):
    grid = GRIDS["sml"]()
    pd = (0, 1, 0, 0, 0, 0)
    sim, p = make_sim(grid, quirk=quirk, pd=pd, op=op, tol=1e-13, inner=inner)
    sim.setup(noise=1e-3, seed=3)
    osim = O.Sim(O.Grid(grid), p)
    osim.init_state(1e-3, 3)
    osim.setup()
    for c in range(3):
        assert np.array_equal(sim.download(capi.T0_U + c), osim.field(capi.T0_U + c))
    for k in range(4):
        st = sim.step()
        ost = osim.step()
        for c in range(3):
            e = rel_l2(sim.download(capi.T0_U + c), osim.field(capi.T0_U + c))
            assert e < 1e-10, (k, c, e)
        assert rel_l2(sim.download(capi.X), osim.field(capi.X)) < 1e-10
        assert rel_l2(sim.download(capi.P), osim.field(capi.P)) < 1e-10
        assert rel_l2(sim.download(capi.F), osim.field(capi.F)) < 1e-9
        for q in range(st.n_passes):
            assert abs(st.cd[q] - ost.cd[q]) <= 1e-9 * max(1.0, abs(ost.cd[q]))
        for c in range(3):
            assert abs(st.maxdiff[c] - ost.maxdiff[c]) <= 1e-10 * max(1e-3, abs(ost.maxdiff[c]))
