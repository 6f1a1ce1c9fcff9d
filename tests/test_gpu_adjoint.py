"""The GPU discrete adjoint (an addition with no reference counterpart).

Pins: the objective and the gradient against the oracle's adjoint (which
solves the transposed pressure system on the explicit transpose of the
assembled ELL operator and records every step instead of recomputing from
checkpoints), with the boundary table and with the workloads' convective
outflow and periodic z, the gradient against central finite differences of the GPU
forward run, independence from the checkpoint interval, the state left
after the sweep, the torch.autograd.Function, and bit-identity of the
forward step at the parameters' reference values.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle_lib as O
from cfd_nssolver_forhpc_amd import capi, configs
from cfd_nssolver_forhpc_amd.adjoint import DragObjective, DragProblem, smoothed_inflow
from cfd_nssolver_forhpc_amd.solver import Simulation

pytestmark = pytest.mark.gpu

STEPS, AVG_FROM = 10, 5


def case(tol=1e-13):
    w = configs.small_cylinder(ncal=(24, 12, 4), lengths=(6.0, 3.0, 1.0), tol=tol)
    w.params.maxit = 20000
    return w, smoothed_inflow(12)


def gpu_problem(w, every):
    sim = Simulation(w.grid, w.params)
    sim.init_state(1e-3, 1)
    sim.DFIB_CylinderZ()
    sim.createPressureMatrix()
    return DragProblem(sim, STEPS, AVG_FROM, every, noise=1e-3, seed=1)


def oracle_gradient(w, uin, gain):
    g = O.Grid(w.grid)
    s = O.Sim(g, w.params)
    s.set_inflow(uin)
    s.set_gain(gain)
    s.init_state(1e-3, 1)
    s.setup()
    J = s.objective(STEPS, AVG_FROM)
    g_uin, g_gain, _ = s.gradient()
    return J, g_uin, g_gain


def rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def test_objective_and_gradient_match_the_oracle(require_gpu):
    w, uin = case()
    pr = gpu_problem(w, every=3)
    J = pr.run(uin, 0.9)
    g_uin, g_gain, st = pr.sim.adjoint_backward()
    Jo, go_uin, go_gain = oracle_gradient(w, uin, 0.9)
    print("J", J, Jo, "gain grad", g_gain, go_gain, "stats", st)
    assert rel(J, Jo) < 1e-10
    assert np.linalg.norm(g_uin - go_uin) <= 1e-8 * np.linalg.norm(go_uin)
    assert rel(g_gain, go_gain) < 1e-8
    assert st["checkpoints"] == 4 and st["recomputed_steps"] == STEPS and st["adjoint_solves"] > 0


def test_gradient_matches_central_differences(require_gpu):
    """About gain 0.9: at gain 1 exactly the fully immersed cells hold exact
    zeros, whose QUICK upwind test flips with the sign of (gain - 1), so the
    discrete objective has a small jump there that central differences
    straddle.

    The bar is the 1e-5 relative agreement the adjoint is specified to.  The
    central difference itself carries noise: the forward runs' pressure
    solves stop at a relative residual of 1e-13, which leaves ~1e-12 in J,
    and (J+ - J-) / 2h with h = 1e-6 turns that into ~1e-6 in the
    difference quotient (0.64 here), so a tighter bar would test the solver's
    stopping point, not the gradient.  The gradient itself is pinned to the
    oracle's adjoint at 1e-8 above."""
    w, uin = case()
    pr = gpu_problem(w, every=4)
    base = 0.9
    pr.run(uin, base)
    g_uin, g_gain, _ = pr.sim.adjoint_backward()
    rng = np.random.default_rng(21)
    h = 1e-6
    for _ in range(3):
        d = rng.standard_normal(uin.size + 1)
        d /= np.linalg.norm(d)
        Jp = pr.run(uin + h * d[:-1], base + h * d[-1])
        Jm = pr.run(uin - h * d[:-1], base - h * d[-1])
        fd = (Jp - Jm) / (2 * h)
        ad = float(np.dot(g_uin, d[:-1]) + g_gain * d[-1])
        print("fd", fd, "adjoint", ad, "rel", rel(fd, ad))
        assert rel(fd, ad) < 1e-5


def outflow_periodic_case(tol=1e-13):
    """The small case with the workloads' boundaries: convective outflow and
    periodic z."""
    w, uin = case(tol)
    configs.outflow_periodic_bc(w.params)
    return w, uin


def test_outflow_periodic_bc_objective_and_gradient_match_the_oracle(require_gpu):
    """With the workloads' boundaries the oracle's gradient (explicit
    transposes, the transposed solve on the explicit transpose of the
    wrapped ELL operator, every step recorded) pins the GPU's (the symmetric
    reformulation of the transposed solve, checkpoints and recomputation)."""
    w, uin = outflow_periodic_case()
    pr = gpu_problem(w, every=3)
    J = pr.run(uin, 0.9)
    g_uin, g_gain, st = pr.sim.adjoint_backward()
    Jo, go_uin, go_gain = oracle_gradient(w, uin, 0.9)
    print("J", J, Jo, "gain grad", g_gain, go_gain, "|g_uin - oracle| / |oracle|",
          np.linalg.norm(g_uin - go_uin) / np.linalg.norm(go_uin))
    assert rel(J, Jo) < 1e-10
    assert np.linalg.norm(g_uin - go_uin) <= 1e-8 * np.linalg.norm(go_uin)
    assert rel(g_gain, go_gain) < 1e-8


def test_outflow_periodic_bc_gradient_matches_central_differences(require_gpu):
    """The transposed step with the convective outflow (its implicit upwind
    update from the ghosts' old values, the projected outflow face) and
    periodic z (ghost planes folded into their images, the wrapped pressure
    coupling, the seam face of w), against central differences of the GPU
    forward run, as for the table conditions above."""
    w, uin = outflow_periodic_case()
    pr = gpu_problem(w, every=4)
    base = 0.9
    pr.run(uin, base)
    g_uin, g_gain, st = pr.sim.adjoint_backward()
    assert st["recomputed_steps"] == STEPS and st["adjoint_solves"] > 0
    rng = np.random.default_rng(23)
    h = 1e-6
    dirs = [np.concatenate([1.0 + 0.5 * rng.standard_normal(uin.size), [0.0]])]  # u_in only
    dirs += [rng.standard_normal(uin.size + 1) for _ in range(2)]
    for d in dirs:
        d = d / np.linalg.norm(d)
        Jp = pr.run(uin + h * d[:-1], base + h * d[-1])
        Jm = pr.run(uin - h * d[:-1], base - h * d[-1])
        fd = (Jp - Jm) / (2 * h)
        ad = float(np.dot(g_uin, d[:-1]) + g_gain * d[-1])
        print("outflow/periodic: fd", fd, "adjoint", ad, "rel", rel(fd, ad))
        assert rel(fd, ad) < 1e-5


def test_outflow_periodic_bc_checkpoint_interval_does_not_change_the_gradient(require_gpu):
    w, uin = outflow_periodic_case(tol=1e-10)
    res = []
    for every in (1, 4, STEPS):
        pr = gpu_problem(w, every)
        J = pr.run(uin, 1.0)
        g_uin, g_gain, _ = pr.sim.adjoint_backward()
        res.append((J, g_uin, g_gain))
    for J, g_uin, g_gain in res[1:]:
        assert J == res[0][0]
        assert np.array_equal(g_uin, res[0][1]) and g_gain == res[0][2]


def test_checkpoint_interval_does_not_change_the_gradient(require_gpu):
    w, uin = case(tol=1e-10)
    res = []
    for every in (1, 4, STEPS):
        pr = gpu_problem(w, every)
        J = pr.run(uin, 1.0)
        g_uin, g_gain, st = pr.sim.adjoint_backward()
        res.append((J, g_uin, g_gain))
        assert st["checkpoints"] == -(-STEPS // every)
    for J, g_uin, g_gain in res[1:]:
        assert J == res[0][0]
        assert np.array_equal(g_uin, res[0][1]) and g_gain == res[0][2]


def test_sweep_leaves_the_final_state(require_gpu):
    w, uin = case(tol=1e-10)
    pr = gpu_problem(w, every=3)
    pr.run(uin, 1.0)
    before = [pr.sim.download(f) for f in (capi.T0_U, capi.T0_V, capi.T0_W, capi.X)]
    loop = pr.sim.loop
    pr.sim.adjoint_backward()
    after = [pr.sim.download(f) for f in (capi.T0_U, capi.T0_V, capi.T0_W, capi.X)]
    assert pr.sim.loop == loop
    assert all(np.array_equal(a, b) for a, b in zip(before, after))
    # and the run continues as if the sweep had not happened
    st = pr.sim.step()
    assert st.loop == loop


def test_autograd_function(require_gpu):
    w, uin = case(tol=1e-10)
    pr = gpu_problem(w, every=5)
    u = torch.tensor(uin, dtype=torch.float64, requires_grad=True)
    gn = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    J = DragObjective.apply(u, gn, pr)
    (2.0 * J).backward()
    pr2 = gpu_problem(w, every=5)
    J2 = pr2.run(uin, 1.0)
    g_uin, g_gain, _ = pr2.sim.adjoint_backward()
    assert float(J) == J2
    assert np.array_equal(u.grad.numpy(), 2.0 * g_uin) and float(gn.grad) == 2.0 * g_gain
    assert pr.last_stats["steps"] == STEPS


def test_stale_objective_run_is_refused(require_gpu):
    """backward differentiates the run its forward recorded: after a second
    forward, or a changed gain / state, the first graph node raises and the
    C-ABI refuses the sweep instead of mixing trajectories."""
    w, uin = case(tol=1e-10)
    pr = gpu_problem(w, every=5)
    u1 = torch.tensor(uin, dtype=torch.float64, requires_grad=True)
    g1 = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    J1 = DragObjective.apply(u1, g1, pr)
    u2 = torch.tensor(uin * 1.01, dtype=torch.float64, requires_grad=True)
    g2 = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    J2 = DragObjective.apply(u2, g2, pr)
    with pytest.raises(RuntimeError, match="another objective"):
        J1.backward()
    J2.backward()  # the latest run is still differentiable
    assert u2.grad is not None and np.all(np.isfinite(u2.grad.numpy()))
    pr.run(uin, 1.0)
    assert pr.sim.adjoint_generation() > 0
    pr.sim.set_forcing_gain(1.001)
    assert pr.sim.adjoint_generation() == 0
    with pytest.raises(capi.CfdError) as e:
        pr.sim.adjoint_backward()
    assert e.value.status == capi.CFD_ERR_STATE
    pr.run(uin, 1.0)
    pr.sim.step()  # the state moved on
    with pytest.raises(capi.CfdError):
        pr.sim.adjoint_backward()


def test_reference_parameter_values_keep_the_step_bit_identical(require_gpu):
    w, _ = case(tol=1e-10)
    runs = []
    for set_params in (False, True):
        sim = Simulation(w.grid, w.params)
        if set_params:
            sim.set_inflow_profile(np.full(w.grid.ncal[1], w.params.bc_dir[0][0]))
            sim.set_forcing_gain(1.0)
        sim.init_state(1e-3, 1)
        sim.DFIB_CylinderZ()
        sim.createPressureMatrix()
        for _ in range(3):
            sim.step()
        runs.append([sim.download(f) for f in (capi.T0_U, capi.T0_V, capi.T0_W, capi.F, capi.X)])
    assert all(np.array_equal(a, b) for a, b in zip(*runs))


def test_bad_adjoint_arguments_are_errors(require_gpu):
    w, uin = case(tol=1e-10)
    sim = Simulation(w.grid, w.params)
    with pytest.raises(capi.CfdInvalidArgument):
        sim.set_inflow_profile(np.ones(5))
    with pytest.raises(capi.CfdError):
        sim.adjoint_backward()  # no objective run yet
    sim.init_state(1e-3, 1)
    sim.createPressureMatrix()
    with pytest.raises(capi.CfdInvalidArgument):
        sim.adjoint_forward(5, 5, 1)
    # the transposed predictor is QUICK's
    p2 = w.params.copy()
    p2.convection = capi.CFD_CONV_LUD
    sim2 = Simulation(w.grid, p2)
    sim2.init_state(1e-3, 1)
    sim2.createPressureMatrix()
    with pytest.raises(capi.CfdInvalidArgument):
        sim2.adjoint_forward(5, 2, 1)
    p3 = w.params.copy()
    p3.les = 1
    sim3 = Simulation(w.grid, p3)
    sim3.init_state(1e-3, 1)
    sim3.createPressureMatrix()
    with pytest.raises(capi.CfdInvalidArgument):
        sim3.adjoint_forward(5, 2, 1)
    # periodic z: the transposed solve needs the wrapped coupling symmetric,
    # i.e. a uniform z axis
    from cfd_nssolver_forhpc_amd.capi import GridArrays
    from cfd_nssolver_forhpc_amd.solver import axis_tanh, axis_uniform
    ax, ay, az = axis_uniform(24, 6.0), axis_uniform(12, 3.0), axis_tanh(8, 2.0, 1.0, 0.2, 0.4, 1.1)
    gz = GridArrays((24, 12, 8), ax[0], ay[0], az[0], ax[1], ay[1], az[1], ax[2], ay[2], az[2], 6.0, 3.0, 2.0)
    p4 = configs.outflow_periodic_bc(w.params.copy())
    sim4 = Simulation(gz, p4)
    sim4.init_state(1e-3, 1)
    sim4.createPressureMatrix()
    with pytest.raises(capi.CfdInvalidArgument):
        sim4.adjoint_forward(5, 2, 1)


@pytest.mark.parametrize("mode", ["fp64", "pc32"])
def test_multigrid_forward_and_adjoint_solves_match_the_oracle(require_gpu, mode):
    """The same objective and gradient with every forward and transposed
    pressure solve preconditioned by the multigrid V-cycle (to the same
    tolerance)."""
    w, uin = case()
    w.params.precond = capi.CFD_PRECOND_MG
    if mode == "pc32":
        w.params.precision = capi.CFD_MIXED
        w.params.mixed_mode = capi.CFD_MIXED_PC32
    pr = gpu_problem(w, every=4)
    J = pr.run(uin, 1.0)
    g_uin, g_gain, st = pr.sim.adjoint_backward()
    w0, _ = case()
    Jo, go_uin, go_gain = oracle_gradient(w0, uin, 1.0)
    print(mode, "J", J, Jo, "adjoint iterations", st["adjoint_iters"], "solves", st["adjoint_solves"])
    assert rel(J, Jo) < 1e-9
    assert np.linalg.norm(g_uin - go_uin) <= 1e-7 * np.linalg.norm(go_uin)
    assert rel(g_gain, go_gain) < 1e-7
