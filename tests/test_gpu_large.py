"""The largest workload at its full size: 1024x512x128 cells (67.1M), x/y
stretched by tanh clustering (growth 1.05) around a D = 40-cell cylinder,
Re 300, dt 2.5e-4 (configs.cylinder_1024x512x128).  The oracle does not finish at this
size, so the checks are size-independent properties:

- every multigrid-preconditioned pressure pass reaches the FP64 relative
  residual 1e-10: FP64 throughout, FP64 Krylov around the FP32 V-cycle, and
  FP64 defect correction around FP32 inner solves to 1e-5;
- the mixed run's fields stay within what two such solves can differ by;
- the discrete adjoint of the drag objective, with checkpoints and
  recomputation, agrees with a central difference of the forward run.
"""
from __future__ import annotations

import sys
import time

import numpy as np
import pytest

from cfd_nssolver_forhpc_amd import capi, configs
from cfd_nssolver_forhpc_amd.adjoint import DragProblem, smoothed_inflow
from cfd_nssolver_forhpc_amd.solver import Simulation

pytestmark = pytest.mark.gpu

FIELDS = (("u", capi.T0_U), ("v", capi.T0_V), ("w", capi.T0_W), ("p", capi.X))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def large(precision, mode):
    w = configs.cylinder_1024x512x128(precision)
    w.params.precond = capi.CFD_PRECOND_MG
    w.params.mixed_mode = mode
    w.params.tol_inner = 1e-5
    w.params.maxit = 200  # a multigrid solve takes ~10 iterations; no run-away at this size
    return w


def test_large_jacobi_bicgstab_residual_is_the_true_one(require_gpu):
    """The reference's Jacobi-scaled BiCGSTAB on the 67M-row ELL operator,
    where every kernel of the iteration runs its grid-stride loop (more rows
    than its grid's threads): after 30 iterations from x = 0 the solver's
    recursive relative residual must equal ||b - A x|| / ||b|| recomputed
    from the returned x (cfd_spmv), which holds only if every row's SpMV and
    update ran.  A skipped row block leaves a gap of order its share of
    ||b||; rounding leaves ~1e-13."""
    w = large(capi.CFD_FP64, capi.CFD_MIXED_RESTARTED)
    w.params.precond = capi.CFD_PRECOND_JACOBI
    w.params.op = capi.CFD_OP_ELL
    say(f"[large] jacobi bicgstab: setup {w.name}")
    sim = Simulation(w.grid, w.params)
    sim.init_state()
    sim.createPressureMatrix()
    assert sim.N == 1024 * 512 * 128
    b = np.random.default_rng(5).standard_normal(sim.N)
    sim.upload(capi.RHS, b)
    sim.upload(capi.X, np.zeros(sim.N))
    it, rel = sim.bicgstab(tol=1e-300, maxit=30)
    sim.spmv(capi.X, capi.VEC_A)
    ax = sim.download(capi.VEC_A)
    true = float(np.linalg.norm(b - ax) / np.linalg.norm(b))
    say(f"[large] jacobi bicgstab: {it} iterations, relres {rel:.6e}, true {true:.6e}")
    sim.close()
    assert 25 <= it <= 31
    assert 0.0 < rel < 0.9
    assert abs(rel - true) <= 1e-9 * true


def test_large_stretched_mixed_against_fp64(require_gpu):
    """Three steps of the 67M-cell case in FP64 and in the two mixed schemes
    (the first step has no pressure solve, as in the reference).  Each pass stops at a
    relative residual below 1e-10, so the two pressure fields may differ by
    up to kappa * 1e-10 in the modes the residual weighs least; the bars (velocity 1e-9, pressure 1e-6 relative L2) sit
    far below what a wrong operator, a lost FP64 correction or a stale
    preconditioner produce (1e-4 and more)."""
    fields = {}
    for tag, prec, mode in (("fp64_mg", capi.CFD_FP64, capi.CFD_MIXED_RESTARTED),
                            ("mixed_mg", capi.CFD_MIXED, capi.CFD_MIXED_PC32),
                            ("mixed_mg_dc", capi.CFD_MIXED, capi.CFD_MIXED_RESTARTED)):
        w = large(prec, mode)
        assert tuple(w.grid.ncal) == (1024, 512, 128)
        say(f"[large] {tag}: setup {w.name}")
        sim = Simulation(w.grid, w.params)
        sim.setup(noise=w.noise, seed=w.seed)
        for k in range(3):
            t0 = time.perf_counter()
            st = sim.step()
            ms = 1e3 * (time.perf_counter() - t0)
            rel = list(st.relres[:st.n_passes])
            its = list(st.iters[:st.n_passes])
            say(f"[large] {tag} step {k + 1} (loop {st.loop}): iters {its} inner {list(st.inner_iters[:st.n_passes])} "
                f"relres {max(rel):.2e} {ms:.0f} ms")
            if st.loop == 1:  # the reference's first loop has no pressure solve
                assert its == [-1] * st.n_passes
                continue
            assert all(0 < it < w.params.maxit for it in its)
            assert max(rel) < 1e-10
        fields[tag] = {name: sim.download(fid) for name, fid in FIELDS}
        sim.close()
    for tag in ("mixed_mg", "mixed_mg_dc"):
        for name, _ in FIELDS:
            x, y = fields[tag][name], fields["fp64_mg"][name]
            assert np.all(np.isfinite(x))
            err = float(np.linalg.norm(x - y) / np.linalg.norm(y))
            say(f"[large] {tag} vs fp64_mg {name}: rel L2 {err:.3e}")
            assert err < (1e-6 if name == "p" else 1e-9)


def test_large_adjoint_against_central_difference(require_gpu):
    """Discrete adjoint of the mean drag over steps [2, 4) at the full size,
    checkpoints every 2 steps (recomputation in between), against a central
    difference in one random direction of (u_in, gain) about gain 1.001 (at
    gain 1 exactly the immersed cells' QUICK upwind test meets exact zeros
    and the objective has a jump).  The step h = 1e-5 keeps the difference
    quotient's noise, ~1e-12 in J from solves stopped at 1e-12 divided by
    2h, near 1e-7 relative, while its truncation error is O(h^2) ~ 1e-10;
    the bar is the adjoint's 1e-5."""
    w = large(capi.CFD_FP64, capi.CFD_MIXED_RESTARTED)
    w.params.tol = 1e-12
    say(f"[large] adjoint: setup {w.name}")
    sim = Simulation(w.grid, w.params)
    sim.init_state(w.noise, w.seed)
    sim.DFIB_CylinderZ()
    sim.createPressureMatrix()
    pr = DragProblem(sim, 4, 2, 2, noise=w.noise, seed=w.seed)
    uin = smoothed_inflow(w.grid.ncal[1])
    gain, h = 1.001, 1e-5
    J = pr.run(uin, gain)
    say(f"[large] adjoint: J {J:.12f}")
    g_uin, g_gain, st = sim.adjoint_backward()
    say(f"[large] adjoint: |g_uin| {np.linalg.norm(g_uin):.6e} g_gain {g_gain:.6e} stats {st}")
    assert np.all(np.isfinite(g_uin)) and np.isfinite(g_gain)
    assert st["recomputed_steps"] == 4 and st["adjoint_solves"] > 0
    d = np.random.default_rng(5).standard_normal(uin.size + 1)
    d /= np.linalg.norm(d)
    Jp = pr.run(uin + h * d[:-1], gain + h * d[-1])
    Jm = pr.run(uin - h * d[:-1], gain - h * d[-1])
    fd = (Jp - Jm) / (2 * h)
    ad = float(np.dot(g_uin, d[:-1]) + g_gain * d[-1])
    rel = abs(fd - ad) / max(abs(fd), abs(ad))
    say(f"[large] adjoint: fd {fd:.9e} adjoint {ad:.9e} rel {rel:.2e}")
    sim.close()
    assert rel < 1e-5
