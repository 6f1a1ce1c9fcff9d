"""Per-phase device timers of the timestep (cfd_phase_timing /
cfd_phase_times), the reference's clockstruct stopwatches
(inc/profiling/STL_clock.hpp:35-47) around the same parts of its loop
(RUN_CPU_PC_METHOD.hpp:296-607): they account for the steps' device time,
see the solve only where there is one, and leave the step's results as they
are."""
from __future__ import annotations

import time

import numpy as np
import pytest

from cfd_nssolver_forhpc_amd import capi, configs
from cfd_nssolver_forhpc_amd.solver import Simulation

pytestmark = pytest.mark.gpu


def case():
    w = configs.small_cylinder(ncal=(64, 32, 16), lengths=(8.0, 4.0, 2.0), quirk=0, tol=1e-10)
    w.params.maxit = 20000
    return w


def test_phase_times_account_for_the_steps(require_gpu):
    w = case()
    sim = Simulation(w.grid, w.params)
    sim.setup(noise=w.noise, seed=w.seed)
    sim.phase_timing(True)
    sim.step()  # loop 1: no pressure solve
    first = sim.phase_times()
    assert first["steps"] == 1
    t0 = time.perf_counter()
    for _ in range(3):
        sim.step()
    wall = 1e3 * (time.perf_counter() - t0)
    ph = sim.phase_times()
    ms = {k: ph["ms"][k] - first["ms"][k] for k in sim.PHASES}
    total = sum(ms.values())
    print("phases (ms, 3 steps):", ms, "wall", wall)
    assert ph["steps"] == 4
    assert all(v >= 0.0 for v in ms.values())
    assert 0.3 * wall <= total <= 1.05 * wall
    # the solve dominates a step with solves; without one there is almost no pressure time
    assert ms["pressure"] > 0.5 * total
    assert first["ms"]["pressure"] < 0.05 * ms["pressure"] / 3
    assert ms["convection"] > 0.0 and ms["update"] > 0.0 and ms["bc"] > 0.0
    # switching off clears the totals and stops them
    sim.phase_timing(False)
    sim.step()
    assert sim.phase_times() == {"steps": 0, "ms": {k: 0.0 for k in sim.PHASES}}
    sim.close()


def test_timing_leaves_the_step_unchanged(require_gpu):
    w = case()
    runs = []
    for timed in (False, True):
        sim = Simulation(w.grid, w.params)
        sim.setup(noise=w.noise, seed=w.seed)
        sim.phase_timing(timed)
        stats = [sim.step() for _ in range(3)]
        runs.append(([list(s.iters[:s.n_passes]) for s in stats],
                     [sim.download(f) for f in (capi.T0_U, capi.T0_V, capi.T0_W, capi.X)]))
        sim.close()
    assert runs[0][0] == runs[1][0]
    assert all(np.array_equal(a, b) for a, b in zip(runs[0][1], runs[1][1]))
