"""Exact checkpoint / resume (cfd_checkpoint_save / cfd_checkpoint_load; no
reference counterpart: the reference's only restart reads its float Plot3D
file back, and that block is commented out at RUN_CPU_PC_METHOD.hpp:98-111).

A run that is saved after a few steps and resumed in a fresh context
continues bit for bit like the run that never stopped: the same iteration
counts, residuals, drag and max-differences per pass, the same fields and
solver counters.  The
resumed part crosses the tenth solve, where the reference's restart rule
starts to apply, so the solve count travels too.  A file of another grid, a
truncated, corrupted or foreign file is rejected and leaves the context as
it was.
"""
from __future__ import annotations

import numpy as np
import pytest

from cfd_nssolver_forhpc_amd import capi, configs
from cfd_nssolver_forhpc_amd.solver import Simulation

pytestmark = pytest.mark.gpu

FIELDS = (capi.T0_U, capi.T0_V, capi.T0_W, capi.T3_U, capi.P, capi.X)
BEFORE, AFTER = 3, 3  # steps before the checkpoint (loop 1 has no solve) and after it


def workload(kind):
    w = configs.small_cylinder(ncal=(32, 16, 8), lengths=(8.0, 4.0, 2.0), quirk=0, tol=1e-10)
    p = w.params
    p.maxit = 20000
    if kind == "spe":
        p.op = capi.CFD_OP_SPE
    elif kind == "mg":
        p.precond = capi.CFD_PRECOND_MG
    elif kind == "mixed":
        p.precision = capi.CFD_MIXED
    return w


def stats_tuple(st):
    n = st.n_passes
    return (st.loop, n, list(st.iters[:n]), list(st.relres[:n]), list(st.cd[:n]), list(st.cl[:n]),
            list(st.maxdiff))


@pytest.mark.parametrize("kind", ["ell", "spe", "mg", "mixed"])
def test_resume_continues_bit_for_bit(require_gpu, tmp_path, kind):
    w = workload(kind)
    path = tmp_path / "state.ckpt"
    a = Simulation(w.grid, w.params)
    a.setup(noise=w.noise, seed=w.seed)
    for _ in range(BEFORE):
        a.step()
    a.save_checkpoint(path)
    ref = [stats_tuple(a.step()) for _ in range(AFTER)]
    ref_fields = [a.download(f) for f in FIELDS]
    ref_counters = a.solver_counters()
    loop_a = a.loop
    a.close()

    b = Simulation(w.grid, w.params)
    b.setup(noise=0.0, seed=99)  # a different initial state, replaced by the checkpoint
    b.load_checkpoint(path)
    assert b.loop == BEFORE + 1
    got = [stats_tuple(b.step()) for _ in range(AFTER)]
    assert got == ref
    for fid, x in zip(FIELDS, ref_fields):
        assert np.array_equal(b.download(fid), x), fid
    assert b.loop == loop_a
    # bicgstabRe2's counters continue too (timestep_, numRestarts_, pResets_, totalNumIter_)
    assert b.solver_counters() == ref_counters, (b.solver_counters(), ref_counters)
    b.close()


def test_bad_checkpoints_are_rejected(require_gpu, tmp_path):
    w = workload("ell")
    sim = Simulation(w.grid, w.params)
    sim.setup(noise=w.noise, seed=w.seed)
    sim.step()
    sim.step()
    good = tmp_path / "good.ckpt"
    sim.save_checkpoint(good)
    data = good.read_bytes()
    before = [sim.download(f) for f in FIELDS], sim.loop

    # another grid: same extents, stretched differently
    w2 = configs.small_cylinder(ncal=(32, 16, 8), lengths=(8.0, 4.0, 2.5), quirk=0, tol=1e-10)
    other = Simulation(w2.grid, w2.params)
    other.setup()
    other.save_checkpoint(tmp_path / "other.ckpt")
    other.close()
    with pytest.raises(capi.CfdInvalidArgument):
        sim.load_checkpoint(tmp_path / "other.ckpt")

    bad = {
        "truncated": data[: len(data) // 2],
        "flipped": data[:-100] + bytes([data[-100] ^ 0x01]) + data[-99:],
        "trailing": data + b"\0",
        "foreign": b"not a checkpoint" * 8,
    }
    for name, blob in bad.items():
        p = tmp_path / f"{name}.ckpt"
        p.write_bytes(blob)
        with pytest.raises(capi.CfdError) as e:
            sim.load_checkpoint(p)
        assert not isinstance(e.value, capi.CfdInvalidArgument), name
    with pytest.raises(capi.CfdError):
        sim.load_checkpoint(tmp_path / "missing.ckpt")

    after = [sim.download(f) for f in FIELDS], sim.loop
    assert after[1] == before[1]
    assert all(np.array_equal(x, y) for x, y in zip(after[0], before[0]))
    # and the good file still loads
    sim.load_checkpoint(good)
    sim.close()
