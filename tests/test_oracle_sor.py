"""The oracle's SOR restatements against what the mathematics fixes.

SorPipeLine_seq (inc/pressure/solver/SorPipeLine.seq.hpp:4-122) is pinned to
the compiled reference through the lexicographic restatement
(test_oracle_golden.py).  The red-black restatement of the same
volume-weighted update, which the GPU runs, is pinned here by: the same
fixed point as the lexicographic sweep (also with non-zero ghost
pressures), the dense solution of the Dirichlet ELL system when the ghosts
are zero, and the reference's reporting (iters = sweeps - 1, error = the
last sweep's max |pNEW - p|).  Red-black SOR on the CSR operator is pinned
by the consistent singular system it solves.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

import oracle_lib as O
from test_oracle_closed_forms import dense, ell, sgrid, ugrid


def predicted(g, seed):
    rng = np.random.default_rng(seed)
    return [1.0 + 0.05 * rng.standard_normal(g.G), 0.05 * rng.standard_normal(g.G),
            0.05 * rng.standard_normal(g.G)]


def interior(g, p):
    return p.reshape(g.nx, g.ny, g.nz)[2:-2, 2:-2, 2:-2].ravel()


def ghost_pressure(g, seed):
    """Random ghost layers, zero interior (a lagged BC's ghost values)."""
    p = np.random.default_rng(seed).standard_normal((g.nx, g.ny, g.nz))
    p[2:-2, 2:-2, 2:-2] = 0.0
    return np.ascontiguousarray(p.ravel())


def test_red_black_and_lexicographic_share_the_fixed_point():
    g = sgrid()
    u1 = predicted(g, 8)
    for ghosts in (np.zeros(g.G), ghost_pressure(g, 4)):
        p_lex, p_rb = ghosts.copy(), ghosts.copy()
        it = C.c_int()
        r1 = O.lib().orc_sor_lex(g.ref, 0.002, 1.8, 1e-13, 200000, *u1, p_lex, C.byref(it))
        r2 = O.lib().orc_sor_rb_vw(g.ref, 0.002, 1.8, 1e-13, 200000, *u1, p_rb, C.byref(it))
        assert r1 <= 1e-13 and r2 <= 1e-13
        a, b = interior(g, p_lex), interior(g, p_rb)
        assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-10
        # the ghost layers are read, never written
        mask = np.ones((g.nx, g.ny, g.nz), bool)
        mask[2:-2, 2:-2, 2:-2] = False
        assert np.array_equal(p_rb.reshape(g.nx, g.ny, g.nz)[mask], ghosts.reshape(g.nx, g.ny, g.nz)[mask])


def test_red_black_fixed_point_is_the_dirichlet_ell_solution():
    """Ghost p = 0: the volume-weighted system is the Jacobi-scaled ELL
    system with p = 0 on every face and the RHS of createBMatrix."""
    g = sgrid()
    u1 = predicted(g, 9)
    p = np.zeros(g.G)
    it = C.c_int()
    O.lib().orc_sor_rb_vw(g.ref, 0.002, 1.8, 1e-13, 200000, *u1, p, C.byref(it))
    col, val, jp = ell(g, (1,) * 6)
    b = np.zeros(g.N)
    O.lib().orc_rhs(g.ref, 0.002, jp, *u1, b)
    x = np.linalg.solve(dense(g, col, val), b)
    assert np.linalg.norm(interior(g, p) - x) / np.linalg.norm(x) < 1e-9


def test_iters_and_error_follow_the_reference():
    """iters = sweeps - 1 and error = the last sweep's max |pNEW - p|, which is
    max |p_after - p_before| / omega up to rounding."""
    g = ugrid((10, 8, 6), (2.5, 2.0, 1.5))
    u1 = predicted(g, 10)
    omega = 1.6
    ps = {}
    for sweeps in (5, 6):
        p = ghost_pressure(g, 11)
        it = C.c_int()
        err = O.lib().orc_sor_rb_vw(g.ref, 0.002, omega, 0.0, sweeps, *u1, p, C.byref(it))
        assert it.value == sweeps - 1
        ps[sweeps] = (p, err)
    step = np.abs(ps[6][0] - ps[5][0]).max() / omega
    assert abs(ps[6][1] - step) <= 1e-9 * step
    # the stop test: a criterion just above the sixth sweep's change stops there
    p = ghost_pressure(g, 11)
    it = C.c_int()
    err = O.lib().orc_sor_rb_vw(g.ref, 0.002, omega, ps[6][1] * (1 + 1e-12), 1000, *u1, p, C.byref(it))
    assert it.value <= 5 and err <= ps[6][1] * (1 + 1e-12)


def csr_operator(g):
    ptr = np.zeros(g.N + 1, np.int32)
    col = np.zeros(7 * g.N, np.int32)
    val = np.zeros(7 * g.N)
    jp = np.zeros(g.N)
    O.lib().orc_csr_assemble(g.ref, ptr, col, val, jp)
    nnz = int(ptr[-1])
    return ptr, col[:nnz].copy(), val[:nnz].copy()


def test_csr_red_black_sor_solves_the_consistent_system():
    g = ugrid((8, 6, 4), (2.0, 1.5, 1.0))
    ptr, col, val = csr_operator(g)
    A = np.zeros((g.N, g.N))
    for r in range(g.N):
        A[r, col[ptr[r]:ptr[r + 1]]] = val[ptr[r]:ptr[r + 1]]
    x_true = np.random.default_rng(6).standard_normal(g.N)
    b = A @ x_true  # consistent with the all-Neumann operator's null space
    x = np.zeros(g.N)
    it = C.c_int()
    rel = O.lib().orc_rbsor_csr(g.N, g.ref, ptr, col, val, b, x, 1.7, 1e-11, 200000, 10, C.byref(it))
    assert rel < 1e-11
    d = (x - x.mean()) - (x_true - x_true.mean())
    assert np.linalg.norm(d) / np.linalg.norm(x_true - x_true.mean()) < 1e-8
