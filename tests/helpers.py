"""Shared helpers of the parity tests: run the oracle and the HIP library on
the same inputs and compare."""
from __future__ import annotations

import ctypes as C

import numpy as np

import oracle_lib as O
from cfd_nssolver_forhpc_amd import capi
from cfd_nssolver_forhpc_amd.capi import GridArrays


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    if nb == 0.0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / nb)


def bc_tables(params):
    Dir = ((C.c_double * 3) * 6)()
    Num = ((C.c_double * 3) * 6)()
    for b in range(6):
        for c in range(3):
            Dir[b][c] = params.bc_dir[b][c]
            Num[b][c] = params.bc_num[b][c]
    return Dir, Num


def pd6(params):
    return O.ints6([params.p_dirichlet[f] for f in range(6)])


def ghost_mask(shape):
    """Boolean mask over a ghosted field: True for interior cells."""
    nx, ny, nz = shape
    m = np.zeros((nx, ny, nz), dtype=bool)
    m[2:nx - 2, 2:ny - 2, 2:nz - 2] = True
    return m.ravel()


def random_state(grid: GridArrays, seed: int, amp=0.3):
    """Smooth-ish random ghosted velocity fields for operator-level tests."""
    rng = np.random.default_rng(seed)
    G = grid.G
    base = [1.0, 0.0, 0.0]
    return [base[c] + amp * rng.standard_normal(G) for c in range(3)]


# ---- bicgstabRe2's failure-handling case (the reference harness's `rules` mode)
RULES_N = 200
RULES_SEED = 20240611


def rules_system():
    """The 200-row tridiagonal system of the harness's rules mode in the
    7-slot slot-major ELL layout: each row's entries in the order the
    harness sets them (i-1, i, i+1), the rest padded like ELL_matrix does,
    (column N-1, 0.0) (ELL_sparseMatrix.hpp:157-158)."""
    n = RULES_N
    col = np.full((7, n), n - 1, dtype=np.int32)
    val = np.zeros((7, n))
    for i in range(n):
        entries = ([(i - 1, -1.0)] if i > 0 else []) + [(i, 2.0 + 0.01 * (i % 7))] + \
                  ([(i + 1, -1.0)] if i < n - 1 else [])
        for s, (c, v) in enumerate(entries):
            col[s, i] = c
            val[s, i] = v
    return col, val


def rules_rhs(tag, amp=1.0):
    return O.noise(RULES_N, RULES_SEED, tag, 0.0, amp)
