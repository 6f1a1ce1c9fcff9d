"""The allocation-free BiCGSTAB entry point against cfd_bicgstab.

cfd_bicgstab_ell_noalloc solves on caller-owned device buffers with its work
space in the library's static device array.  It runs the kernels of
cfd_bicgstab on a fresh context in the same order, so its solution,
iteration count and residual must equal that entry point's bit for bit.
Two calls in a row share the array, the second on a smaller system; a third
stops at the iteration cap.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from cfd_nssolver_forhpc_amd import capi, configs
from cfd_nssolver_forhpc_amd.solver import Simulation, uniform_grid

pytestmark = pytest.mark.gpu


def reference_solve(ncal, seed, tol, maxit):
    """cfd_bicgstab on a new context: the ELL operator of a channel with a
    Dirichlet outflow face, a seeded N(0,1) right-hand side and warm start."""
    grid = uniform_grid(ncal, (10.0, 6.0, 3.0))
    p = configs.cylinder_params(100.0, 0.002, 2.5, 3.0, 1.0, tol)
    sim = Simulation(grid, p)
    sim.init_state()
    sim.createPressureMatrix()
    rng = np.random.default_rng(seed)
    b = rng.standard_normal(sim.N)
    x0 = 0.1 * rng.standard_normal(sim.N)
    sim.upload(capi.RHS, b)
    sim.upload(capi.X, x0)
    it, rel = sim.bicgstab(tol=tol, maxit=maxit)
    col, val = sim.ell()
    x = sim.download(capi.X)
    sim.close()
    return col, val, b, x0, x, it, rel


def hip_runtime():
    """The HIP runtime the library links (device buffers for the caller's side
    of the call; torch carries a HIP runtime of its own, which must not be
    initialised in the same process)."""
    capi.load()
    L = C.CDLL("libamdhip64.so.7")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipFree.argtypes = [C.c_void_p]
    return L


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, host: np.ndarray):
        self.L = hip_runtime()
        self.host = np.ascontiguousarray(host)
        self.ptr = C.c_void_p()
        assert self.L.hipMalloc(C.byref(self.ptr), self.host.nbytes) == 0
        assert self.L.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, H2D) == 0

    def get(self) -> np.ndarray:
        out = np.empty_like(self.host)
        assert self.L.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr.value:
            self.L.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def noalloc_solve(col, val, b, x0, tol, maxit):
    arrs = [DeviceArray(a) for a in (col, val, b, x0.copy())]
    try:
        it, rel = C.c_int(), C.c_double()
        rc = capi.load().cfd_bicgstab_ell_noalloc(arrs[0].ptr, arrs[1].ptr, arrs[2].ptr, arrs[3].ptr, b.size, tol,
                                                 maxit, None, C.byref(it), C.byref(rel))
        capi.check(rc, "cfd_bicgstab_ell_noalloc")
        return arrs[3].get(), it.value, rel.value
    finally:
        for a in arrs:
            a.free()


@pytest.mark.parametrize("case", [((40, 24, 12), 11, 1e-10, 20000),   # larger system first
                                  ((16, 8, 8), 12, 1e-10, 20000),     # then a smaller one
                                  ((24, 12, 8), 13, 1e-12, 5)])       # stops at the iteration cap
def test_noalloc_matches_cfd_bicgstab(require_gpu, case):
    ncal, seed, tol, maxit = case
    print(f"noalloc case ncal={ncal} seed={seed} tol={tol} maxit={maxit}", flush=True)
    col, val, b, x0, x_ref, it_ref, rel_ref = reference_solve(ncal, seed, tol, maxit)
    x, it, rel = noalloc_solve(col, val, b, x0, tol, maxit)
    assert it == it_ref and rel == rel_ref, (it, it_ref, rel, rel_ref)
    assert np.array_equal(x, x_ref)
    if maxit > 100:
        assert rel < tol


def test_noalloc_twice_in_a_row(require_gpu):
    """The array keeps the previous call's contents; a repeated and then a
    smaller call still give cfd_bicgstab's results."""
    big = reference_solve((32, 16, 16), 21, 1e-10, 20000)
    small = reference_solve((8, 8, 4), 22, 1e-10, 20000)
    for col, val, b, x0, x_ref, it_ref, rel_ref in (big, small, big, small):
        x, it, rel = noalloc_solve(col, val, b, x0, 1e-10, 20000)
        assert it == it_ref and rel == rel_ref
        assert np.array_equal(x, x_ref)


def test_noalloc_rejects_oversized_system(require_gpu):
    t = DeviceArray(np.zeros(8))
    try:
        it, rel = C.c_int(), C.c_double()
        rc = capi.load().cfd_bicgstab_ell_noalloc(t.ptr, t.ptr, t.ptr, t.ptr, capi.CFD_NOALLOC_MAX_ROWS + 1, 1e-10,
                                                 10, None, C.byref(it), C.byref(rel))
        assert rc == capi.CFD_ERR_ARG
        assert b"exceeds the maximum" in capi.load().cfd_last_error()
    finally:
        t.free()
