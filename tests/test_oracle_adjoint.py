"""The oracle's discrete adjoint (an addition with no reference counterpart)
against what fixes it independently of its own formulas:

* every transposed piece satisfies <ly, J dx> = <J^T ly, dx>, with J dx taken
  from the forward oracle function itself: exact for the affine pieces (BC,
  DFIB update, RHS, the ELL operator), central differences for the
  piecewise-quadratic QUICK predictor (exact up to rounding while no upwind
  choice flips);
* the gradient of the time-averaged drag with respect to the inflow profile
  and the forcing gain matches central finite differences of the forward
  oracle run.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from cfd_nssolver_forhpc_amd import capi, configs
from test_oracle_closed_forms import cd_mask, dense, ell, sgrid

TABLE = (C.c_double * 3) * 6


def table(rows):
    return TABLE(*[(C.c_double * 3)(*r) for r in rows])


def eta_of(g, cx, cy, r):
    eta = np.zeros(g.G)
    O.lib().orc_eta_cylinder_z(g.ref, cx, cy, r, 100, eta)
    return eta


def rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


def test_ell_transpose_is_exact():
    g = sgrid()
    col, val, _ = ell(g, (0, 1, 0, 0, 0, 0))
    colT = np.zeros_like(col)
    valT = np.zeros_like(val)
    assert O.lib().orc_ell_transpose(g.N, col, val, colT, valT) == 0
    assert np.array_equal(dense(g, colT, valT), dense(g, col, val).T)


@pytest.mark.parametrize("seed", [0, 1])
def test_bc_adjoint_dot_product(seed):
    g = sgrid()
    rng = np.random.default_rng(seed)
    num = rng.choice([0.0, 1.0, 0.5, -0.25], size=(6, 3))
    zero = table(np.zeros((6, 3)))
    N_ = table(num)
    dx = [rng.standard_normal(g.G) for _ in range(3)]
    Mdx = [a.copy() for a in dx]
    O.lib().orc_bc_main_ex(g.ref, C.byref(zero), C.byref(N_), O.ints6([0] * 6), None, *Mdx, np.zeros(g.G))
    ly = [rng.standard_normal(g.G) for _ in range(3)]
    lx = [a.copy() for a in ly]
    g_uin = np.zeros(g.nyc)
    O.lib().orc_bc_main_adjoint(g.ref, C.byref(N_), g_uin.ctypes.data_as(C.c_void_p), *lx)
    lhs = sum(np.dot(a, b) for a, b in zip(ly, Mdx))
    rhs = sum(np.dot(a, b) for a, b in zip(lx, dx))
    assert rel(lhs, rhs) < 1e-13
    # the inflow profile enters through the face-0 u ghosts only
    du = rng.standard_normal(g.nyc)
    y = [np.zeros(g.G) for _ in range(3)]
    O.lib().orc_bc_main_ex(g.ref, C.byref(zero), C.byref(N_), O.ints6([0] * 6), du.ctypes.data_as(C.c_void_p),
                           *y, np.zeros(g.G))
    assert rel(sum(np.dot(a, b) for a, b in zip(ly, y)), np.dot(g_uin, du)) < 1e-13


def _bl_params(g, seed):
    """The workloads' boundary mode (convective outflow, periodic z) with
    random Neumann factors and no Dirichlet constants, so that BC is linear."""
    p = configs.outflow_periodic_bc(configs.cylinder_params(100.0, 0.002, 1.5, 1.0, 0.45, 1e-10))
    p.outflow_speed = 0.7
    rng = np.random.default_rng(seed)
    for f in range(6):
        for c in range(3):
            p.bc_dir[f][c] = 0.0
            p.bc_num[f][c] = float(rng.choice([0.0, 1.0, 0.5, -0.25]))
    return p


@pytest.mark.parametrize("seed", [0, 1])
def test_bc_adjoint_dot_product_convective_periodic(seed):
    """The transposed boundary update of the workloads' boundaries: the
    convective outflow (linear in the interior face and the old ghost) and
    the periodic z images, after the table's x and y faces."""
    g = sgrid()
    p = _bl_params(g, seed)
    rng = np.random.default_rng(10 + seed)
    dx = [rng.standard_normal(g.G) for _ in range(3)]
    Mdx = [a.copy() for a in dx]
    O.lib().orc_bc_main_bl(g.ref, C.byref(p), None, *Mdx, np.zeros(g.G))
    ly = [rng.standard_normal(g.G) for _ in range(3)]
    lx = [a.copy() for a in ly]
    g_uin = np.zeros(g.nyc)
    O.lib().orc_bc_main_adjoint_bl(g.ref, C.byref(p), g_uin.ctypes.data_as(C.c_void_p), *lx)
    lhs = sum(np.dot(a, b) for a, b in zip(ly, Mdx))
    rhs = sum(np.dot(a, b) for a, b in zip(lx, dx))
    assert rel(lhs, rhs) < 1e-13
    du = rng.standard_normal(g.nyc)
    y = [np.zeros(g.G) for _ in range(3)]
    O.lib().orc_bc_main_bl(g.ref, C.byref(p), du.ctypes.data_as(C.c_void_p), *y, np.zeros(g.G))
    assert rel(sum(np.dot(a, b) for a, b in zip(ly, y)), np.dot(g_uin, du)) < 1e-13


def test_dfib_adjoint_dot_product_convective_periodic():
    """The transposed update with the outflow face (no forcing, p = 0
    outside) and the seam face of w (its + neighbour the first plane, through
    the periodic pressure ghosts)."""
    g = sgrid()
    p = _bl_params(g, 2)
    rng = np.random.default_rng(13)
    eta = eta_of(g, 1.5, 1.0, 0.45)
    gain = 0.8
    dT1 = [rng.standard_normal(g.G) for _ in range(3)]
    dX = rng.standard_normal(g.N)
    dp_ = np.zeros(g.G)
    O.lib().orc_pressure_scatter(g.ref, dX, dp_)
    O.lib().orc_periodic_refresh_p(g.ref, dp_)

    def fwd(T1, pr, gn):
        u3 = [np.zeros(g.G) for _ in range(3)]
        f = np.zeros(3 * g.N)
        t2 = np.zeros(3 * g.N)
        O.lib().orc_dfib_update_bl(g.ref, C.byref(p), gn, eta, pr, *T1, *u3, f, t2.ctypes.data_as(C.c_void_p))
        return u3, f, t2

    u3, f, t2 = fwd(dT1, dp_, gain)
    lu3 = [rng.standard_normal(g.G) for _ in range(3)]
    lf = rng.standard_normal(3 * g.N)
    lT1 = [np.zeros(g.G) for _ in range(3)]
    lX = np.zeros(g.N)
    gg = C.c_double(0.0)
    O.lib().orc_dfib_adjoint_bl(g.ref, C.byref(p), gain, eta, t2, *[a.ctypes.data_as(C.c_void_p) for a in lu3], lf,
                                *lT1, lX, C.byref(gg))
    # (u3 stays 0 where the update does not write, so lu3 meets it only at
    # the updated faces, the outflow and seam faces among them)
    lhs = sum(np.dot(a, b) for a, b in zip(lu3, u3)) + np.dot(lf, f)
    rhs = sum(np.dot(a, b) for a, b in zip(lT1, dT1)) + np.dot(lX, dX)
    assert rel(lhs, rhs) < 1e-12
    u3a, fa, _ = fwd(dT1, dp_, 1.0)
    u3b, fb, _ = fwd(dT1, dp_, 0.0)
    dg = sum(np.dot(a, b - c_) for a, b, c_ in zip(lu3, u3a, u3b)) + np.dot(lf, fa - fb)
    assert rel(dg, gg.value) < 1e-12


def test_dfib_adjoint_dot_product():
    g = sgrid()
    rng = np.random.default_rng(3)
    eta = eta_of(g, 1.5, 1.0, 0.45)
    assert 0.0 < eta.max() and eta.min() == 0.0
    dt, gain = 0.002, 0.8
    dT1 = [rng.standard_normal(g.G) for _ in range(3)]
    dX = rng.standard_normal(g.N)
    dp_ = np.zeros(g.G)
    O.lib().orc_pressure_scatter(g.ref, dX, dp_)

    def fwd(T1, p, gn):
        u3 = [np.zeros(g.G) for _ in range(3)]
        f = np.zeros(3 * g.N)
        t2 = np.zeros(3 * g.N)
        O.lib().orc_dfib_update_ex(g.ref, dt, gn, eta, p, *T1, *u3, f, t2.ctypes.data_as(C.c_void_p))
        return u3, f, t2

    u3, f, t2 = fwd(dT1, dp_, gain)
    lu3 = [rng.standard_normal(g.G) for _ in range(3)]
    lf = rng.standard_normal(3 * g.N)
    lT1 = [np.zeros(g.G) for _ in range(3)]
    lX = np.zeros(g.N)
    gg = C.c_double(0.0)
    O.lib().orc_dfib_adjoint(g.ref, dt, gain, eta, t2, *[a.ctypes.data_as(C.c_void_p) for a in lu3], lf, *lT1, lX,
                             C.byref(gg))
    lhs = sum(np.dot(a, b) for a, b in zip(lu3, u3)) + np.dot(lf, f)
    rhs = sum(np.dot(a, b) for a, b in zip(lT1, dT1)) + np.dot(lX, dX)
    assert rel(lhs, rhs) < 1e-12
    # the outputs are affine in the gain at fixed T2
    u3a, fa, _ = fwd(dT1, dp_, 1.0)
    u3b, fb, _ = fwd(dT1, dp_, 0.0)
    dg = sum(np.dot(a, b - c) for a, b, c in zip(lu3, u3a, u3b)) + np.dot(lf, fa - fb)
    assert rel(dg, gg.value) < 1e-12


def test_rhs_adjoint_dot_product():
    g = sgrid()
    rng = np.random.default_rng(4)
    _, _, jp = ell(g, (0, 1, 0, 0, 0, 0))
    dT1 = [rng.standard_normal(g.G) for _ in range(3)]
    b = np.zeros(g.N)
    O.lib().orc_rhs(g.ref, 0.002, jp, *dT1, b)
    lb = rng.standard_normal(g.N)
    lT1 = [np.zeros(g.G) for _ in range(3)]
    O.lib().orc_rhs_adjoint(g.ref, 0.002, jp, lb, *lT1)
    assert rel(np.dot(lb, b), sum(np.dot(a, c) for a, c in zip(lT1, dT1))) < 1e-13


@pytest.mark.parametrize("seed", [5, 6, 7])
def test_predictor_adjoint_dot_product(seed):
    """Central differences of the forward predictor are exact for its
    piecewise-quadratic form (rounding aside) as long as no face velocity
    changes sign within the step; mixed-sign flow exercises every upwind
    branch of both QUICK forms on a non-uniform grid."""
    g = sgrid()
    rng = np.random.default_rng(seed)
    x = [0.3 + rng.standard_normal(g.G), 0.5 * rng.standard_normal(g.G), 0.5 * rng.standard_normal(g.G)]
    dx = [rng.standard_normal(g.G) for _ in range(3)]
    nu, dt = 0.01, 0.002
    eps = 1e-6

    def P(xx):
        out = [np.zeros(g.G) for _ in range(3)]
        O.lib().orc_predictor(g.ref, nu, dt, *xx, *out)
        return out

    Pp = P([a + eps * b for a, b in zip(x, dx)])
    Pm = P([a - eps * b for a, b in zip(x, dx)])
    Jdx = [(a - b) / (2 * eps) for a, b in zip(Pp, Pm)]
    ly = [rng.standard_normal(g.G) * cd_mask(g, c) for c in range(3)]
    lx = [np.zeros(g.G) for _ in range(3)]
    O.lib().orc_predictor_adjoint(g.ref, nu, dt, *x, *ly, *lx)
    lhs = sum(np.dot(a, b) for a, b in zip(ly, Jdx))
    rhs = sum(np.dot(a, b) for a, b in zip(lx, dx))
    assert rel(lhs, rhs) < 1e-8


def _case(tol=1e-13):
    w = configs.small_cylinder(ncal=(24, 12, 4), lengths=(6.0, 3.0, 1.0), tol=tol)
    w.params.maxit = 20000
    g = O.Grid(w.grid)
    rng = np.random.default_rng(3)
    base = 1.0 + 0.05 * rng.standard_normal(g.nyc)
    for _ in range(4):
        base = np.convolve(np.pad(base, 1, mode="edge"), [0.25, 0.5, 0.25], mode="valid")
    return w, g, base


def _objective(w, g, uin, gain, steps, avg_from, grad=False):
    s = O.Sim(g, w.params)
    s.set_inflow(uin)
    s.set_gain(gain)
    s.init_state(1e-3, 1)
    s.setup()
    J = s.objective(steps, avg_from)
    return (J, s.gradient()) if grad else J


def test_gradient_matches_central_differences():
    w, g, base = _case()
    steps, avg_from = 10, 5
    J, (g_uin, g_gain, it) = _objective(w, g, base, 1.0, steps, avg_from, grad=True)
    assert np.isfinite(J) and it > 0
    assert np.any(g_uin != 0.0) and g_gain != 0.0
    rng = np.random.default_rng(11)
    h = 1e-6
    for _ in range(4):
        d = rng.standard_normal(g.nyc + 1)
        d /= np.linalg.norm(d)
        Jp = _objective(w, g, base + h * d[:-1], 1.0 + h * d[-1], steps, avg_from)
        Jm = _objective(w, g, base - h * d[:-1], 1.0 - h * d[-1], steps, avg_from)
        fd = (Jp - Jm) / (2 * h)
        ad = np.dot(g_uin, d[:-1]) + g_gain * d[-1]
        assert rel(fd, ad) < 1e-5, (fd, ad)


def test_gradient_matches_central_differences_convective_periodic():
    """The workloads' boundaries (convective outflow, periodic z) through the
    whole objective: the oracle's gradient against its own forward run."""
    w, g, base = _case()
    configs.outflow_periodic_bc(w.params)
    steps, avg_from = 10, 5
    J, (g_uin, g_gain, it) = _objective(w, g, base, 0.9, steps, avg_from, grad=True)
    assert np.isfinite(J) and it > 0
    rng = np.random.default_rng(12)
    h = 1e-6
    dirs = [np.concatenate([1.0 + 0.5 * rng.standard_normal(g.nyc), [0.0]])]
    dirs += [rng.standard_normal(g.nyc + 1) for _ in range(2)]
    for d in dirs:
        d /= np.linalg.norm(d)
        Jp = _objective(w, g, base + h * d[:-1], 0.9 + h * d[-1], steps, avg_from)
        Jm = _objective(w, g, base - h * d[:-1], 0.9 - h * d[-1], steps, avg_from)
        fd = (Jp - Jm) / (2 * h)
        ad = np.dot(g_uin, d[:-1]) + g_gain * d[-1]
        assert rel(fd, ad) < 1e-5, (fd, ad)


def test_table_inflow_and_unit_gain_reproduce_the_reference_step():
    """The adjoint's parameters at their reference values leave the forward
    run bit-identical to the plain step."""
    w, g, base = _case(tol=1e-10)
    runs = []
    for uin in (None, np.full(g.nyc, w.params.bc_dir[0][0])):
        s = O.Sim(g, w.params)
        if uin is not None:
            s.set_inflow(uin)
            s.set_gain(1.0)
        s.init_state(1e-3, 1)
        s.setup()
        for _ in range(3):
            s.step()
        runs.append([s.field(f).copy() for f in (capi.T0_U, capi.T0_V, capi.T0_W, capi.F)])
    assert all(np.array_equal(a, b) for a, b in zip(*runs))
