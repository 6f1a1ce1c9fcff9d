"""Host-side mirror of the reference solver's interface for the hot path.

``Simulation`` owns one device context of ``libcfd_hip.so`` and exposes the
reference's functions under their own names (``ConvectionDifussion``,
``createBMatrix``, ``createPressureMatrix``, ``update_UandF``,
``BC_staggered_main`` ...), each a single call through the C-ABI.  Errors
arrive as status codes and are raised as :class:`CfdError` (or
:class:`CfdInvalidArgument` where the reference throws
``std::invalid_argument``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import capi
from .capi import CfdParams, CfdStepStats, GridArrays, check, dptr


# ---------------------------------------------------------------------------
# grids (inc/grid/structureGrid.hpp, inc/grid/Generator.hpp)
# ---------------------------------------------------------------------------

def axis_uniform(ncal: int, L: float):
    """grid::init_uniform for one axis: returns (X, D, Ds)."""
    n = ncal + 4
    X, D, Ds = np.zeros(n + 1), np.zeros(n), np.zeros(n)
    check(capi.load().cfd_axis_uniform(ncal, L, dptr(X), dptr(D), dptr(Ds)), "cfd_axis_uniform")
    return X, D, Ds


def axis_sml(ncal: int, L: float, lS: float, nS: int, off_nS: int):
    """grid::unit_SML + init_gC_dx for one axis."""
    n = ncal + 4
    X, D, Ds = np.zeros(n + 1), np.zeros(n), np.zeros(n)
    check(capi.load().cfd_axis_sml(ncal, L, lS, nS, off_nS, dptr(X), dptr(D), dptr(Ds)), "cfd_axis_sml")
    return X, D, Ds


def axis_tanh(ncal: int, L: float, xc: float, h0: float, band: float, ratio: float):
    """Stretched axis (large configuration): uniform band around xc, growth `ratio` outside."""
    n = ncal + 4
    X, D, Ds = np.zeros(n + 1), np.zeros(n), np.zeros(n)
    check(capi.load().cfd_axis_tanh(ncal, L, xc, h0, band, ratio, dptr(X), dptr(D), dptr(Ds)), "cfd_axis_tanh")
    return X, D, Ds


def uniform_grid(ncal, lengths) -> GridArrays:
    ax = [axis_uniform(n, L) for n, L in zip(ncal, lengths)]
    return GridArrays(tuple(ncal), ax[0][0], ax[1][0], ax[2][0], ax[0][1], ax[1][1], ax[2][1],
                      ax[0][2], ax[1][2], ax[2][2], *[float(L) for L in lengths])


def native_grid() -> GridArrays:
    """The reference's compiled-in grid: 190x210x40 'non_uniform-SML' over
    17x12x3 (structureGrid.hpp:26-46, 388-405)."""
    lx, ly, lz = 17.0, 12.0, 3.0
    ax = axis_sml(190, lx, 0.3 * lx, 100, -4)
    ay = axis_sml(210, ly, 0.3 * ly, 100, 0)
    az = axis_sml(40, lz, 0.5 * lz, 0, 0)
    return GridArrays((190, 210, 40), ax[0], ay[0], az[0], ax[1], ay[1], az[1], ax[2], ay[2], az[2], lx, ly, lz)


# ---------------------------------------------------------------------------
# parameters
# ---------------------------------------------------------------------------

def default_params() -> CfdParams:
    """The reference's native settings (src/CFD_MX.cpp:35-57, simuClass.hpp,
    BC_Cavity_A table, all-Neumann pressure, ELL BiCGSTAB)."""
    p = CfdParams()
    check(capi.load().cfd_params_default(C.byref(p)), "cfd_params_default")
    return p


# ---------------------------------------------------------------------------
# the simulation object
# ---------------------------------------------------------------------------

class Simulation:
    """One device context: fields, pressure operator and solver state."""

    def __init__(self, grid: GridArrays, params: CfdParams, device: int = 0):
        self.lib = capi.load()
        self.grid = grid
        self.params = params.copy()
        self._desc = grid.desc()
        h = C.c_void_p()
        check(self.lib.cfd_create(C.byref(self._desc), C.byref(self.params), device, C.byref(h)), "cfd_create")
        self.h = h
        G, N = C.c_size_t(), C.c_size_t()
        check(self.lib.cfd_sizes(self.h, C.byref(G), C.byref(N)), "cfd_sizes")
        self.G, self.N = G.value, N.value

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.cfd_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- configuration ----------------------------------------------------
    def set_params(self, params: CfdParams):
        self.params = params.copy()
        check(self.lib.cfd_set_params(self.h, C.byref(self.params)), "cfd_set_params")

    # ---- data movement ------------------------------------------------------
    def field_len(self, field: int) -> int:
        if field in capi.GHOSTED_FIELDS:
            return self.G
        if field in capi.FORCE_FIELDS:
            return 3 * self.N
        return self.N

    def upload(self, field: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        check(self.lib.cfd_upload(self.h, field, dptr(a), a.size), "cfd_upload")

    def download(self, field: int) -> np.ndarray:
        out = np.empty(self.field_len(field), dtype=np.float64)
        check(self.lib.cfd_download(self.h, field, dptr(out), out.size), "cfd_download")
        return out

    def ell(self):
        """(col, val) slot-major [7][N] arrays of the assembled ELL operator."""
        col = np.empty(7 * self.N, dtype=np.int32)
        val = np.empty(7 * self.N, dtype=np.float64)
        check(self.lib.cfd_ell_download(self.h, col.ctypes.data_as(C.POINTER(C.c_int32)), dptr(val), self.N, 7),
              "cfd_ell_download")
        return col, val

    def csr(self):
        """(ptr, col, val) of the assembled CSR operator (createPressureMatrixCSR)."""
        nnz = C.c_size_t()
        check(self.lib.cfd_csr_nnz(self.h, C.byref(nnz)), "cfd_csr_nnz")
        ptr = np.empty(self.N + 1, dtype=np.int32)
        col = np.empty(nnz.value, dtype=np.int32)
        val = np.empty(nnz.value, dtype=np.float64)
        i32 = C.POINTER(C.c_int32)
        check(self.lib.cfd_csr_download(self.h, ptr.ctypes.data_as(i32), col.ctypes.data_as(i32), dptr(val),
                                        self.N, nnz.value), "cfd_csr_download")
        return ptr, col, val

    def set_ell(self, col: np.ndarray, val: np.ndarray):
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        check(self.lib.cfd_ell_upload(self.h, col.ctypes.data_as(C.POINTER(C.c_int32)), dptr(val), self.N, 7),
              "cfd_ell_upload")

    # ---- set-up (RUN_CPU_PC_METHOD.hpp:59-264) --------------------------------
    def init_state(self, noise: float = 0.0, seed: int = 1):
        """iniU + seeded interior noise, p = 0, first BC_staggered_main/copy."""
        check(self.lib.cfd_init_state(self.h, float(noise), int(seed)), "cfd_init_state")

    def DFIB_CylinderZ(self):
        check(self.lib.cfd_set_eta_cylinder_z(self.h), "cfd_set_eta_cylinder_z")

    def createPressureMatrix(self):
        check(self.lib.cfd_ell_assemble(self.h), "cfd_ell_assemble")

    def setupPressure(self):
        """SPE_matrix1::setupPressure + set_Jp_pc (matrix-free operator)."""
        check(self.lib.cfd_spe_assemble(self.h), "cfd_spe_assemble")

    def createPressureMatrixCSR(self):
        """grid::createPressureMatrixCSR on the device; jp follows it when op is CSR."""
        check(self.lib.cfd_csr_assemble(self.h), "cfd_csr_assemble")

    def setup(self, noise: float = 0.0, seed: int = 1):
        """Everything the reference does before its main loop."""
        self.init_state(noise, seed)
        self.createPressureMatrix()
        if self.params.op == capi.CFD_OP_SPE:
            self.setupPressure()
        elif self.params.op == capi.CFD_OP_CSR:
            self.createPressureMatrixCSR()
        self.DFIB_CylinderZ()

    # ---- operators and solves ---------------------------------------------------
    def spmv(self, x_field=capi.X, y_field=capi.VEC_A, op=None, precision=capi.CFD_FP64):
        op = self.params.op if op is None else op
        check(self.lib.cfd_spmv(self.h, op, precision, x_field, y_field), "cfd_spmv")

    def bicgstab(self, tol=None, maxit=None):
        it, rel = C.c_int(), C.c_double()
        tol = self.params.tol if tol is None else tol
        maxit = self.params.maxit if maxit is None else maxit
        check(self.lib.cfd_bicgstab(self.h, tol, maxit, C.byref(it), C.byref(rel)), "cfd_bicgstab")
        return it.value, rel.value

    def SorPipeLine(self, omega=None, crit=None, itmax=None):
        """SorPipeLine_seq on the device's T1 and ghosted p (in place):
        (simu.iters, simu.error) = (sweeps - 1, last sweep's max |dp|)."""
        it, err = C.c_int(), C.c_double()
        omega = self.params.sor_omega if omega is None else omega
        crit = self.params.tol if crit is None else crit
        itmax = self.params.maxit if itmax is None else itmax
        check(self.lib.cfd_sor_pipeline(self.h, omega, crit, itmax, C.byref(it), C.byref(err)), "cfd_sor_pipeline")
        return it.value, err.value

    def rbsor(self, omega=None, tol=None, maxit=None, check_every=None):
        it, rel = C.c_int(), C.c_double()
        omega = self.params.sor_omega if omega is None else omega
        tol = self.params.tol if tol is None else tol
        maxit = self.params.maxit if maxit is None else maxit
        check_every = self.params.sor_check if check_every is None else check_every
        check(self.lib.cfd_rbsor(self.h, omega, tol, maxit, check_every, C.byref(it), C.byref(rel)), "cfd_rbsor")
        return it.value, rel.value

    def defect_correction(self, tol=None, tol_inner=None, max_outer=50):
        o, inner, rel = C.c_int(), C.c_int(), C.c_double()
        tol = self.params.tol if tol is None else tol
        tol_inner = self.params.tol_inner if tol_inner is None else tol_inner
        check(self.lib.cfd_defect_correction(self.h, tol, tol_inner, max_outer, C.byref(o), C.byref(inner),
                                             C.byref(rel)), "cfd_defect_correction")
        return o.value, inner.value, rel.value

    def solve(self):
        """pSolver(Mx.matB, Mx.X_result) with the configured solver."""
        it, rel = C.c_int(), C.c_double()
        check(self.lib.cfd_solve_pressure(self.h, C.byref(it), C.byref(rel)), "cfd_solve_pressure")
        return it.value, rel.value

    # ---- timestep phases ------------------------------------------------------------
    def ConvectionDifussion(self):
        check(self.lib.cfd_predictor(self.h), "cfd_predictor")

    def add_source(self):
        check(self.lib.cfd_source_add(self.h), "cfd_source_add")

    def createBMatrix(self):
        check(self.lib.cfd_build_rhs(self.h), "cfd_build_rhs")

    def Pressure_transform_X_result(self):
        check(self.lib.cfd_pressure_scatter(self.h), "cfd_pressure_scatter")

    def update_UandF(self):
        check(self.lib.cfd_dfib_update(self.h), "cfd_dfib_update")

    def accumulate_source(self):
        check(self.lib.cfd_source_accumulate(self.h), "cfd_source_accumulate")

    def BC_staggered_main(self):
        check(self.lib.cfd_bc_main(self.h), "cfd_bc_main")

    def BC_staggered_copy(self):
        check(self.lib.cfd_bc_copy(self.h), "cfd_bc_copy")

    def drag(self):
        cd, cl = C.c_double(), C.c_double()
        check(self.lib.cfd_drag(self.h, C.byref(cd), C.byref(cl)), "cfd_drag")
        return cd.value, cl.value

    def divergence(self, vel_field: int = capi.T0_U) -> dict:
        """FP64 divergence check of a velocity triple (the pressure RHS's
        finite-volume divergence): max and L2 over every interior cell and
        over the fluid cells."""
        out = (C.c_double * 4)()
        check(self.lib.cfd_divergence(self.h, int(vel_field), out), "cfd_divergence")
        return {"max": out[0], "l2": out[1], "fluid_max": out[2], "fluid_l2": out[3]}

    def write_plot3d_q(self, path, time: float, vel_field: int = capi.T3_U):
        """OutputPLOT3D_Qfile (inc/io/Plot3D/qfile.hpp:7-104) from the device
        fields: p, the cell-centred velocity of vel_field's triple, eta."""
        check(self.lib.cfd_plot3d_write_q(self.h, os.fsencode(path), int(vel_field), float(time)),
              "cfd_plot3d_write_q")

    PHASES = ("convection", "pressure", "update", "bc")

    def phase_timing(self, on: bool = True):
        """Per-phase device timers of step() (the reference's clockstruct
        stopwatches); switching clears the totals."""
        check(self.lib.cfd_phase_timing(self.h, 1 if on else 0), "cfd_phase_timing")

    def phase_times(self) -> dict:
        """Milliseconds per phase summed over the timed steps, and their number."""
        ms = np.zeros(len(self.PHASES))
        n = C.c_int()
        check(self.lib.cfd_phase_times(self.h, dptr(ms), C.byref(n)), "cfd_phase_times")
        return {"steps": n.value, "ms": dict(zip(self.PHASES, ms.tolist()))}

    def solver_counters(self) -> dict:
        """bicgstabRe2's timestep_, numRestarts_, pResets_ and totalNumIter_
        over this context's FP64 BiCGSTAB solves."""
        out = capi.CfdSolverCounters()
        check(self.lib.cfd_get_solver_counters(self.h, C.byref(out)), "cfd_get_solver_counters")
        return {"solves": out.solves, "restarts": out.restarts, "resets": out.resets, "iters": out.iters}

    def save_checkpoint(self, path):
        """Exact FP64 snapshot of the step state (cfd_checkpoint_save)."""
        check(self.lib.cfd_checkpoint_save(self.h, os.fsencode(path)), "cfd_checkpoint_save")

    def load_checkpoint(self, path):
        """Resume from a snapshot written for this grid (cfd_checkpoint_load);
        the context must be set up with the same parameters and operator."""
        check(self.lib.cfd_checkpoint_load(self.h, os.fsencode(path)), "cfd_checkpoint_load")

    def CheckSteadyStatebyMaxVal(self):
        out = (C.c_double * 3)()
        check(self.lib.cfd_maxdiff(self.h, out), "cfd_maxdiff")
        return tuple(out)

    def copy_T3_to_T0(self):
        check(self.lib.cfd_copy_t3_to_t0(self.h), "cfd_copy_t3_to_t0")

    # ---- timestep ---------------------------------------------------------------------
    def step(self) -> CfdStepStats:
        st = CfdStepStats()
        check(self.lib.cfd_step(self.h, C.byref(st)), "cfd_step")
        return st

    @property
    def loop(self) -> int:
        lp, t = C.c_int(), C.c_double()
        check(self.lib.cfd_loop(self.h, C.byref(lp), C.byref(t)), "cfd_loop")
        return lp.value

    def set_loop(self, loop: int, time: float = 0.0):
        check(self.lib.cfd_set_loop(self.h, loop, time), "cfd_set_loop")

    # ---- discrete adjoint -------------------------------------------------------------
    def set_inflow_profile(self, u_in):
        """Face-0 u Dirichlet value per interior row (ny_cal values); None: the table value."""
        if u_in is None:
            check(self.lib.cfd_set_inflow_profile(self.h, None, 0), "cfd_set_inflow_profile")
            return
        a = np.ascontiguousarray(u_in, dtype=np.float64)
        check(self.lib.cfd_set_inflow_profile(self.h, dptr(a), a.size), "cfd_set_inflow_profile")

    def set_forcing_gain(self, gain: float):
        check(self.lib.cfd_set_forcing_gain(self.h, float(gain)), "cfd_set_forcing_gain")

    def adjoint_forward(self, nsteps: int, avg_from: int, ckpt_every: int) -> float:
        """Run nsteps timesteps with checkpoints; the mean corrector cD over [avg_from, nsteps)."""
        J = C.c_double()
        check(self.lib.cfd_adjoint_forward(self.h, nsteps, avg_from, ckpt_every, C.byref(J)), "cfd_adjoint_forward")
        return J.value

    def adjoint_generation(self) -> int:
        """Number of the objective run adjoint_backward would differentiate (0: none)."""
        g = C.c_ulonglong()
        check(self.lib.cfd_adjoint_generation(self.h, C.byref(g)), "cfd_adjoint_generation")
        return g.value

    def adjoint_backward(self):
        """(dJ/du_in, dJ/dgain, stats) of the last adjoint_forward run."""
        g = np.zeros(self.grid.ncal[1])
        gg = C.c_double()
        st = capi.CfdAdjointStats()
        check(self.lib.cfd_adjoint_backward(self.h, dptr(g), g.size, C.byref(gg), C.byref(st)),
              "cfd_adjoint_backward")
        return g, gg.value, st.as_dict()

    # ---- measurement ------------------------------------------------------------------
    def time_kernel(self, what: int, nrep: int) -> float:
        ms = C.c_double()
        check(self.lib.cfd_time_kernel(self.h, what, nrep, C.byref(ms)), "cfd_time_kernel")
        return ms.value
