"""ctypes view of the C-ABI declared in include/cfd_capi.h.

The HIP library (``_lib/libcfd_hip.so``) is the product; this module only
describes its structs and loads it.  Loading fails loudly when the library
has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
# CFD_HIP_LIB names another build of the same library (e.g. one whose host
# code is built with sanitizers)
LIB_PATH = Path(os.environ["CFD_HIP_LIB"]) if os.environ.get("CFD_HIP_LIB") else PKG / "_lib" / "libcfd_hip.so"

# status codes
CFD_OK = 0
CFD_ERR_ARG = -1
CFD_ERR_HIP = -2
CFD_ERR_STATE = -3
CFD_ERR_NOMEM = -4
CFD_ERR_RUNTIME = -5

CFD_SOLVER_BICGSTAB = 0
CFD_SOLVER_RBSOR = 1
CFD_SOLVER_SOR_PIPELINE = 2
CFD_OP_ELL = 0
CFD_OP_SPE = 1
CFD_OP_CSR = 2
CFD_FP64 = 0
CFD_MIXED = 1
CFD_MIXED_RELIABLE = 0
CFD_MIXED_RESTARTED = 1
CFD_MIXED_VALUES32 = 2
CFD_MIXED_PC32 = 3
CFD_PRECOND_JACOBI = 0
CFD_PRECOND_MG = 1
CFD_CONV_QUICK = 0
CFD_CONV_LUD = 1
CFD_CONV_UD = 2
CFD_CDCL_REFERENCE = 0
CFD_CDCL_BUFFERED = 1
CFD_CDCL_EVERY_STEP = 2
CFD_NOALLOC_MAX_ROWS = 1048576

# field ids
T0_U, T0_V, T0_W, T1_U, T1_V, T1_W, T3_U, T3_V, T3_W = range(9)
P, ETA, F, SRC, RHS, X, JP, VEC_A, VEC_B = range(9, 18)
NFIELDS = 18
GHOSTED_FIELDS = {T0_U, T0_V, T0_W, T1_U, T1_V, T1_W, T3_U, T3_V, T3_W, P, ETA}
FORCE_FIELDS = {F, SRC}
MAX_PASSES = 8


class CfdGridDesc(C.Structure):
    _fields_ = [
        ("nx_cal", C.c_int), ("ny_cal", C.c_int), ("nz_cal", C.c_int),
        ("X", C.POINTER(C.c_double)), ("Y", C.POINTER(C.c_double)), ("Z", C.POINTER(C.c_double)),
        ("Dx", C.POINTER(C.c_double)), ("Dy", C.POINTER(C.c_double)), ("Dz", C.POINTER(C.c_double)),
        ("Dxs", C.POINTER(C.c_double)), ("Dys", C.POINTER(C.c_double)), ("Dzs", C.POINTER(C.c_double)),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
    ]


class CfdParams(C.Structure):
    _fields_ = [
        ("Re", C.c_double), ("dt", C.c_double),
        ("bc_dir", (C.c_double * 3) * 6), ("bc_num", (C.c_double * 3) * 6),
        ("p_dirichlet", C.c_int * 6),
        ("source_index_quirk", C.c_int),
        ("inner_loops", C.c_int),
        ("solver", C.c_int), ("op", C.c_int), ("precision", C.c_int),
        ("tol", C.c_double), ("tol_inner", C.c_double), ("mixed_mode", C.c_int), ("mixed_delta", C.c_double),
        ("maxit", C.c_int),
        ("sor_omega", C.c_double), ("sor_check", C.c_int),
        ("ib_on", C.c_int), ("ib_cx", C.c_double), ("ib_cy", C.c_double), ("ib_r", C.c_double),
        ("ib_nsub", C.c_int),
        ("u_init", C.c_double * 3),
        ("precond", C.c_int), ("mg_nu", C.c_int),
        ("convection", C.c_int), ("les", C.c_int),
        ("periodic_z", C.c_int), ("outflow_convective", C.c_int), ("outflow_speed", C.c_double),
    ]

    def set_bc(self, dir_table, num_table):
        for b in range(6):
            for c in range(3):
                self.bc_dir[b][c] = float(dir_table[b][c])
                self.bc_num[b][c] = float(num_table[b][c])

    def copy(self) -> "CfdParams":
        out = CfdParams()
        C.pointer(out)[0] = self
        return out


class CfdStepStats(C.Structure):
    _fields_ = [
        ("loop", C.c_int), ("n_passes", C.c_int),
        ("iters", C.c_int * MAX_PASSES), ("relres", C.c_double * MAX_PASSES),
        ("cd", C.c_double * MAX_PASSES), ("cl", C.c_double * MAX_PASSES),
        ("maxdiff", C.c_double * 3),
        ("inner_iters", C.c_int * MAX_PASSES),
        ("div_max", C.c_double), ("div_l2", C.c_double), ("div_fluid_max", C.c_double),
        ("div_fluid_l2", C.c_double),
    ]

    def as_dict(self) -> dict:
        n = self.n_passes
        return {
            "loop": self.loop,
            "iters": list(self.iters[:n]),
            "relres": list(self.relres[:n]),
            "cd": list(self.cd[:n]),
            "cl": list(self.cl[:n]),
            "maxdiff": list(self.maxdiff),
            "inner_iters": list(self.inner_iters[:n]),
            "div": {"max": self.div_max, "l2": self.div_l2, "fluid_max": self.div_fluid_max,
                    "fluid_l2": self.div_fluid_l2},
        }


class CfdSolverCounters(C.Structure):
    _fields_ = [("solves", C.c_int), ("restarts", C.c_int), ("resets", C.c_int), ("pad", C.c_int),
                ("iters", C.c_longlong)]


class CfdAdjointStats(C.Structure):
    _fields_ = [
        ("steps", C.c_int), ("checkpoints", C.c_int), ("recomputed_steps", C.c_int),
        ("adjoint_solves", C.c_int), ("adjoint_iters", C.c_int),
        ("forward_ms", C.c_double), ("backward_ms", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# Velocity BC tables Dir[face][component], Num[face][component]; faces 0..5 =
# x-, x+, y-, y+, z-, z+ (inc/BoundaryCondition/Boundary_Condition.hpp:13-29).
# The lid-driven-cavity table the reference compiles in (BC_Cavity_A.hpp).
CAVITY_A_DIR = [[0, 0, 0], [0, 0, 0], [0, 0, 0], [1, 0, 0], [0, 0, 0], [0, 0, 0]]
CAVITY_A_NUM = [[0, 0, 0]] * 6
# Flow past a body: uniform inflow on x-, zero-gradient outflow on x+,
# zero-gradient elsewhere (inc/BoundaryCondition/FLOWpassing.hpp).
FLOWPASSING_DIR = [[1, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0]]
FLOWPASSING_NUM = [[0, 0, 0], [1, 1, 1], [1, 1, 1], [1, 1, 1], [1, 1, 1], [1, 1, 1]]
# Cylinder-in-channel table: inflow x-, zero-gradient outflow x+, slip walls
# on y and z (normal component 0, tangential zero-gradient).
CHANNEL_SLIP_DIR = [[1, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0]]
CHANNEL_SLIP_NUM = [[0, 0, 0], [1, 1, 1], [1, 0, 1], [1, 0, 1], [1, 1, 0], [1, 1, 0]]


@dataclass
class GridArrays:
    """Node coordinates and metrics of one structured grid (grid class of the
    reference, inc/grid/structureGrid.hpp:14-863)."""

    ncal: tuple
    X: np.ndarray
    Y: np.ndarray
    Z: np.ndarray
    Dx: np.ndarray
    Dy: np.ndarray
    Dz: np.ndarray
    Dxs: np.ndarray
    Dys: np.ndarray
    Dzs: np.ndarray
    lx: float
    ly: float
    lz: float
    _keep: list = field(default_factory=list, repr=False)

    def desc(self) -> CfdGridDesc:
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
                (self.X, self.Y, self.Z, self.Dx, self.Dy, self.Dz, self.Dxs, self.Dys, self.Dzs)]
        self._keep = arrs
        ptr = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs]
        return CfdGridDesc(self.ncal[0], self.ncal[1], self.ncal[2], *ptr, self.lx, self.ly, self.lz)

    @property
    def shape(self):
        return tuple(n + 4 for n in self.ncal)

    @property
    def G(self):
        nx, ny, nz = self.shape
        return nx * ny * nz

    @property
    def N(self):
        return self.ncal[0] * self.ncal[1] * self.ncal[2]

    def delta(self) -> np.ndarray:
        """Cell volumes Dx*Dy*Dz over interior cells (Generator.hpp:47-53)."""
        i = slice(2, self.ncal[0] + 2)
        j = slice(2, self.ncal[1] + 2)
        k = slice(2, self.ncal[2] + 2)
        return (self.Dx[i][:, None, None] * self.Dy[j][None, :, None] * self.Dz[k][None, None, :]).ravel()


_LIB = None


def load() -> C.CDLL:
    """Load the HIP library; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"HIP library {LIB_PATH} is missing: build it with __graft_entry__.build() "
            "(there is no CPU fallback for the product path)")
    L = C.CDLL(str(LIB_PATH))
    _declare(L)
    _LIB = L
    return L


def _declare(L: C.CDLL) -> None:
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    i32p = C.POINTER(C.c_int32)
    sig = {
        "cfd_abi_version": ([], C.c_int),
        "cfd_last_error": ([], C.c_char_p),
        "cfd_params_default": ([C.POINTER(CfdParams)], C.c_int),
        "cfd_axis_uniform": ([C.c_int, C.c_double, dp, dp, dp], C.c_int),
        "cfd_axis_sml": ([C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp, dp], C.c_int),
        "cfd_axis_tanh": ([C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp, dp],
                          C.c_int),
        "cfd_device_count": ([ip], C.c_int),
        "cfd_create": ([C.POINTER(CfdGridDesc), C.POINTER(CfdParams), C.c_int, C.POINTER(vp)], C.c_int),
        "cfd_destroy": ([vp], C.c_int),
        "cfd_set_params": ([vp, C.POINTER(CfdParams)], C.c_int),
        "cfd_sizes": ([vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)], C.c_int),
        "cfd_upload": ([vp, C.c_int, dp, C.c_size_t], C.c_int),
        "cfd_download": ([vp, C.c_int, dp, C.c_size_t], C.c_int),
        "cfd_init_state": ([vp, C.c_double, C.c_uint64], C.c_int),
        "cfd_set_eta_cylinder_z": ([vp], C.c_int),
        "cfd_ell_assemble": ([vp], C.c_int),
        "cfd_spe_assemble": ([vp], C.c_int),
        "cfd_csr_assemble": ([vp], C.c_int),
        "cfd_csr_nnz": ([vp, C.POINTER(C.c_size_t)], C.c_int),
        "cfd_csr_download": ([vp, i32p, i32p, dp, C.c_size_t, C.c_size_t], C.c_int),
        "cfd_ell_download": ([vp, i32p, dp, C.c_size_t, C.c_int], C.c_int),
        "cfd_ell_upload": ([vp, i32p, dp, C.c_size_t, C.c_int], C.c_int),
        "cfd_spmv": ([vp, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
        "cfd_bicgstab": ([vp, C.c_double, C.c_int, ip, dp], C.c_int),
        "cfd_rbsor": ([vp, C.c_double, C.c_double, C.c_int, C.c_int, ip, dp], C.c_int),
        "cfd_sor_pipeline": ([vp, C.c_double, C.c_double, C.c_int, ip, dp], C.c_int),
        "cfd_defect_correction": ([vp, C.c_double, C.c_double, C.c_int, ip, ip, dp], C.c_int),
        "cfd_solve_pressure": ([vp, ip, dp], C.c_int),
        "cfd_predictor": ([vp], C.c_int),
        "cfd_source_add": ([vp], C.c_int),
        "cfd_build_rhs": ([vp], C.c_int),
        "cfd_pressure_scatter": ([vp], C.c_int),
        "cfd_dfib_update": ([vp], C.c_int),
        "cfd_source_accumulate": ([vp], C.c_int),
        "cfd_bc_main": ([vp], C.c_int),
        "cfd_bc_copy": ([vp], C.c_int),
        "cfd_drag": ([vp, dp, dp], C.c_int),
        "cfd_force_sum": ([vp, C.c_int, dp], C.c_int),
        "cfd_maxdiff": ([vp, dp], C.c_int),
        "cfd_copy_t3_to_t0": ([vp], C.c_int),
        "cfd_divergence": ([vp, C.c_int, dp], C.c_int),
        "cfd_step": ([vp, C.POINTER(CfdStepStats)], C.c_int),
        "cfd_loop": ([vp, ip, dp], C.c_int),
        "cfd_set_loop": ([vp, C.c_int, C.c_double], C.c_int),
        "cfd_time_kernel": ([vp, C.c_int, C.c_int, dp], C.c_int),
        "cfd_set_inflow_profile": ([vp, dp, C.c_size_t], C.c_int),
        "cfd_set_forcing_gain": ([vp, C.c_double], C.c_int),
        "cfd_adjoint_forward": ([vp, C.c_int, C.c_int, C.c_int, dp], C.c_int),
        "cfd_adjoint_backward": ([vp, dp, C.c_size_t, dp, C.POINTER(CfdAdjointStats)], C.c_int),
        "cfd_adjoint_generation": ([vp, C.POINTER(C.c_ulonglong)], C.c_int),
        "cfd_cdcl_open": ([C.c_char_p, C.c_int, C.c_char_p, C.c_int, C.POINTER(vp)], C.c_int),
        "cfd_cdcl_record": ([vp, C.c_int, C.c_double, C.c_double, C.c_double], C.c_int),
        "cfd_cdcl_close": ([vp], C.c_int),
        "cfd_plot3d_write_x": ([C.POINTER(CfdGridDesc), C.c_char_p], C.c_int),
        "cfd_plot3d_write_q": ([vp, C.c_char_p, C.c_int, C.c_double], C.c_int),
        "cfd_plot3d_read_q": ([C.c_char_p, C.POINTER(CfdGridDesc), ip, C.POINTER(C.c_float), dp, dp, dp, dp, dp],
                              C.c_int),
        "cfd_checkpoint_save": ([vp, C.c_char_p], C.c_int),
        "cfd_get_solver_counters": ([vp, C.POINTER(CfdSolverCounters)], C.c_int),
        "cfd_phase_timing": ([vp, C.c_int], C.c_int),
        "cfd_phase_times": ([vp, dp, ip], C.c_int),
        "cfd_checkpoint_load": ([vp, C.c_char_p], C.c_int),
        "cfd_bicgstab_ell_noalloc": ([vp, vp, vp, vp, C.c_size_t, C.c_double, C.c_int, vp, ip, dp], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res


class CfdError(RuntimeError):
    """Raised for a non-zero status; carries the library's message."""

    def __init__(self, status: int, where: str):
        msg = load().cfd_last_error()
        text = msg.decode() if msg else ""
        super().__init__(f"{where} failed with status {status}: {text}")
        self.status = status


class CfdInvalidArgument(CfdError, ValueError):
    """Status CFD_ERR_ARG: the reference throws std::invalid_argument here."""


def check(status: int, where: str) -> None:
    if status == CFD_OK:
        return
    if status == CFD_ERR_ARG:
        raise CfdInvalidArgument(status, where)
    raise CfdError(status, where)


def dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))
