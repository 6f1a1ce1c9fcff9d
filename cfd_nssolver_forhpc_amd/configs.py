"""Workload definitions: the reference's native case and the five cylinder
configurations (128x64x4 CPU case, 512x256x64 FP64 and mixed, 256x128x32
adjoint case, 1024x512x128 stretched case)."""
from __future__ import annotations

from dataclasses import dataclass

from . import capi
from .capi import CfdParams, GridArrays
from .solver import axis_tanh, default_params, native_grid, uniform_grid
from .capi import GridArrays as _GA


@dataclass
class Workload:
    name: str
    grid: GridArrays
    params: CfdParams
    steps: int
    noise: float
    seed: int


def reference_native() -> Workload:
    """The case the reference compiles in: 190x210x40 SML grid, cavity
    table, cylinder (6.5, 6) R 0.5, Re 40, dt 0.002, BiCGSTAB to 1e-3."""
    return Workload("native-190x210x40", native_grid(), default_params(), 20000, 0.0, 0)


def cylinder_params(Re: float, dt: float, cx: float, cy: float, D: float, tol: float,
                    precision: int = capi.CFD_FP64) -> CfdParams:
    """Flow past a z-cylinder: uniform inflow u = 1 on x-, zero-gradient
    outflow on x+, slip walls on y and z; pressure Neumann except p = 0 on
    the outflow face; source term added at the cell it belongs to."""
    p = default_params()
    p.Re = Re
    p.dt = dt
    p.set_bc(capi.CHANNEL_SLIP_DIR, capi.CHANNEL_SLIP_NUM)
    for f in range(6):
        p.p_dirichlet[f] = 0
    p.p_dirichlet[1] = 1
    p.source_index_quirk = 0
    p.inner_loops = 3
    p.solver = capi.CFD_SOLVER_BICGSTAB
    p.op = capi.CFD_OP_ELL
    p.precision = precision
    p.tol = tol
    p.tol_inner = 1e-5
    # FP32 operator values, FP64 Krylov vectors, FP64 residual replacement:
    # the scheme that stays convergent at 8M+ unknowns (FP32 Krylov vectors
    # stall near 3e-4 and diverge there; see the mixed-precision notes)
    p.mixed_mode = capi.CFD_MIXED_VALUES32
    p.mixed_delta = 0.1
    p.maxit = 5000
    p.ib_on = 1
    p.ib_cx, p.ib_cy, p.ib_r = cx, cy, 0.5 * D
    p.ib_nsub = 100
    p.u_init[0], p.u_init[1], p.u_init[2] = 1.0, 0.0, 0.0
    return p


def outflow_periodic_bc(p: CfdParams, speed: float = 1.0) -> CfdParams:
    """The cylinder workloads' boundaries: convective outflow on x+ at the
    inflow speed (pressure p = 0 there) and periodic z; uniform inflow and
    slip y-walls stay in the table."""
    p.periodic_z = 1
    p.outflow_convective = 1
    p.outflow_speed = speed
    return p


def cylinder_128x64x4() -> Workload:
    """128x64x4 over 20D x 10D x 0.5D, D = 1 at (5, 5), Re 100, dt 0.002, 100 steps, FP64, tol 1e-10;
    convective outflow, slip y-walls, periodic z."""
    g = uniform_grid((128, 64, 4), (20.0, 10.0, 0.5))
    return Workload("cyl-128x64x4-Re100", g,
                    outflow_periodic_bc(cylinder_params(100.0, 0.002, 5.0, 5.0, 1.0, 1e-10)), 100, 1e-3, 1)


def cylinder_512x256x64(precision: int = capi.CFD_FP64) -> Workload:
    """512x256x64 over 20 x 10 x 2.5 (D = 25 cells), Re 200, dt 5e-4, 200 steps, tol 1e-10;
    convective outflow, slip y-walls, periodic z."""
    g = uniform_grid((512, 256, 64), (20.0, 10.0, 2.5))
    name = "cyl-512x256x64-Re200" + ("-mixed" if precision == capi.CFD_MIXED else "-fp64")
    return Workload(name, g, outflow_periodic_bc(cylinder_params(200.0, 5e-4, 5.0, 5.0, 1.0, 1e-10, precision)),
                    200, 1e-3, 1)


def cylinder_256x128x32() -> Workload:
    """256x128x32 over 20 x 10 x 2.5, Re 100, dt 1e-3, 400 steps (adjoint case); convective outflow,
    slip y-walls, periodic z."""
    g = uniform_grid((256, 128, 32), (20.0, 10.0, 2.5))
    return Workload("cyl-256x128x32-Re100", g, outflow_periodic_bc(cylinder_params(100.0, 1e-3, 5.0, 5.0, 1.0, 1e-10)),
                    400, 1e-3, 1)


def cylinder_1024x512x128(precision: int = capi.CFD_MIXED) -> Workload:
    """1024x512x128, x/y stretched (growth 1.05) around a D = 40-cell cylinder, Re 300, dt 2.5e-4, 100 steps."""
    D = 1.0
    h0 = D / 40.0
    ax = axis_tanh(1024, 40.0, 10.0, h0, 4.0 * D, 1.05)
    ay = axis_tanh(512, 20.0, 10.0, h0, 4.0 * D, 1.05)
    from .solver import axis_uniform
    az = axis_uniform(128, 128 * h0)
    g = _GA((1024, 512, 128), ax[0], ay[0], az[0], ax[1], ay[1], az[1], ax[2], ay[2], az[2], 40.0, 20.0, 128 * h0)
    return Workload("cyl-1024x512x128-Re300", g,
                    outflow_periodic_bc(cylinder_params(300.0, 2.5e-4, 10.0, 10.0, D, 1e-10, precision)), 100, 1e-3, 1)


def small_cylinder(ncal=(32, 16, 8), lengths=(8.0, 4.0, 2.0), quirk=0, tol=1e-10, inner=3) -> Workload:
    """A few-thousand-cell cylinder case for parity tests."""
    g = uniform_grid(ncal, lengths)
    p = cylinder_params(100.0, 0.002, lengths[0] / 4.0, lengths[1] / 2.0, 1.0, tol)
    p.source_index_quirk = quirk
    p.inner_loops = inner
    return Workload("small", g, p, 10, 1e-3, 7)


# ---------------------------------------------------------------------------
# runtime configuration under the reference's option names
# ---------------------------------------------------------------------------
# The reference fixes its configuration at compile time (inc/controlPanel.hpp
# macros, the BC table #included in Boundary_Condition.hpp:9-11) and in main
# (simuClass setters, src/CFD_MX.cpp:35-57).  Here the same names select the
# same things in cfd_params at run time.

_P_SOLVER = {  # controlPanel.hpp:17-29
    "P_SOLVER_BICG_ELL": (capi.CFD_SOLVER_BICGSTAB, capi.CFD_OP_ELL),
    "P_SOLVER_BICG_SPE": (capi.CFD_SOLVER_BICGSTAB, capi.CFD_OP_SPE),
    "P_SOLVER_BICG_CSR": (capi.CFD_SOLVER_BICGSTAB, capi.CFD_OP_CSR),
    # SorPipeLine's volume-weighted update on the ghosted pressure with its
    # max |dp| stop; the sweep order is red-black instead of lexicographic
    # (same fixed point, race-free on the device)
    "P_SOLVER_SOR": (capi.CFD_SOLVER_SOR_PIPELINE, capi.CFD_OP_ELL),
    # this build's red-black SOR on the Jacobi-scaled ELL / CSR operator
    # with a relative-residual stop (no reference counterpart)
    "RBSOR_ELL": (capi.CFD_SOLVER_RBSOR, capi.CFD_OP_ELL),
    "RBSOR_CSR": (capi.CFD_SOLVER_RBSOR, capi.CFD_OP_CSR),
}
_CONVECTION = {  # controlPanel.hpp:125-129
    "CONVECTION_DIFUSSION_QUICK": capi.CFD_CONV_QUICK,
    "CONVECTION_DIFUSSION_LUD": capi.CFD_CONV_LUD,
    "CONVECTION_DIFUSSION_UD": capi.CFD_CONV_UD,
}
_BC_TABLES = {  # the table headers Boundary_Condition.hpp includes
    "BC_Cavity_A": (capi.CAVITY_A_DIR, capi.CAVITY_A_NUM),
    "FLOWpassing": (capi.FLOWPASSING_DIR, capi.FLOWPASSING_NUM),
}
_METHOD = {"PC_METHOD": 3, "SINGLE_PASS": 0}  # runSeqOmp_predictionMethod / runSeqOmp


def params_from_options(options: dict, base: CfdParams | None = None) -> CfdParams:
    """cfd_params from the reference's option names:

    * simuClass: Re, dt (set_Re / set_dt), p_criteria, p_sor_omega,
      DfibMethod ("DFIB_Cylinder-Z" or "OFF"), with the cylinder as
      Dfib centre ib_cx, ib_cy and radius ib_r;
    * controlPanel.hpp: P_SOLVER (P_SOLVER_BICG_ELL / _SPE / _CSR,
      P_SOLVER_SOR), CONVECTION_DIFUSSION (CONVECTION_DIFUSSION_QUICK / _LUD
      / _UD), TERBULENCE_SMAGORINSKY (true / false);
    * the BC table, BC ("BC_Cavity_A" or "FLOWpassing"), and the driver,
      method ("PC_METHOD": 3 predictor passes and a corrector, or
      "SINGLE_PASS");
    * the additions: precision ("FP64" / "MIXED"), precond ("JACOBI" / "MG"),
      mg_nu, tol_inner, maxit.

    Unknown names and values are errors, as are options this build does not
    have (other IB shapes, the AMGCL and Eigen solvers)."""
    p = (base or default_params()).copy()
    for key, val in options.items():
        if key == "Re":
            p.Re = float(val)
        elif key == "dt":
            p.dt = float(val)
        elif key == "p_criteria":
            p.tol = float(val)
        elif key == "p_sor_omega":
            p.sor_omega = float(val)
        elif key == "DfibMethod":
            if val not in ("DFIB_Cylinder-Z", "OFF"):
                raise ValueError(f"DfibMethod {val!r}: only DFIB_Cylinder-Z and OFF are built")
            p.ib_on = 1 if val == "DFIB_Cylinder-Z" else 0
        elif key in ("ib_cx", "ib_cy", "ib_r"):
            setattr(p, key, float(val))
        elif key == "P_SOLVER":
            if val not in _P_SOLVER:
                raise ValueError(f"P_SOLVER {val!r}: one of {sorted(_P_SOLVER)}")
            p.solver, p.op = _P_SOLVER[val]
        elif key == "CONVECTION_DIFUSSION":
            if val not in _CONVECTION:
                raise ValueError(f"CONVECTION_DIFUSSION {val!r}: one of {sorted(_CONVECTION)}")
            p.convection = _CONVECTION[val]
        elif key == "TERBULENCE_SMAGORINSKY":
            p.les = 1 if bool(val) else 0
        elif key == "BC":
            if val not in _BC_TABLES:
                raise ValueError(f"BC {val!r}: one of {sorted(_BC_TABLES)}")
            p.set_bc(*_BC_TABLES[val])
        elif key == "method":
            if val not in _METHOD:
                raise ValueError(f"method {val!r}: one of {sorted(_METHOD)}")
            p.inner_loops = _METHOD[val]
        elif key == "precision":
            p.precision = {"FP64": capi.CFD_FP64, "MIXED": capi.CFD_MIXED}[val]
        elif key == "precond":
            p.precond = {"JACOBI": capi.CFD_PRECOND_JACOBI, "MG": capi.CFD_PRECOND_MG}[val]
        elif key in ("mg_nu", "maxit"):
            setattr(p, key, int(val))
        elif key == "tol_inner":
            p.tol_inner = float(val)
        else:
            raise ValueError(f"unknown option {key!r}")
    return p


def params_from_json(path) -> CfdParams:
    """params_from_options of a JSON object read from path."""
    import json
    from pathlib import Path
    return params_from_options(json.loads(Path(path).read_text()))
