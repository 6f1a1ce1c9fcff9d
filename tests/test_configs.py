"""Runtime configuration under the reference's option names
(configs.params_from_options / params_from_json): the controlPanel.hpp
macros, the simuClass settings of src/CFD_MX.cpp:35-57 and the BC table
select the same things in cfd_params; unknown or unbuilt options are
errors."""
from __future__ import annotations

import json

import pytest

from cfd_nssolver_forhpc_amd import capi, configs
from cfd_nssolver_forhpc_amd.solver import default_params

# src/CFD_MX.cpp:35-57 and controlPanel.hpp as shipped
REFERENCE_MAIN = {"Re": 40.0, "dt": 0.002, "DfibMethod": "DFIB_Cylinder-Z", "P_SOLVER": "P_SOLVER_BICG_SPE",
                  "CONVECTION_DIFUSSION": "CONVECTION_DIFUSSION_QUICK", "TERBULENCE_SMAGORINSKY": False,
                  "BC": "BC_Cavity_A", "method": "PC_METHOD", "p_criteria": 1e-3}


def as_bytes(p):
    return bytes(memoryview(p))


def test_no_options_is_the_default():
    assert as_bytes(configs.params_from_options({})) == as_bytes(default_params())


def test_reference_main_settings():
    p = configs.params_from_options(REFERENCE_MAIN)
    d = default_params()
    assert (p.Re, p.dt, p.tol) == (40.0, 0.002, 1e-3)
    assert p.solver == capi.CFD_SOLVER_BICGSTAB and p.op == capi.CFD_OP_SPE
    assert p.convection == capi.CFD_CONV_QUICK and p.les == 0 and p.ib_on == 1 and p.inner_loops == 3
    assert [list(p.bc_dir[f]) for f in range(6)] == [list(d.bc_dir[f]) for f in range(6)]


@pytest.mark.parametrize("opts, field, want", [
    ({"P_SOLVER": "P_SOLVER_BICG_ELL"}, "op", capi.CFD_OP_ELL),
    ({"P_SOLVER": "P_SOLVER_BICG_CSR"}, "op", capi.CFD_OP_CSR),
    ({"P_SOLVER": "P_SOLVER_SOR"}, "solver", capi.CFD_SOLVER_SOR_PIPELINE),
    ({"P_SOLVER": "RBSOR_CSR"}, "op", capi.CFD_OP_CSR),
    ({"CONVECTION_DIFUSSION": "CONVECTION_DIFUSSION_LUD"}, "convection", capi.CFD_CONV_LUD),
    ({"CONVECTION_DIFUSSION": "CONVECTION_DIFUSSION_UD"}, "convection", capi.CFD_CONV_UD),
    ({"TERBULENCE_SMAGORINSKY": True}, "les", 1),
    ({"DfibMethod": "OFF"}, "ib_on", 0),
    ({"method": "SINGLE_PASS"}, "inner_loops", 0),
    ({"p_sor_omega": 1.5}, "sor_omega", 1.5),
    ({"precision": "MIXED"}, "precision", capi.CFD_MIXED),
    ({"precond": "MG", "mg_nu": 1}, "precond", capi.CFD_PRECOND_MG),
])
def test_each_option(opts, field, want):
    assert getattr(configs.params_from_options(opts), field) == want


def test_bc_table():
    p = configs.params_from_options({"BC": "FLOWpassing"})
    assert [list(p.bc_dir[f]) for f in range(6)] == capi.FLOWPASSING_DIR
    assert [list(p.bc_num[f]) for f in range(6)] == capi.FLOWPASSING_NUM


@pytest.mark.parametrize("opts", [{"P_SOLVER": "P_SOLVER_AMGCL_BUILTIN"}, {"DfibMethod": "DFIB_Cylinder-X"},
                                  {"CONVECTION_DIFUSSION": "CENTRAL"}, {"BC": "BC_Cavity_B"},
                                  {"no_such_option": 1}])
def test_unknown_or_unbuilt_options_are_errors(opts):
    with pytest.raises(ValueError):
        configs.params_from_options(opts)


def test_json(tmp_path):
    f = tmp_path / "case.json"
    f.write_text(json.dumps(REFERENCE_MAIN))
    assert as_bytes(configs.params_from_json(f)) == as_bytes(configs.params_from_options(REFERENCE_MAIN))


@pytest.mark.gpu
def test_configured_run(require_gpu):
    """A configuration in the reference's terms runs: the matrix-free operator,
    LUD convection and the Smagorinsky model on a small cylinder case."""
    from cfd_nssolver_forhpc_amd.solver import Simulation
    w = configs.small_cylinder(ncal=(32, 16, 8), lengths=(8.0, 4.0, 2.0))
    p = configs.params_from_options({"P_SOLVER": "P_SOLVER_BICG_SPE", "CONVECTION_DIFUSSION":
                                     "CONVECTION_DIFUSSION_LUD", "TERBULENCE_SMAGORINSKY": True,
                                     "p_criteria": 1e-10}, base=w.params)
    sim = Simulation(w.grid, p)
    sim.setup(noise=w.noise, seed=w.seed)
    for _ in range(2):
        st = sim.step()
    assert max(st.relres[:st.n_passes]) < 1e-10
    sim.close()
