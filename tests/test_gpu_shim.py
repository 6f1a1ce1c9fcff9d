"""The C++ reference-interface shim (include/cfd_reference_shim.hpp) drives
the reference's PC loop with host vectors; it must reproduce the fused
device-resident timestep."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

pytestmark = pytest.mark.gpu


def test_cpp_shim_driver(require_gpu, tmp_path):
    exe = tmp_path / "test_shim"
    lib = ROOT / "cfd_nssolver_forhpc_amd" / "_lib"
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", str(ROOT / "include"), str(ROOT / "tests/cpp/test_shim.cpp"),
                    "-L", str(lib), "-lcfd_hip", f"-Wl,-rpath,{lib}", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "PASS" in r.stdout, r.stdout + r.stderr
