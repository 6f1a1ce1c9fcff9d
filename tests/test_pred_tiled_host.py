"""The LDS-tiled QUICK predictor's phases on the host, bit for bit against the
oracle.

`tests/cpp/pred_tiled_host.cpp` includes the kernel's own phase functions
(`csrc/pred_tiled.h`, `__host__ __device__`) and runs each workgroup one
thread at a time between the kernel's barriers, with the LDS poisoned to
NaN at every workgroup.  Built with contraction off, like the oracle, the
kernel's arithmetic must reproduce the oracle's predictor exactly in every
cell it writes, and write no other cell: 44 grids covering every patch depth
(4 ... 64), patches and plane ranges the grid does not fill, two patches in
z, stretched x / y axes, closed and periodic z, the eddy viscosity, and the
kept prediction with the source added after it.  A wrong index, halo or
table entry, or a read of LDS no phase wrote, is a mismatch here.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

import oracle_lib as O

ROOT = Path(__file__).resolve().parents[1]
HIPCC = Path("/opt/rocm/bin/hipcc")


def test_tiled_predictor_phases_match_oracle_bit_for_bit(tmp_path):
    if not HIPCC.exists():
        pytest.skip("hipcc is not available")
    O.lib()  # builds the oracle if needed
    exe = tmp_path / "pred_tiled_host"
    subprocess.run([str(HIPCC), "-x", "hip", "--offload-host-only", "-std=c++17", "-O2", "-ffp-contract=off",
                    "-I", str(ROOT / "include"), str(ROOT / "tests" / "cpp" / "pred_tiled_host.cpp"), "-o", str(exe),
                    "-ldl"], check=True, capture_output=True, text=True, timeout=300)
    lib = ROOT / "oracle" / "_build" / "liboracle.so"
    r = subprocess.run([str(exe), str(lib)], capture_output=True, text=True, timeout=600)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "all cases bit-identical to the oracle" in r.stdout
    assert r.stdout.count("mismatches 0 ") == 44
