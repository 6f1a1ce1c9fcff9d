"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer.

* The oracle: plain C with OpenMP and a great deal of index arithmetic over
  ghosted and interior arrays.  Its tests (reference golden vectors and
  closed forms) are run again, in a child process, against
  `make -C oracle sanitize` with gcc's ASan runtime preloaded into the
  interpreter.
* The library's host-side entry points that need no GPU (the Plot3D writer
  and reader, which parses sizes from the file, the drag-history log, the
  grid axes, the shim's compile check): their tests are run against
  `build.sh sanitize` (host code instrumented, device code unchanged) with
  clang's ASan runtime preloaded.

Any out-of-bounds access, use after free or undefined operation aborts the
child and fails the test.  The ASan runtime must be the first library
loaded, so where the environment already preloads one the tests are skipped.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "cfd_nssolver_forhpc_amd"
OPTS = {"ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1", "UBSAN_OPTIONS": "print_stacktrace=1:halt_on_error=1"}


def preload_blocked():
    return bool(os.environ.get("LD_PRELOAD")) or Path("/etc/ld.so.preload").exists()


def run_child(tests, env_extra):
    env = dict(os.environ, **OPTS, **env_extra)
    r = subprocess.run([sys.executable, "-m", "pytest", *tests, "-x", "-q", "-m", "not gpu", "-p", "no:cacheprovider"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]


def test_oracle_under_asan_and_ubsan():
    if preload_blocked():
        pytest.skip("the environment preloads a library, so the ASan runtime cannot be loaded first")
    try:
        asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                              check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        pytest.skip("gcc is not available")
    if not Path(asan).is_absolute():
        pytest.skip("gcc's libasan is not available")
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "sanitize"], check=True)
    run_child(["tests/test_oracle_closed_forms.py", "tests/test_oracle_golden.py"],
              {"LD_PRELOAD": asan, "CFD_ORACLE_LIB": str(ROOT / "oracle" / "_build" / "liboracle_san.so")})


def test_library_host_code_under_asan_and_ubsan():
    if preload_blocked():
        pytest.skip("the environment preloads a library, so the ASan runtime cannot be loaded first")
    runtimes = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not runtimes or not Path("/opt/rocm/bin/hipcc").exists():
        pytest.skip("hipcc or clang's ASan runtime is not available")
    lib = PKG / "_lib" / "libcfd_hip_san.so"
    sources = list((PKG / "csrc").iterdir()) + list((ROOT / "include").glob("*.h"))
    if not lib.exists() or lib.stat().st_mtime < max(p.stat().st_mtime for p in sources):
        subprocess.run(["sh", str(PKG / "build.sh"), "sanitize"], check=True, capture_output=True)
    run_child(["tests/test_io_golden.py", "tests/test_abi.py"], {"LD_PRELOAD": runtimes[-1], "CFD_HIP_LIB": str(lib)})
