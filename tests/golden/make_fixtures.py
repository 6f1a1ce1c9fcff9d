"""Writes tests/golden/native_fixtures.json from the reference solver itself.

Builds oracle/_ref (the reference's own header-only sources compiled where
they lie, see oracle/Makefile), runs its harness at the reference's native
case (190x210x40 SML grid, cavity table, Re 40, dt 0.002) and records:
  * SHA-256 of the grid metrics, the IB mask, the ELL operator (columns,
    values), gA.jp, the SPE Jacobi scaling and the CSR operator (row
    pointers, columns, values and its Jacobi scaling);
  * the predictor with the LUD and UD convection schemes, and with the
    Smagorinsky eddy viscosity (which it also records), on the same seeded
    velocities (ref_harness_lud / _ud / _les);
  * the output writers (Plot3D grid and solution files, Time_cDcL.dat in
    both of its modes) and the Plot3D reader on seeded fields;
  * for every hot-path function applied once to seeded inputs (splitmix64
    noise, reproduced bit for bit by the oracle's orc_fill_noise): SHA-256 of
    its output, the L2 norm, and 256 seeded sample values;
  * drag/lift, max-difference, BiCGSTAB (iterations, residual) and
    SorPipeLine_seq scalars;
  * bicgstabRe2's failure handling on a 200-row tridiagonal system in the
    reference's own ELL_matrix: the NaN / Inf reset after iters_max (which
    leaves px2) and the shadow-residual restart after the tenth solve;
  * the per-pass statistics of three PC-method timesteps at the reference
    tolerance 1e-3;
  * whole trajectories of the native case with the reference's default
    matrix-free (SPE) operator: 50 steps at its tolerance 1e-3 (serially and
    with 8 OpenMP threads, whose different dot-product order alone moves the
    loosely stopped solves), and 20 steps
    at 1e-8 with the source term added at its own cell (per pass:
    iterations, residual, cD, cL; per step: max-diff).  With the driver's
    icelCal source index the right-hand side is inconsistent with the
    singular all-Neumann operator and the reference's own solve diverges
    below ~1e-4, so the tight run needs the corrected index.  It uses 8
    OpenMP threads (an hour serially); its values then repeat to ~1e-12
    relative, far inside the bars of the tests that use them.
Only this data is committed; reference sources and binaries never are.
usage: python tests/golden/make_fixtures.py [--reuse-dumps]  (needs /root/reference;
       --reuse-dumps keeps trajectory dumps already under oracle/_ref/dumps)
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
ORACLE = ROOT / "oracle"
DUMPS = ORACLE / "_ref" / "dumps"
OUT = Path(__file__).resolve().parent / "native_fixtures.json"


def sha(path: Path) -> str:
    return hashlib.sha256(path.read_bytes()).hexdigest()


def samples(a: np.ndarray, n=256, seed=11):
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(a.size, size=min(n, a.size), replace=False))
    return {"idx": idx.tolist(), "val": [float(v) for v in a[idx]]}


def main():
    if not Path("/root/reference/inc").is_dir():
        sys.exit("needs the reference sources at /root/reference")
    subprocess.run(["make", "-s", "-C", str(ORACLE), "ref"], check=True)
    # one thread: the reference's BC face loops (updateBC.hpp:14-81) are
    # `omp for ... nowait`, so with several threads the edge cells that two
    # face loops write race and its BiCGSTAB reductions change with the
    # partition; serially every dump is reproducible and the BC follows the
    # x -> y -> z face order its code spells out
    env = dict(os.environ, OMP_NUM_THREADS="1")
    h = str(ORACLE / "_ref" / "ref_harness")
    DUMPS.mkdir(parents=True, exist_ok=True)
    subprocess.run([h, str(DUMPS / "native"), "3", "1e-3", "ell", "2"], check=True, env=env,
                   stdout=subprocess.DEVNULL)
    subprocess.run([h, str(DUMPS / "synth"), "synth"], check=True, env=env, stdout=subprocess.DEVNULL)
    subprocess.run([h, str(DUMPS / "io"), "io"], check=True, env=env, stdout=subprocess.DEVNULL)
    # the predictor with the reference's other convection schemes and with its
    # Smagorinsky model
    for scheme in ("lud", "ud", "les"):
        subprocess.run([h + "_" + scheme, str(DUMPS / f"synth_{scheme}"), "synth"], check=True, env=env,
                       stdout=subprocess.DEVNULL)
    # bicgstabRe2's NaN reset and shadow-residual restart on a small system
    subprocess.run([h, str(DUMPS / "rules"), "rules"], check=True, env=env, stdout=subprocess.DEVNULL)
    reuse = "--reuse-dumps" in sys.argv
    # (steps, tolerance, OpenMP threads, source term at its own cell)
    # spe_1e-3_omp8 repeats spe_1e-3 with 8 OpenMP threads: its dot products
    # sum in another order, which measures how far the reference's own 1e-3
    # solves move with the summation order alone
    trajs = {"spe_1e-3": ("50", "1e-3", "1", "0"), "spe_1e-8_cell": ("20", "1e-8", "8", "1"),
             "spe_1e-3_omp8": ("50", "1e-3", "8", "0")}
    for key, (nsteps, tol, threads, at_cell) in trajs.items():
        out = DUMPS / f"traj_{key}"
        done = (out / "stats.txt").exists() and sum(
            1 for line in (out / "stats.txt").read_text().splitlines() if line.startswith("maxdiff")) == int(nsteps)
        if not (reuse and done):
            subprocess.run([h, str(out), nsteps, tol, "spe", "0", "0", at_cell], check=True,
                           env=dict(os.environ, OMP_NUM_THREADS=threads), stdout=subprocess.DEVNULL)
    nat, syn, io = DUMPS / "native", DUMPS / "synth", DUMPS / "io"
    fx = {"case": "native 190x210x40 SML, BC_Cavity_A, Re 40, dt 0.002, cylinder (6.5, 6) R 0.5",
          "seed": 20240611, "setup": {}, "synth": {}, "scalars": {}, "steps": []}
    for name in ("grid.bin", "eta.bin", "ell_col.bin", "ell_val.bin", "jp.bin", "spe_jp.bin", "csr_ptr.bin",
                 "csr_idx.bin", "csr_val.bin", "csr_jp.bin"):
        fx["setup"][name] = sha(nat / name)
    eta = np.fromfile(nat / "eta.bin")
    fx["setup"]["eta_sum"] = float(eta.sum())
    fx["setup"]["eta_cut_cells"] = int(((eta > 0) & (eta < 1)).sum())
    for f in sorted(syn.glob("syn_*.bin")):
        a = np.fromfile(f)
        fx["synth"][f.name] = {"sha256": sha(f), "l2": float(np.linalg.norm(a)), **samples(a)}
    for line in (syn / "syn_scalars.txt").read_text().splitlines():
        k, *v = line.split()
        fx["scalars"][k] = [float(x) for x in v]
    for line in (nat / "stats.txt").read_text().splitlines():
        k, *v = line.split()
        if k == "pass":
            fx["steps"].append({"step": int(v[0]), "pass": int(v[1]), "iters": int(v[2]), "err": float(v[3]),
                                "cd": float(v[4]), "cl": float(v[5])})
        elif k == "maxdiff":
            fx["steps"].append({"step": int(v[0]), "maxdiff": [float(x) for x in v[1:]]})
    for scheme in ("lud", "ud", "les"):
        fx[f"synth_{scheme}"] = {}
        for name in [f"syn_pred_T1_{n}.bin" for n in "uvw"] + (["syn_les_viseff.bin"] if scheme == "les" else []):
            f = DUMPS / f"synth_{scheme}" / name
            a = np.fromfile(f)
            fx[f"synth_{scheme}"][f.name] = {"sha256": sha(f), "l2": float(np.linalg.norm(a)), **samples(a)}
    # output files: SHA-256 of the Plot3D files and of what QfileRead returns,
    # the reader's header, and the two Time_cDcL.dat texts verbatim
    fx["io"] = {"P3D.x": sha(io / "mx_out" / "P3D.x"), "P3D01.q": sha(io / "mx_out" / "P3D01.q"),
                "read_header": (io / "read_header.txt").read_text().split(),
                "cdcl_delay": (io / "cdcl_delay.dat").read_text(),
                "cdcl_nodelay": (io / "cdcl_nodelay.dat").read_text()}
    for q in "p u v w eta".split():
        fx["io"][f"read_{q}"] = sha(io / f"read_{q}.bin")
    fx["rules"] = {}
    for line in (DUMPS / "rules" / "rules.txt").read_text().splitlines():
        k, *v = line.split()
        if k == "nan":
            fx["rules"]["nan"] = {"iters": [int(v[0]), int(v[2]), int(v[4])], "relres": [float(v[1]), float(v[3])],
                                  "relres3_is_nan": v[5] == "nan", "x3_equals_x1": v[6] == "1"}
        elif k == "restart":
            fx["rules"]["restart"] = {"iters": int(v[0]), "relres": float(v[1]), "iters_fresh": int(v[2]),
                                      "relres_fresh": float(v[3]), "restarts": int(v[4]),
                                      "restarts_fresh": int(v[5])}
        else:
            fx["rules"][k] = [float(x) for x in v]
    fx["trajectories"] = {}
    for key, (nsteps, tol, threads, at_cell) in trajs.items():
        t = {"op": "spe", "tol": float(tol), "steps": int(nsteps), "omp_threads": int(threads),
             "source_at_cell": int(at_cell), "pass": [], "maxdiff": []}
        for line in (DUMPS / f"traj_{key}" / "stats.txt").read_text().splitlines():
            k, *v = line.split()
            if k == "pass":
                t["pass"].append([int(v[0]), int(v[1]), int(v[2]), float(v[3]), float(v[4]), float(v[5])])
            elif k == "maxdiff":
                t["maxdiff"].append([int(v[0])] + [float(x) for x in v[1:]])
        fx["trajectories"][key] = t
    OUT.write_text(json.dumps(fx, indent=0))
    print(f"wrote {OUT} ({OUT.stat().st_size // 1024} KiB)")


if __name__ == "__main__":
    main()
