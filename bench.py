"""Benchmark of the prediction-correction timestep on one MI355X.

Workload: 3-D flow past a z-cylinder, 512x256x64 uniform cells over
20D x 10D x 2.5D (D = 25 cells), Re 200, dt 5e-4, uniform inflow, convective
outflow, slip y-walls, periodic z.  A step is one iteration of the
reference's main loop: 3 predictor passes and a corrector, each with its
pressure solve.

Prints one JSON line.  `value` is FP64 timesteps per second with the
reference's own pressure solve: bicgstabRe2 on the Jacobi-scaled ELL
operator, tol 1e-10, iters_max 5000.  At this size many of its passes stop
at iters_max above 1e-10; the line states how many and the worst residual.
`same_residual` compares FP64 and mixed precision where every pass reaches
the FP64 relative residual 1e-10 (multigrid-preconditioned BiCGSTAB: FP64
throughout, FP64 Krylov around an FP32 V-cycle, FP64 defect correction
around FP32 inner solves).  Reported beside them: the reference's default
matrix-free (SPE) operator, the standalone Poisson problem on the grid, the
256x128x32 discrete adjoint, the kernel table, the roofline of the kernel
the timed step spends most in (the SpMV-dot of the BiCGSTAB iteration,
timed with HIP events on the library's stream), and the CPU baseline on the
host's cores, a bounded sample of the same workload: the reference's own
bicgstabRe2 iterations (compiled from its sources) on the workload's operator,
and the non-solver part of a step from the FP64 oracle.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from cfd_nssolver_forhpc_amd import capi, configs  # noqa: E402
from cfd_nssolver_forhpc_amd.solver import Simulation  # noqa: E402

HBM_PEAK_GBS = 8000.0
try:  # the host cores this process may use, before an OpenMP runtime binds its first thread to one of them
    HOST_CPUS = os.sched_getaffinity(0)
except (AttributeError, OSError):
    HOST_CPUS = None
# algorithmic bytes per row: ELL SpMV 7 x (val + col) + x + y; SPE x + jp + y;
# BiCGSTAB iteration (5 fused kernels) 320; RB-SOR sweep 84 + x + b + x(w) + x(nbr)
# SpMV-dot (K1 of the iteration): the SpMV's 100 plus r0 8
BYTES = {"spmv_fp64": 100, "spmv_fp32": 64, "spe_fp64": 24, "bicg_iter": 320, "sor_sweep": 116,
         "copy": 16, "predictor": 48, "spmv_dot": 108}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profiled_traffic(key="spmv_dot_kernel<double, 512> traffic_bytes_per_launch:"):
    """HBM bytes per launch of a kernel from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel calibration)."""
    path = ROOT / "profiles" / "kernels_512x256x64.md"
    try:
        for line in path.read_text().splitlines():
            if line.startswith(key):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def host_cpu():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_steps(w, warmup, steps, tag):
    sim = Simulation(w.grid, w.params)
    sim.setup(noise=w.noise, seed=w.seed)
    for k in range(warmup):
        st = sim.step()
        log(f"  [{tag}] warmup step {k + 1}: iters {list(st.iters[:st.n_passes])}")
    stats = []
    sim.phase_timing(True)  # HIP events at the step's phase boundaries
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(sim.step().as_dict())
    dt = time.perf_counter() - t0
    for k, s in enumerate(stats):
        log(f"  [{tag}] step {warmup + k + 1}: iters {s['iters']} inner {s['inner_iters']} "
            f"relres {max(s['relres']):.2e} cD {s['cd'][-1]:.4f}")
    return sim, dt, stats


def phase_report(sim):
    """Device milliseconds per timed step in each phase of the PC loop (the
    reference's clockstruct stopwatches: convectionDifussion, pressure,
    updateT1toT3 with the fused checkL2norm / updateT3toT0 / getCdCl, BC)."""
    ph = sim.phase_times()
    n = max(ph["steps"], 1)
    ms = {k: v / n for k, v in ph["ms"].items()}
    total = sum(ms.values())
    return {"unit": "ms per step", "steps": ph["steps"], **ms,
            "pressure_share": ms["pressure"] / total if total > 0 else None}


def per_step(stats, key):
    return float(np.mean([sum(max(i, 0) for i in s[key]) for s in stats]))


def poisson_benchmark(w, sor_full=False):
    """Standalone Poisson solve on the 512x256x64 grid: ELL operator with
    Neumann walls and p = 0 on the outflow face, seeded zero-mean N(0,1) RHS;
    BiCGSTAB FP64 and mixed, red-black SOR (omega 1.8) capped at 2000 sweeps
    or, with sor_full, to 1e-10."""
    out = {}
    rng = np.random.default_rng(2)
    b = rng.standard_normal(w.grid.N)
    b -= b.mean()
    for tag, prec, solver in (("bicgstab_fp64", capi.CFD_FP64, 0), ("bicgstab_mixed", capi.CFD_MIXED, 0),
                              ("rbsor_fp64", capi.CFD_FP64, 1)):
        p = w.params.copy()
        p.precision, p.solver, p.maxit = prec, solver, 20000
        p.periodic_z = p.outflow_convective = 0  # Neumann walls, Dirichlet outflow face
        sim = Simulation(w.grid, p)
        sim.init_state()
        sim.createPressureMatrix()
        sim.upload(capi.RHS, b)
        sim.upload(capi.X, np.zeros(w.grid.N))
        t0 = time.perf_counter()
        if solver == 1 and sor_full:
            # to 1e-10 in chunks of 100000 sweeps (each continues from the last
            # one's x), with a progress line per chunk
            it, rel = 0, 1.0
            while rel >= 1e-10 and it < 3000000:
                k, rel = sim.rbsor(omega=1.8, tol=1e-10, maxit=100000, check_every=1000)
                it += k
                log(f"[poisson] rbsor_fp64: {it} sweeps, relres {rel:.2e}")
        elif solver == 1:
            it, rel = sim.rbsor(omega=1.8, tol=1e-10, maxit=2000, check_every=100)
        else:
            it, rel = sim.solve()
        dt = time.perf_counter() - t0
        rec = {"seconds": dt, "iters": it, "relres": rel}
        if solver == 1:
            rec["note"] = ("to 1e-10" if sor_full else
                           "capped at 2000 sweeps (omega 1.8 needs ~8e5 sweeps for 1e-10 at this size; --poisson-sor-full)")
            rec["us_per_sweep"] = dt / max(it, 1) * 1e6
        out[tag] = rec
        log(f"[poisson] {tag}: {it} iters, relres {rel:.2e}, {dt:.2f} s")
        sim.close()
    return out


# The restatement against the compiled reference in the build container (its
# own ELL BiCGSTAB loop, 190x210x40, tol 1e-3, steps 2-6, 8 OpenMP threads,
# Xeon @ 2.70 GHz): 33.7 s for 3373 iterations against the restatement's
# 40.2 s for 3249, i.e. 10.0 against 12.4 ms per iteration with its step work.
CALIBRATION = {"reference_ms_per_iteration": 10.0, "restatement_ms_per_iteration": 12.4, "threads": 8,
               "case": "190x210x40 native grid, ELL BiCGSTAB tol 1e-3, steps 2-6"}


def reference_iterations(col, val, n, nit, threads, timeout=600):
    """Milliseconds per iteration of the reference's own bicgstabRe2 on its own
    ELL_matrix (oracle/_ref/ref_bench, compiled from the reference's sources),
    on the given slot-major operator; None where the binary is absent."""
    exe = ROOT / "oracle" / "_ref" / "ref_bench"
    if not exe.exists():
        return None
    import shutil
    import tempfile
    tmp = tempfile.mkdtemp(prefix="cfd_ref_bench_")
    try:
        cf, vf = os.path.join(tmp, "col.bin"), os.path.join(tmp, "val.bin")
        col.astype(np.int32).tofile(cf)
        val.astype(np.float64).tofile(vf)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
        # the child inherits the calling thread's CPU set, which the oracle's
        # OpenMP runtime (OMP_PROC_BIND) has narrowed to one core: widen it
        # for the launch, so that the reference's threads get every core
        prev = os.sched_getaffinity(0) if HOST_CPUS else None
        try:
            if HOST_CPUS:
                os.sched_setaffinity(0, HOST_CPUS)
            out = subprocess.run([str(exe), tmp, "bench", cf, vf, str(n), str(nit)], env=env, capture_output=True,
                                 text=True, timeout=timeout)
        finally:
            if prev:
                os.sched_setaffinity(0, prev)
        for line in out.stdout.splitlines():
            if line.startswith("ref_bench"):
                f = line.split()
                rec = {f[k]: float(f[k + 1]) for k in range(1, len(f) - 1, 2)}
                return rec
        log(f"[bench] ref_bench: rc {out.returncode}: {out.stderr.strip()[-300:]}")
    except Exception as e:
        log(f"[bench] ref_bench: {e}")
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return None


def cpu_baseline(w, iters_per_step, threads, nit=300, ref_nit=600):
    """The workload on the host's cores at 512x256x64 with its boundary
    conditions, FP64: one timestep without pressure solves (the loop == 1
    step: 4 passes of predictor, source, DFIB update, BC) of the oracle, plus
    ref_nit iterations of the reference's own bicgstabRe2 (oracle/_ref/ref_bench)
    on the workload's ELL operator, and nit iterations of the oracle's
    restatement of it for comparison on the same host.  The step time is
    t_nonsolver + iters_per_step * t_iteration (iters_per_step: the GPU run's
    count, the same algorithm); t_iteration is the reference's where its
    binary runs, else the restatement's."""
    sys.path.insert(0, str(ROOT / "tests"))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import oracle_lib as O
    flags = "-O3 -march=x86-64-v3 -fopenmp"
    lib = None
    try:
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "native"], check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=120)
        lib = O.lib(native=True)
        flags = "-O3 -march=native -fopenmp"
    except Exception:
        lib = O.lib()
    O._LIB = lib
    g = O.Grid(w.grid)
    sim = O.Sim(g, w.params)
    sim.init_state(w.noise, w.seed)
    sim.setup()
    t0 = time.perf_counter()
    sim.step()  # loop 1: no pressure solve
    t_nonsolver = time.perf_counter() - t0
    del sim
    N = g.N
    col = np.zeros(7 * N, np.int32)
    val = np.zeros(7 * N)
    jp = np.zeros(N)
    lib.orc_ell_assemble_bl(g.ref, O.ints6([w.params.p_dirichlet[f] for f in range(6)]), w.params.periodic_z,
                            col, val, jp)
    b = np.random.default_rng(2).standard_normal(N)
    x = np.zeros(N)
    s = lib.orc_bicg_create(N, 0, col, val, g.ref, None, None)
    lib.orc_bicg_set_tolerance(s, 1e-300)
    it = C.c_int()
    # a short untimed solve first: the solver's work vectors are allocated
    # zeroed and their pages are touched by the first iterations (the
    # reference's solver fills its vectors when it is built, untimed)
    lib.orc_bicg_set_maxit(s, 9)
    lib.orc_bicg_solve(s, b, x, C.byref(it))
    x[:] = 0.0
    lib.orc_bicg_set_maxit(s, nit - 1)
    t0 = time.perf_counter()
    lib.orc_bicg_solve(s, b, x, C.byref(it))
    t_port = (time.perf_counter() - t0) / nit
    lib.orc_bicg_destroy(s)
    del b, x, jp
    ref = reference_iterations(col, val, N, ref_nit, threads)
    kind = "reference" if ref else "port"
    t_iter = ref["ms_per_iteration"] * 1e-3 if ref else t_port
    step_s = t_nonsolver + iters_per_step * t_iter
    it_src = (f"{int(ref['nit'])} iterations of the reference's own bicgstabRe2 on its own ELL_matrix "
              f"(oracle/_ref/ref_bench, g++ -O3 -march=x86-64-v3 -fopenmp from the reference's sources) on the "
              f"workload's operator, timed ({t_iter * 1e3:.1f} ms each; the restatement's {t_port * 1e3:.1f} ms "
              f"on the same host)") if ref else (
              f"{nit} iterations of the restatement of bicgstabRe2 on the workload's operator, timed "
              f"({t_port * 1e3:.1f} ms each)")
    return {
        "value": 1.0 / step_s,
        "unit": "timesteps/s",
        "cores": threads,
        "kind": kind,
        "sample": (f"512x256x64 with the workload's boundaries, FP64, {threads} OpenMP threads on {host_cpu()}: one "
                   f"PC-method step without solves of the oracle ({flags}), timed ({t_nonsolver:.2f} s), plus "
                   f"{it_src}; step = non-solver + {iters_per_step:.0f} iterations (the GPU run's count per step, "
                   f"the same algorithm)"),
        "t_nonsolver_s": t_nonsolver,
        "t_iteration_ms": t_iter * 1e3,
        "t_iteration_restatement_ms": t_port * 1e3,
        "reference_bench": ref,
        "host_cpu": host_cpu(),
        "calibration_build_host": CALIBRATION,
    }


def kernel_table(sim, reps):
    N = sim.N
    rows = {}
    nbytes = dict(BYTES)
    try:
        # CSR SpMV: 12 B per stored entry (value + column), the row pointers,
        # x and y (boundary rows hold fewer than 7 entries)
        nnz = len(sim.csr()[1])
        nbytes["spmv_csr"] = (12 * nnz + 4 * (N + 1) + 16 * N) / N
    except Exception as e:  # CSR operator not assembled for this context
        log(f"[bench] kernel spmv_csr: {e}")
    for what, name in ((8, "spmv_dot"), (0, "spmv_fp64"), (1, "spmv_fp32"), (2, "spe_fp64"), (7, "spmv_csr"),
                       (6, "copy"), (5, "predictor"), (4, "bicg_iter"), (3, "sor_sweep")):
        if name not in nbytes:
            continue
        try:
            ms = sim.time_kernel(what, reps if what != 4 else max(10, reps // 4))
        except Exception as e:  # an operator not assembled for this context
            log(f"[bench] kernel {name}: {e}")
            continue
        gbs = nbytes[name] * N / (ms * 1e-3) / 1e9
        rows[name] = {"avg_us": ms * 1e3, "gbs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS}
        log(f"[bench] {name:10s} {ms * 1e3:9.1f} us  {gbs:7.0f} GB/s  ({gbs / HBM_PEAK_GBS:.2f} of peak)")
    return rows


def adjoint_benchmark(steps, avg_from, every, fd_dirs, precond, tol=1e-10, h=1e-6, gain=1.0):
    """Discrete adjoint on the 256x128x32 Re-100 cylinder (dt 1e-3): J = mean
    corrector cD over steps [avg_from, steps), gradient w.r.t. the seeded
    smoothed inflow profile (128 values) and the DFIB forcing gain; timed
    objective run (with checkpoints) and reverse sweep (with recomputation);
    optionally central differences (h = 1e-6) in fd_dirs random directions,
    about the base forcing gain `gain`."""
    from cfd_nssolver_forhpc_amd.adjoint import DragProblem, smoothed_inflow
    w = configs.cylinder_256x128x32()
    w.params.precond = precond
    w.params.tol = tol
    sim = Simulation(w.grid, w.params)
    sim.init_state(w.noise, w.seed)
    sim.DFIB_CylinderZ()
    sim.createPressureMatrix()
    pr = DragProblem(sim, steps, avg_from, every, noise=w.noise, seed=w.seed)
    uin = smoothed_inflow(w.grid.ncal[1])
    log(f"[bench] adjoint {w.name}: {steps} steps, J over [{avg_from}, {steps}), checkpoint every {every}")
    t0 = time.perf_counter()
    J = pr.run(uin, gain)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    g_uin, g_gain, st = sim.adjoint_backward()
    t_bwd = time.perf_counter() - t0
    log(f"[bench] adjoint: J {J:.6f}, forward {t_fwd:.2f} s, reverse sweep {t_bwd:.2f} s")
    checks = []
    rng = np.random.default_rng(8)
    for q in range(fd_dirs):
        d = rng.standard_normal(uin.size + 1)
        d /= np.linalg.norm(d)
        Jp = pr.run(uin + h * d[:-1], gain + h * d[-1])
        Jm = pr.run(uin - h * d[:-1], gain - h * d[-1])
        fd = (Jp - Jm) / (2 * h)
        ad = float(np.dot(g_uin, d[:-1]) + g_gain * d[-1])
        checks.append({"fd": fd, "adjoint": ad, "rel_diff": abs(fd - ad) / max(abs(fd), abs(ad), 1e-300)})
        log(f"[bench] adjoint FD direction {q}: fd {fd:.9e} adjoint {ad:.9e}")
    sim.close()
    return {
        "workload": w.name, "steps": steps, "avg_from": avg_from, "checkpoint_every": every,
        "pressure": f"BiCGSTAB to {tol:g}, " + ("multigrid-preconditioned" if precond else "Jacobi-scaled") + ", FP64",
        "J_mean_cd": J, "grad_uin_norm": float(np.linalg.norm(g_uin)), "grad_gain": g_gain,
        "forward_s": t_fwd, "reverse_sweep_s": t_bwd, "reverse_over_forward": t_bwd / t_fwd,
        "forward_ms_per_step": 1e3 * t_fwd / steps, "reverse_ms_per_step": 1e3 * t_bwd / steps,
        "stats": st, "fd_step": h, "fd_base_gain": gain, "fd_checks": checks,
    }


def passes_report(stats, maxit):
    """What the timed solves reached: bicgstabRe2 stops at iters_max as the
    reference does, whatever the residual then is."""
    passes = [(it, rel) for s in stats for it, rel in zip(s["iters"], s["relres"]) if it >= 0]
    if not passes:
        return None
    return {"passes": len(passes), "passes_at_iteration_cap": sum(1 for it, _ in passes if it >= maxit),
            "iteration_cap": maxit, "max_final_relres": max(rel for _, rel in passes),
            "median_final_relres": float(np.median([rel for _, rel in passes]))}


def field_errors(a, b):
    err = {}
    for name, fid in (("u", capi.T0_U), ("v", capi.T0_V), ("w", capi.T0_W), ("p", capi.X)):
        x, y = a.download(fid), b.download(fid)
        err[name] = float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))
    return err


def same_residual_legs(steps, warmup, op=None, nu=4):
    """FP64 and mixed precision at the same FP64 relative residual 1e-10 in
    every pass: BiCGSTAB on the ELL operator right-preconditioned by the
    multigrid V-cycle (FP64; FP64 Krylov around the FP32 cycle; FP64 defect
    correction around FP32 preconditioned BiCGSTAB to 1e-5 on FP32 ELL values)."""
    out, sims = {}, {}
    for tag, prec, mode in (("fp64_mg", capi.CFD_FP64, capi.CFD_MIXED_RESTARTED),
                            ("mixed_mg", capi.CFD_MIXED, capi.CFD_MIXED_PC32),
                            ("mixed_mg_dc", capi.CFD_MIXED, capi.CFD_MIXED_RESTARTED)):
        wg = configs.cylinder_512x256x64(prec)
        wg.params.precond = capi.CFD_PRECOND_MG
        wg.params.mixed_mode = mode
        wg.params.tol_inner = 1e-5
        wg.params.maxit = 200
        wg.params.mg_nu = nu
        if op is not None:
            wg.params.op = op
        wg.name += "-mg"
        simg, tg, stg = run_steps(wg, warmup, steps, tag)
        rep = passes_report(stg, wg.params.maxit)
        out[tag] = {"timesteps_per_s": steps / tg, "ms_per_step": 1e3 * tg / steps,
                    "solver_iters_per_step": per_step(stg, "iters"),
                    "inner_iters_per_step": per_step(stg, "inner_iters"),
                    "max_final_relres": rep["max_final_relres"],
                    "all_passes_at_1e-10": rep["max_final_relres"] < 1e-10 and rep["passes_at_iteration_cap"] == 0,
                    "divergence_fluid_max": max(s["div"]["fluid_max"] for s in stg),
                    "phases": phase_report(simg)}
        sims[tag] = simg
    for tag in ("mixed_mg", "mixed_mg_dc"):
        out[tag]["rel_l2_vs_fp64_mg_after_timed_steps"] = field_errors(sims[tag], sims["fp64_mg"])
        out[tag]["speedup_vs_fp64_mg"] = out[tag]["timesteps_per_s"] / out["fp64_mg"]["timesteps_per_s"]
    for simg in sims.values():
        simg.close()
    out["residual"] = 1e-10
    out["operator"] = "matrix-free SPE" if op == capi.CFD_OP_SPE else "ELL"
    out["sweeps"] = nu
    out["note"] = ("every pass of every leg stops at ||b - Ax||/||b|| < 1e-10 in FP64 (all_passes_at_1e-10); "
                   "fp64_mg: FP64 throughout; mixed_mg: FP64 BiCGSTAB with the V-cycle in FP32; mixed_mg_dc: FP64 "
                   "defect correction around FP32 preconditioned BiCGSTAB (FP32 ELL values) to 1e-5; "
                   "speedup_vs_fp64_mg is at that common residual")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6, help="timed steps of the headline (reference-algorithm) leg")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mg-steps", type=int, default=20, help="timed steps of each same-residual leg")
    ap.add_argument("--spe-steps", type=int, default=3, help="timed steps of the matrix-free (SPE) leg (0: skip)")
    ap.add_argument("--no-spe", action="store_true")
    ap.add_argument("--no-poisson", action="store_true")
    ap.add_argument("--poisson-sor-full", action="store_true",
                    help="run the Poisson problem's red-black SOR (omega 1.8) to 1e-10 instead of 2000 sweeps "
                         "(about 8e5 sweeps, minutes)")
    ap.add_argument("--poisson-only", action="store_true", help="run only the standalone Poisson problem")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=300,
                    help="iterations of the restatement's BiCGSTAB the CPU baseline times")
    ap.add_argument("--cpu-ref-iters", type=int, default=600,
                    help="iterations of the reference's own BiCGSTAB (oracle/_ref/ref_bench) the CPU baseline times")
    ap.add_argument("--no-mg", action="store_true")
    ap.add_argument("--mg-only", action="store_true", help="run only the same-residual legs and print their JSON")
    ap.add_argument("--mg-op", choices=("ell", "spe"), default="spe",
                    help="operator of the same-residual legs' FP64 Krylov and residual (default: the matrix-free "
                         "SPE form, the reference's default operator; the defect correction's FP32 inner solves "
                         "use the FP32 ELL values either way)")
    ap.add_argument("--mg-nu", type=int, default=4, help="red-black sweeps before and after the coarse correction")
    ap.add_argument("--reps", type=int, default=200)
    ap.add_argument("--adjoint-steps", type=int, default=40,
                    help="steps of the 256x128x32 adjoint run (0: skip); the drag is averaged over the second half")
    ap.add_argument("--adjoint-every", type=int, default=20, help="checkpoint interval of the adjoint run")
    ap.add_argument("--adjoint-fd", type=int, default=0, help="central-difference directions to check the gradient")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the adjoint leg and print its JSON")
    ap.add_argument("--adjoint-tol", type=float, default=1e-10, help="pressure tolerance of the adjoint run")
    ap.add_argument("--adjoint-h", type=float, default=1e-6, help="central-difference step of the gradient check")
    ap.add_argument("--adjoint-gain", type=float, default=1.0,
                    help="base DFIB forcing gain of the adjoint run and its central differences")
    ap.add_argument("--adjoint-jacobi", action="store_true",
                    help="solve the adjoint run's pressure systems with the reference's Jacobi scaling (default: MG)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="time the operator kernels on the 512x256x64 grid and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported in this build "
              "(the hot path runs on a single MI355X; multi-GPU x-slab sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    warmup = max(args.warmup, 2)  # step 1 of the reference loop has no pressure solve
    mg_op = capi.CFD_OP_SPE if args.mg_op == "spe" else capi.CFD_OP_ELL

    if args.adjoint_only:
        steps = max(args.adjoint_steps, 2)
        pc = capi.CFD_PRECOND_JACOBI if args.adjoint_jacobi else capi.CFD_PRECOND_MG
        print(json.dumps({"adjoint": adjoint_benchmark(steps, steps // 2, args.adjoint_every, args.adjoint_fd, pc,
                                                       args.adjoint_tol, args.adjoint_h, args.adjoint_gain)}),
              flush=True)
        return

    w64 = configs.cylinder_512x256x64(capi.CFD_FP64)
    if args.poisson_only:
        print(json.dumps({"poisson": poisson_benchmark(w64, args.poisson_sor_full)}), flush=True)
        return
    if args.kernels_only:
        # operator kernels alone on the 512x256x64 grid (for counter collection)
        sim = Simulation(w64.grid, w64.params)
        sim.init_state()
        sim.createPressureMatrix()
        sim.setupPressure()
        sim.createPressureMatrixCSR()
        rng = np.random.default_rng(2)
        sim.upload(capi.X, rng.standard_normal(sim.N))
        sim.upload(capi.RHS, rng.standard_normal(sim.N))
        print(json.dumps({"kernels": kernel_table(sim, args.reps)}), flush=True)
        return
    if args.mg_only:
        print(json.dumps({"same_residual": same_residual_legs(args.mg_steps, warmup, mg_op, args.mg_nu)}), flush=True)
        return

    # ---- headline: the reference's algorithm, FP64 ----
    log(f"[bench] {w64.name}: warmup {warmup}, timed {args.steps}")
    sim64, t64, st64 = run_steps(w64, warmup, args.steps, "fp64")
    tps64 = args.steps / t64
    phases64 = phase_report(sim64)
    iters64 = per_step(st64, "iters")
    solves = passes_report(st64, w64.params.maxit)
    counters = sim64.solver_counters()
    # kernel table; the timed step spends most in the iteration's SpMV-dot kernels
    sim64.setupPressure()
    sim64.createPressureMatrixCSR()
    kt = kernel_table(sim64, args.reps)
    sim64.close()

    mg = None if args.no_mg else same_residual_legs(args.mg_steps, warmup, mg_op, args.mg_nu)

    spe = None
    if not args.no_spe and args.spe_steps > 0:
        ws = configs.cylinder_512x256x64(capi.CFD_FP64)
        ws.params.op = capi.CFD_OP_SPE
        sims, ts, sts = run_steps(ws, warmup, args.spe_steps, "spe")
        spe = {"operator": "matrix-free SPE (the reference's default P_SOLVER_BICG_SPE), bicgstabRe2 tol 1e-10, "
                           "iters_max 5000", "steps": args.spe_steps,
               "timesteps_per_s": args.spe_steps / ts, "bicgstab_iters_per_step": per_step(sts, "iters"),
               "solves": passes_report(sts, ws.params.maxit)}
        sims.close()

    poisson = None if args.no_poisson else poisson_benchmark(w64, args.poisson_sor_full)

    adjoint = None
    if args.adjoint_steps > 1:
        pc = capi.CFD_PRECOND_JACOBI if args.adjoint_jacobi else capi.CFD_PRECOND_MG
        adjoint = adjoint_benchmark(args.adjoint_steps, args.adjoint_steps // 2, args.adjoint_every, args.adjoint_fd, pc,
                                    args.adjoint_tol, args.adjoint_h)

    cpu = None
    if not args.no_cpu:
        threads = len(HOST_CPUS) if HOST_CPUS else (os.cpu_count() or 1)
        threads = max(1, min(threads, int(os.environ.get("OMP_NUM_THREADS", threads))))
        log(f"[bench] CPU baseline on {threads} threads")
        cpu = cpu_baseline(w64, iters64, threads, args.cpu_iters, args.cpu_ref_iters)
        cpu["gpu_speedup"] = tps64 / cpu["value"]  # the same algorithm on both sides

    kd = kt["spmv_dot"]
    pressure = (f"bicgstabRe2 (the reference's BiCGSTAB) on the Jacobi-scaled ELL operator, FP64, tol 1e-10, iters_max "
                f"{w64.params.maxit} as the reference: {solves['passes_at_iteration_cap']} of {solves['passes']} "
                f"timed passes stopped at iters_max, worst final relative residual {solves['max_final_relres']:.1e} "
                f"(median {solves['median_final_relres']:.1e}); every pass at 1e-10: see same_residual")
    out = {
        "metric": "timesteps/s @512x256x64 (mixed vs FP64, same residual); Poisson SpMV HBM GB/s",
        "value": tps64,
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warmup,
        "ms_per_step": 1e3 / tps64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (uniform inflow, seeded +-1e-3 noise on the initial velocity; seeded N(0,1) Poisson RHS)",
        "config": {"workload": w64.name, "grid": [512, 256, 64], "Re": 200, "dt": 5e-4,
                   "boundaries": "uniform inflow u = 1, convective outflow (U = 1, p = 0), slip y-walls, periodic z",
                   "pressure": pressure, "passes_per_step": 4, "bicgstab_iters_per_step": iters64},
        "roofline": {"bound": "hbm", "achieved": kd["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": kd["gbs"] / HBM_PEAK_GBS, "traffic": profiled_traffic(),
                     "kernel": "spmv_dot_kernel<double, 512>: K1 of the BiCGSTAB iteration, ap = A p (ELL) with the partial "
                               "ap . r0; 108 B/row x 8388608 rows per launch",
                     "avg_us": kd["avg_us"]},
        "fp64_solves": solves,
        "solver_counters": counters,
        "phases": phases64,
        "kernels": kt,
        "same_residual": mg,
        "spe": spe,
        "poisson": poisson,
        "adjoint": adjoint,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
