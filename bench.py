"""Throughput of the GPU Krylov solve on one MI355X, with the reference CPU
solver timed beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 2]

A step is one complete solve of the workload's linear system, as the
reference's Process makes it after every assembly: Prec_Form (the factor is
rebuilt, the values having been rewritten in place), ||b||, the residual and
the Krylov iterations to tolerance.  Matrix, right-hand side and work vectors
are resident in HBM before the timed region; each step restores r1 = b and
x = 0 on the device (part of the timed work).

Workload 2 (default): momentum matrix on a 128^3 hex mesh (tanh clustering
beta 2, +-10% node jitter, random cell numbering, upwind convection-diffusion
nu 1e-3, inertia V/dt), Bicg with the reference's incomplete Cholesky (prec
2) to relative residual 1e-10, FP64.

One JSON line on stdout; progress goes to stderr.
"""
import argparse
import importlib.util
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_package():
    if "tflows_amd" in sys.modules:
        return sys.modules["tflows_amd"]
    pkg = ROOT / "t-flows_amd"
    spec = importlib.util.spec_from_file_location(
        "tflows_amd", pkg / "__init__.py", submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["tflows_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def spmv_bytes(n, nnz, val_bytes=8, idx_bytes=4):
    """Algorithmic HBM bytes of one CSR SpMV: val + col per entry, row pointer,
    x read once (compulsory), y written."""
    return (val_bytes + idx_bytes) * nnz + idx_bytes * (n + 1) + 2 * val_bytes * n


def cpu_baseline(case, solve, tol, reps, x_gpu):
    """The reference's own solver (oracle/_ref, compiled from its sources with
    its FAST flags), or the C restatement if that library is absent, timed on
    this host's cores, 1 thread."""
    import numpy as np
    sys.path.insert(0, str(ROOT))
    from oracle.oracle import Oracle, Reference, reference_available
    kind = "reference" if reference_available() else "port"
    times, out = [], None
    if kind == "reference":
        R = Reference()
        R.set_matrix(case.row, case.col, case.dia, case.val)
        for _ in range(reps):
            out = R.solve(solve["method"], case.b, prec=solve["prec"], niter=solve["cap"], tol=tol)
            times.append(out["seconds"])
    else:
        O = Oracle()
        for _ in range(reps):
            t0 = time.perf_counter()
            out = O.solve(solve["method"], case.row, case.col, case.dia, case.val, case.b,
                          prec=solve["prec"], niter=solve["cap"], tol=tol)
            times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    err = float(np.linalg.norm(x_gpu - out["x"]) / np.linalg.norm(out["x"]))
    cpu = os.popen("grep -m1 'model name' /proc/cpuinfo").read().split(":")[-1].strip()
    return dict(value=1.0 / med, unit="solves/s", cores=1, kind=kind,
                sample=f"{reps} full solves of the same system (median {med:.3f} s), "
                       f"{out['niter']} iterations, 1 thread on {cpu}",
                niter=out["niter"], seconds_per_solve=med), err


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", type=int, default=2, choices=[1, 2])
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(the solver path runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    load_package()
    from tflows_amd import cases, mesh, solvers

    k = args.workload
    kw = cases.config_kwargs(k)
    solve = cases.SOLVE[k]
    log(f"generating workload {k}: {kw}")
    t0 = time.perf_counter()
    c = mesh.make_case(**kw)
    log(f"  n={c.n} nnz={c.nnz} ({time.perf_counter() - t0:.1f}s)")
    tol = solve["rtol"] * float(np.linalg.norm(c.b))
    method = solvers._METHODS[solve["method"]]

    A = solvers.Matrix(c.row, c.col, c.dia, c.val)
    info = A.info()
    b = torch.from_numpy(c.b).cuda()
    x = torch.zeros(c.n, dtype=torch.float64, device="cuda")
    r = torch.empty_like(b)
    torch.cuda.synchronize()

    def step():
        r.copy_(b)
        x.zero_()
        A.values_changed()  # Prec_Form runs inside the solve, as in the reference
        return solvers._solve(method, A, x, r, solve["prec"], solve["cap"], tol)

    for _ in range(args.warmup):
        res = step()
    torch.cuda.synchronize()
    lat = []
    lib = solvers._lib()
    lib.tfs_timing_start(0)  # HIP events around every SpMV launch of the timed region
    t_start = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        res = step()
        lat.append(time.perf_counter() - ts)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    import ctypes
    spmv_ms, spmv_n = ctypes.c_double(), ctypes.c_longlong()
    lib.tfs_timing_stop(ctypes.byref(spmv_ms), ctypes.byref(spmv_n))
    niter, ini_res, fin_res = res
    ms_per_step = 1000.0 * elapsed / args.steps
    value = args.steps / elapsed
    lat_ms = np.array(lat) * 1000.0
    log(f"  {value:.2f} solves/s, {ms_per_step:.3f} ms/solve, niter={niter}, "
        f"fin_res={fin_res:.3e} (tol {tol:.3e})")

    # Per-class breakdown from one instrumented solve per class (outside the timed region).
    breakdown = {}
    for kc, name in ((0, "spmv"), (1, "sweeps"), (2, "factor")):
        lib.tfs_timing_start(kc)
        step()
        ms, nl = ctypes.c_double(), ctypes.c_longlong()
        lib.tfs_timing_stop(ctypes.byref(ms), ctypes.byref(nl))
        breakdown[name] = dict(ms_per_solve=ms.value, launches_per_solve=nl.value)

    # Dominant streaming kernel: the SpMV.  Back-to-back launches give the
    # clean per-launch duration that a kernel trace reports.
    y = torch.empty_like(b)
    spmv_b2b_ms = A.spmv_time(x, y, reps=50)
    algo = spmv_bytes(c.n, c.nnz)
    spmv_inloop_ms = spmv_ms.value / max(spmv_n.value, 1)
    achieved = algo / (spmv_b2b_ms * 1e-3) / 1e9
    log(f"  spmv {spmv_b2b_ms * 1e3:.1f} us back-to-back ({achieved:.0f} GB/s), "
        f"{spmv_inloop_ms * 1e3:.1f} us inside solves; breakdown {breakdown}")

    x_gpu = x.cpu().numpy()
    cpu, err_vs_cpu = (None, None)
    if not args.no_cpu:
        log(f"timing the CPU baseline ({args.cpu_reps} solves)...")
        cpu, err_vs_cpu = cpu_baseline(c, solve, tol, args.cpu_reps, x_gpu)
        log(f"  cpu {cpu['value']:.3f} solves/s ({cpu['kind']}), rel-L2(x_gpu - x_cpu) "
            f"= {err_vs_cpu:.2e}")

    out = {
        "metric": "Poisson solves/s & SpMV HBM GB/s vs plain FP64 at 1e-10 resid; p50/p99 latency",
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded hex mesh and RHS)",
        "config": {
            "workload": f"config {k}: " + ("momentum 128^3 hex (tanh beta 2, 10% jitter, "
                                            "permuted), Bicg + DIC (prec 2) to rel 1e-10"
                                            if k == 2 else "Poisson 32^3, Cg + DIC to rel 1e-8"),
            "n_cells": c.n, "nnz": c.nnz, "method": solve["method"], "prec": solve["prec"],
            "rtol": solve["rtol"], "niter": niter, "fin_res": fin_res, "tol": tol,
            "levels_lower": info["levels_lower"], "levels_upper": info["levels_upper"],
        },
        "latency_ms": {"p50": float(np.percentile(lat_ms, 50)),
                       "p99": float(np.percentile(lat_ms, 99))},
        "roofline": {
            "kernel": "k_spmv (FP64 CSR SpMV)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes": algo,
            "launch_us": spmv_b2b_ms * 1e3,
            "launch_us_in_solves": spmv_inloop_ms * 1e3,
        },
        "breakdown_ms_per_solve": {k2: v["ms_per_solve"] for k2, v in breakdown.items()},
        "cpu_baseline": cpu,
        "rel_l2_vs_cpu": err_vs_cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
