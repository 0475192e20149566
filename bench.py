#!/usr/bin/env python3
"""Benchmark of the grid transport solver on one MI355X.

Workload (default): one 1024 x 1024 solve to the 1e-7 residual threshold on
the seeded smoothed-shapes density (3 discs/rings, Gaussian blur 5 px, floor
1e-3, mean 1), conjugate-Jacobian beta, max_iter 1000.  A step is one full
solve() -- density already resident in HBM, forward vertex positions out --
excluding init(n), which is reported separately, as the reference does.

Prints ONE JSON line: value = ms per solve of the headline precision mode
(mixed when available, else FP64), plus the FP64 figure, the roofline of the
dominant kernel (measured live with HIP events on the solver's stream) and
the CPU baseline (the compiled reference, 1 thread, one solve of the same
density).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/solve to 1e-7 residual at 1024² (BF16-split vs FP64 GPU); batched solves/s"
PEAK_FP64_MFMA_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--profile-reps", type=int, default=20)
    return ap.parse_args()


def cpu_baseline(n, density):
    """The compiled reference (oracle/_ref/libotref.so), one solve, 1 thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        from oracle import Reference, reference_available
    except Exception as e:  # pragma: no cover
        return {"value": None, "unit": "ms/solve", "cores": 1, "kind": "reference", "sample": f"unavailable: {e}"}
    if not reference_available():
        return {"value": None, "unit": "ms/solve", "cores": 1, "kind": "reference",
                "sample": "oracle/_ref/libotref.so not built"}
    R = Reference()
    r = R.solve(n, density, traced=False)
    return {"value": round(r["t_solve"] * 1e3, 1), "unit": "ms/solve", "cores": 1, "kind": "reference",
            "sample": f"one full solve of the same {n}x{n} density (init {r['t_init']:.2f} s excluded), "
                      f"reference built -O3 -DENABLE_SSE_MODE, SimplicialLDLT (no CHOLMOD), 1 thread"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); "
              "the batched and single solves are measured per GPU", file=sys.stderr)
        sys.exit(2)
    import torch

    from otmap_amd import GridBasedTransportSolver, Precision, SolverOptions
    from otmap_amd import _lib
    from otmap_amd import densities as D

    n = args.n
    F, V = n * n, (n + 1) ** 2
    rho = D.smoothed_shapes(n) if n == 1024 else D.smoothed_shapes(n, sigma_px=5.0 * n / 1024)
    dev = torch.device("cuda:0")
    d_rho = torch.from_numpy(rho).to(dev)
    d_fwd = torch.empty(2 * V, dtype=torch.float64, device=dev)
    d_adj = torch.empty(F, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    s = GridBasedTransportSolver(0)
    s.set_verbose_level(0)
    t0 = time.perf_counter()
    s.init(n)
    t_init = time.perf_counter() - t0

    results = {}
    modes = [("fp64", Precision.FP64, 1)]
    if n % 64 == 0:
        modes.append(("mixed", Precision.MIXED, 1))
    for name, prec, refine in modes:
        opt = SolverOptions(precision=prec, refine_steps=refine)
        for _ in range(args.warmup):
            s.solve_device(d_rho.data_ptr(), d_fwd.data_ptr(), d_adj.data_ptr(), opt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stats = []
        for _ in range(args.steps):
            stats.append(s.solve_device(d_rho.data_ptr(), d_fwd.data_ptr(), d_adj.data_ptr(), opt))
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3 / args.steps
        fwd = d_fwd.cpu().numpy()
        results[name] = dict(ms=wall, iters=stats[-1].iterations, res=stats[-1].final_residual, fwd=fwd,
                             dev_ms=float(np.mean([st.ms_solve for st in stats])))

    # parity against the reference's golden positions when they exist for this workload
    err = None
    gpath = os.path.join(ROOT, "tests", "golden", "solve_shapes1024.npz")
    if n == 1024 and os.path.exists(gpath):
        g = np.load(gpath)
        idx = g["vidx"]
        errs = {}
        for name, r in results.items():
            f = r["fwd"]
            errs[name] = float(max(np.abs(f[idx] - g["fx"]).max(), np.abs(f[V + idx] - g["fy"]).max()))
        err = dict(errs, reference_iterations=int(g["iters"]))

    # roofline of the dominant kernel: the FP64-MFMA DCT GEMM stage (2 n^3 flops per launch)
    ms = _lib.C.c_double()
    _lib.check(_lib.lib().otm_solver_profile_kernel(s._h, _lib.OTM_KERNEL_DCT_GEMM, args.profile_reps,
                                                    _lib.C.byref(ms)), "profile")
    gemm_ms = ms.value
    flops = 2.0 * n ** 3
    achieved = flops / (gemm_ms * 1e-3) / 1e12
    kernel_ms = {}
    kernels = [("beta", _lib.OTM_KERNEL_BETA), ("dir_coef", _lib.OTM_KERNEL_DIR_COEF),
               ("update", _lib.OTM_KERNEL_UPDATE), ("dct_solve_f64", _lib.OTM_KERNEL_DCT_SOLVE)]
    if "mixed" in results:
        kernels += [("mixed_gemm_bf16x3", _lib.OTM_KERNEL_MIXED_GEMM), ("mixed_solve", _lib.OTM_KERNEL_MIXED_SOLVE)]
    for kname, kid in kernels:
        _lib.check(_lib.lib().otm_solver_profile_kernel(s._h, kid, args.profile_reps, _lib.C.byref(ms)), "profile")
        kernel_ms[kname] = round(ms.value, 4)

    head = "mixed" if "mixed" in results else "fp64"
    line = {
        "metric": METRIC,
        "value": round(results[head]["ms"], 3),
        "unit": "ms/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(results[head]["ms"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3-split GEMM (f32 accumulate) + f64 refinement/stencils" if head == "mixed" else "f64",
        "data": f"synthetic: seeded smoothed shapes {n}x{n} (3 discs/rings, blur 5 px, floor 1e-3, mean 1)",
        "config": {"workload": f"single {n}x{n} grid transport solve to |r|^2/A <= 1e-7 (conjugate-Jacobian)",
                   "n": n, "threshold": 1e-7, "precision": head},
        "fp64_ms_per_solve": round(results["fp64"]["ms"], 3),
        "fp64_device_ms_per_solve": round(results["fp64"]["dev_ms"], 3),
        "mixed_ms_per_solve": round(results["mixed"]["ms"], 3) if "mixed" in results else None,
        "speedup_mixed_vs_fp64": (round(results["fp64"]["ms"] / results["mixed"]["ms"], 3)
                                  if "mixed" in results else None),
        "iterations": {k: v["iters"] for k, v in results.items()},
        "final_residual": {k: v["res"] for k, v in results.items()},
        "init_s": round(t_init, 4),
        "max_vertex_err_vs_reference": err,
        "kernel_ms": dict(kernel_ms, dct_gemm_f64=round(gemm_ms, 4)),
        "roofline": {"kernel": "k_gemm_f64 (one DCT stage, n^3 FP64 MFMA)", "bound": "mfma",
                     "achieved": round(achieved, 3), "peak": PEAK_FP64_MFMA_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_MFMA_TFLOPS, 4), "traffic": None},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(n, rho)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
