#!/usr/bin/env python3
"""Benchmark of the grid transport solver on one MI355X.

Workload (default): one 1024 x 1024 solve to the 1e-7 residual threshold on
the seeded smoothed-shapes density (3 discs/rings, Gaussian blur 5 px, floor
1e-3, mean 1), conjugate-Jacobian beta, max_iter 1000.  A step is one full
solve() -- density already resident in HBM, forward vertex positions out --
excluding init(n), which is reported separately, as the reference does.

Prints ONE JSON line: value = ms per solve with the default options (FP64,
the Laplacian solve by fast DCTs), beside it the same solve with the DCTs as
folded FP64 GEMMs on MFMA and the mixed form (3-term BF16-split GEMMs, FP64
refinement), each with its error against the reference's positions; the
roofline of the dominant kernel (measured live with HIP events on the
solver's stream) and the CPU baseline (the compiled reference, 1 thread, one
solve of the same density).

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

# HBM bytes per launch of the timed transform kernels at n = 1024 (the
# fast-DCT column pass; mixed M2 with 32- or 64-column tiles, FP64 F1 with
# tile shape c), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per
# dispatch), each scaled by what the counter reads for a 256 MiB device copy
# (FETCH_SIZE x2.0, WRITE_SIZE x1.0); the runs and their commands are in the
# profile named beside each figure.
GEMM_TRAFFIC = "profiles/gemm_1024_hbm_traffic.md"
FFT_TRAFFIC = "profiles/fft_1024_hbm_traffic.md"
STENCIL_TRAFFIC = "profiles/stencil_1024_hbm_traffic.md"
TRAFFIC_BYTES_1024 = {
    "k_dir_coef": ((16623.0 * 2.0 + 8304.0) * 1024, STENCIL_TRAFFIC),
    "k_fft_cols": ((4333.5 * 2.0 + 8192.0) * 1024, FFT_TRAFFIC),
    "k_mgemm<M2,32>": ((9288.0 * 2.0 + 6144.0) * 1024, GEMM_TRAFFIC),
    "k_mgemm<M2,64>": ((9294.0 * 2.0 + 6144.0) * 1024, GEMM_TRAFFIC),
    **{f"k_gemm_fold<F1,cfg{c}>": ((f * 2.0 + 8192.0) * 1024, GEMM_TRAFFIC)
       for c, f in enumerate([20521.6, 20517.0, 20517.0, 20513.0, 20537.5, 20522.0, 20526.5, 20564.3, 20520.0])},
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--profile-reps", type=int, default=20)
    ap.add_argument("--inverse-points", type=int, default=1 << 20)
    ap.add_argument("--batch", type=int, default=64, help="densities of the batched workload (0: skip)")
    ap.add_argument("--no-pipelines", action="store_true", help="skip the barycentre and stippling workloads")
    ap.add_argument("--batch-1024", type=int, default=8,
                    help="1024^2 densities of the batched mixed-vs-FP64 workload (0: skip)")
    ap.add_argument("--kernel", default=None,
                    choices=["dct_gemm_f64", "mixed_gemm", "beta", "dir_coef", "update", "dct_solve_f64",
                             "mixed_solve", "fft_solve", "fft_rows_fwd", "fft_cols", "fft_rows_inv", "dir_coef_pass",
                             "copy"],
                    help="only time this kernel (after one solve that fills the buffers); for profilers")
    ap.add_argument("--tile", type=int, default=None,
                    help="with --kernel mixed_gemm / dct_gemm_f64: force this tile config on the timed stage")
    return ap.parse_args()


KERNEL_IDS = {"dct_gemm_f64": 0, "beta": 1, "dir_coef": 2, "update": 3, "dct_solve_f64": 4, "mixed_gemm": 5,
              "mixed_solve": 6, "fft_solve": 7, "fft_rows_fwd": 8, "fft_cols": 9, "fft_rows_inv": 10, "dir_coef_pass": 11}


def kernel_only(args):
    """One solve to fill the buffers, then `profile_reps` back-to-back launches
    of one kernel between HIP events on the solver's stream."""
    import torch

    if args.kernel == "copy":  # calibration of the HBM traffic counters: 256 MiB read + 256 MiB written
        a = torch.ones(1 << 25, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(args.profile_reps):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"kernel": "copy", "bytes_read": a.numel() * 8, "bytes_written": b.numel() * 8}))
        return

    from otmap_amd import GridBasedTransportSolver, Precision, SolverOptions, Transform
    from otmap_amd import _lib
    from otmap_amd import densities as D

    n = args.n
    s = GridBasedTransportSolver(0)
    s.set_verbose_level(0)
    s.init(n)
    # one solve on the route of the timed kernel fills its intermediate buffers
    opt = (SolverOptions(precision=Precision.MIXED) if args.kernel.startswith("mixed") else
           SolverOptions(transform=Transform.FFT if args.kernel.startswith("fft") else Transform.GEMM))
    s.solve(D.smoothed_shapes(n) if n == 1024 else D.gaussian_mixture(n), opt)
    if args.tile is not None and args.kernel == "mixed_gemm":
        s.set_mixed_config(1, args.tile)
    elif args.tile is not None and args.kernel == "dct_gemm_f64":
        s.set_gemm_config(0, args.tile)
    ms = _lib.C.c_double()
    _lib.check(_lib.lib().otm_solver_profile_kernel(s._h, KERNEL_IDS[args.kernel], args.profile_reps,
                                                    _lib.C.byref(ms)), "profile")
    print(json.dumps({"kernel": args.kernel, "n": n, "reps": args.profile_reps, "ms_per_launch": ms.value,
                      "mixed_tiles": s.mixed_config()[0], "fold_tiles": s.gemm_config()[0]}))


def inverse_throughput(n, d_fwd, d_adj, count, reps):
    """TransportMap::inv (otlib/transport_map.cpp:50-102) of `count` seeded
    uniform points against the forward mesh already in HBM."""
    import torch

    from otmap_amd import _lib
    from otmap_amd import densities as D

    L = _lib.lib()
    # BVH build of a second map of the same mesh: the first build of a process
    # also allocates the build's device workspace (kept for later builds)
    t_again = None
    for _ in range(2):
        h = _lib.C.c_void_p()
        _lib.check(L.otm_map_create_device(_lib.C.byref(h), 0, n, d_fwd.data_ptr(), d_adj.data_ptr()), "map")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _lib.check(L.otm_map_init_inverse(h), "init_inverse")
        t_bvh, t_again = t_again, time.perf_counter() - t0
        if t_bvh is None:
            L.otm_map_destroy(h)
    try:
        dev = d_fwd.device
        pts = torch.from_numpy(D.uniform_points(count, seed=2000 + n).reshape(-1)).to(dev)
        out = torch.empty(2 * count, dtype=torch.float64, device=dev)
        face = torch.empty(count, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        _lib.check(L.otm_map_inverse_device(h, pts.data_ptr(), count, 0, out.data_ptr(), face.data_ptr(), None),
                   "inverse")
        t0 = time.perf_counter()
        for _ in range(reps):
            _lib.check(L.otm_map_inverse_device(h, pts.data_ptr(), count, 0, out.data_ptr(), face.data_ptr(), None),
                       "inverse")
        ms = (time.perf_counter() - t0) * 1e3 / reps
        located = int((face >= 0).sum().item())
    finally:
        L.otm_map_destroy(h)
    return {"queries": count, "ms": round(ms, 4), "queries_per_s": round(count / (ms * 1e-3)),
            "algorithmic_GB_per_s": round(36.0 * count / (ms * 1e-3) / 1e9, 1), "located": located,
            "bvh_build_ms": round(t_again * 1e3, 2), "bvh_build_first_ms": round(t_bvh * 1e3, 2),
            "mode": "inv (multi-hit selection), FP64"}


def batched_throughput(count, steps, warmup):
    """64 seeded Gaussian-mixture densities with grid sizes drawn from
    {128, 192, 256, 384, 512, 768, 1024}, solved together by one BatchSolver
    (device-resident inputs) in FP64 (default options: fast-DCT transform;
    and with the GEMM transform) and in the mixed form, against the same
    densities solved one after the other by single-grid FP64 solvers."""
    import torch

    from otmap_amd import BatchSolver, GridBasedTransportSolver, Precision, SolverOptions, Transform
    from otmap_amd import densities as D

    specs = D.batch_specs(count)
    dens = D.batch_densities(count)
    sizes = [n for n, _, _ in specs]
    dev = torch.device("cuda:0")
    d_in = [torch.from_numpy(d).to(dev) for d in dens]
    d_out = [torch.empty(2 * (n + 1) ** 2, dtype=torch.float64, device=dev) for n in sizes]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b = BatchSolver(sizes)
    t_create = time.perf_counter() - t0
    ip = [t.data_ptr() for t in d_in]
    op = [t.data_ptr() for t in d_out]

    def timed(opt):
        for _ in range(warmup):
            b.solve_device(ip, op, opt=opt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st = b.solve_device(ip, op, opt=opt)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps, [s.iterations for s in st]

    # largest sampled vertex error of each grid against the reference's solves
    gpath = os.path.join(ROOT, "tests", "golden", "batch64.npz")
    gold = np.load(gpath) if count == 64 and os.path.exists(gpath) else None

    def ref_err():
        if gold is None:
            return None
        worst = 0.0
        for p, n in enumerate(sizes):
            V = (n + 1) ** 2
            idx = torch.from_numpy(gold[f"d{p}_vidx"]).to(dev)
            xy = d_out[p]
            worst = max(worst, float((xy[idx].cpu() - torch.from_numpy(gold[f"d{p}_fx"])).abs().max()),
                        float((xy[V + idx].cpu() - torch.from_numpy(gold[f"d{p}_fy"])).abs().max()))
        return worst

    legs = {}
    for name, opt in (("mixed", SolverOptions(precision=Precision.MIXED)),
                      ("bf16x3_all_sizes", SolverOptions(precision=Precision.MIXED, mixed_min_n=0)),
                      ("bf16x3_all_sizes_refine2", SolverOptions(precision=Precision.MIXED, mixed_min_n=0,
                                                                 refine_steps=2)),
                      ("fp64_gemm", SolverOptions(transform=Transform.GEMM)),
                      ("fp64", SolverOptions())):
        t, it = timed(opt)
        legs[name] = dict(t=t, iters=it, fwd=torch.cat(d_out), err=ref_err())
    t_batch, iters, fwd_fp64 = legs["fp64"]["t"], legs["fp64"]["iters"], legs["fp64"]["fwd"]

    # the same densities through single-grid solvers, one handle per size
    solvers = {}
    for n in sorted(set(sizes)):
        s = GridBasedTransportSolver(0)
        s.set_verbose_level(0)
        s.init(n)
        solvers[n] = s
    for k in range(len(sizes)):  # warm every handle once (graph capture)
        solvers[sizes[k]].solve_device(ip[k], op[k])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(len(sizes)):
        solvers[sizes[k]].solve_device(ip[k], op[k])
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    return {"densities": count, "faces": int(sum(n * n for n in sizes)), "sizes": sorted(set(sizes)),
            "solves_per_s": round(count / t_batch, 2), "ms_per_batch": round(t_batch * 1e3, 2),
            "sequential_single_solves_per_s": round(count / t_seq, 2),
            "speedup_vs_sequential": round(t_seq / t_batch, 3),
            "iterations": {"min": min(iters), "max": max(iters), "mean": round(float(np.mean(iters)), 1)},
            "create_s": round(t_create, 3), "precision": "fp64", "transform": "fast DCT (FFT)",
            "max_vertex_err_vs_reference": legs["fp64"]["err"],
            "gemm_tiles_by_size": {str(n): b.gemm_config(sizes.index(n))[0] for n in sorted(set(sizes))},
            **{name: {"solves_per_s": round(count / L["t"], 2), "ms_per_batch": round(L["t"] * 1e3, 2),
                      "iterations_equal_fp64": L["iters"] == iters,
                      "max_vertex_err_vs_reference": L["err"],
                      "max_vertex_diff_vs_fp64": float((L["fwd"] - fwd_fp64).abs().max().item()),
                      "speedup_vs_fp64_batch": round(t_batch / L["t"], 3)}
               for name, L in legs.items() if name != "fp64"},
            "mixed_transform_from_n": SolverOptions().mixed_min_n,
            "mixed_tiles_by_size": {str(n): b.mixed_config(sizes.index(n))[0] for n in sorted(set(sizes))}}


def batched_1024(count, steps, warmup, fp64_single_ms):
    """`count` seeded smoothed-shape densities at 1024^2 (seeds 1..count past
    the single-solve seed) solved together by one BatchSolver, FP64 (fast-DCT
    and GEMM transforms) and mixed, against the single FP64 solve: the
    per-solve time of each batched form."""
    import torch

    from otmap_amd import BatchSolver, Precision, SolverOptions, Transform
    from otmap_amd import densities as D

    n = 1024
    dev = torch.device("cuda:0")
    d_in = [torch.from_numpy(D.smoothed_shapes(n, seed=D.SEED_SHAPES_1024 + 1 + k)).to(dev) for k in range(count)]
    d_out = [torch.empty(2 * (n + 1) ** 2, dtype=torch.float64, device=dev) for _ in range(count)]
    ip = [t.data_ptr() for t in d_in]
    op = [t.data_ptr() for t in d_out]
    b = BatchSolver([n] * count)
    out = {"densities": count, "n": n}
    fwd = {}
    for name, opt in (("fp64", SolverOptions()), ("fp64_gemm", SolverOptions(transform=Transform.GEMM)),
                      ("mixed", SolverOptions(precision=Precision.MIXED))):
        for _ in range(warmup):
            b.solve_device(ip, op, opt=opt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st = b.solve_device(ip, op, opt=opt)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        fwd[name] = torch.cat(d_out)
        out[f"{name}_ms_per_batch"] = round(ms, 2)
        out[f"{name}_ms_per_solve"] = round(ms / count, 3)
        out[f"{name}_iterations"] = [x.iterations for x in st]
    out["max_vertex_diff_mixed_vs_fp64"] = float((fwd["mixed"] - fwd["fp64"]).abs().max().item())
    out["max_vertex_diff_fp64_gemm_vs_fp64"] = float((fwd["fp64_gemm"] - fwd["fp64"]).abs().max().item())
    out["speedup_fp64_batched_vs_fp64_single"] = round(fp64_single_ms / out["fp64_ms_per_solve"], 3)
    out["speedup_mixed_batched_vs_fp64_single"] = round(fp64_single_ms / out["mixed_ms_per_solve"], 3)
    out["mixed_tiles"] = b.mixed_config()[0]
    return out


def barycentre_pipeline(n=512, steps=4):
    """Four seeded smoothed-shape densities at 512^2 solved as one batch
    (FP64), the inverse map of each as a mesh, 5 x 5 bilinear interpolations,
    each rasterised to 512^2 and mean-normalised to FP32."""
    import torch

    from otmap_amd import BatchSolver, pipelines
    from otmap_amd import densities as D

    dens = D.barycentre_inputs(n)
    means = [float(d.mean()) for d in dens]
    b = BatchSolver([n] * len(dens))
    b.solve(dens)  # warm (graph capture, tuning)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    maps = b.solve(dens)
    t_solve = time.perf_counter() - t0
    pipelines.barycentres(maps, steps, n, means)  # warm the BVH builds
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    imgs = pipelines.barycentres(maps, steps, n, means)
    torch.cuda.synchronize()
    t_interp = time.perf_counter() - t0
    return {"densities": len(dens), "n": n, "images": len(imgs), "solve_ms": round(t_solve * 1e3, 2),
            "inverse_interpolate_raster_ms": round(t_interp * 1e3, 2),
            "total_ms": round((t_solve + t_interp) * 1e3, 2),
            "iterations": [s.iterations for s in b.last_stats], "image_dtype": "f32"}


def stippling_pipeline(n=2048, points=1 << 20):
    """2048^2 log-normal x Gaussian-mixture density (contrast capped at 1e4 by
    max_ratio), mixed-precision solve, then 1M uniform points advected through
    the inverse map and pruned on zero density; FP64 solves (fast-DCT and GEMM
    transforms) beside it for the map difference.  Each mode's solve is timed
    after one warm solve."""
    import torch

    from otmap_amd import (GridBasedTransportSolver, GridMesh, Precision, SolverOptions, Transform, TransportMap,
                           pipelines)
    from otmap_amd import densities as D

    rho = D.hdr_lognormal(n)
    dev = torch.device("cuda:0")
    d_rho = torch.from_numpy(rho).to(dev)
    V = (n + 1) ** 2
    d_fwd = {p: torch.empty(2 * V, dtype=torch.float64, device=dev) for p in ("fp64", "fp64_gemm", "mixed")}
    d_adj = torch.empty(n * n, dtype=torch.float64, device=dev)
    s = GridBasedTransportSolver(0)
    s.set_verbose_level(0)
    s.init(n)
    out = {"n": n, "density": "log-normal (sigma 2, corr 16 px) x Gaussian mixture, max_ratio 1e4, floor 1e-4"}
    gpath = os.path.join(ROOT, "tests", "golden", "stipple2048.npz")
    gold = np.load(gpath) if n == 2048 and os.path.exists(gpath) else None
    for name, opt in (("fp64_gemm", SolverOptions(transform=Transform.GEMM, max_ratio=1e4)),
                      ("fp64", SolverOptions(max_ratio=1e4)),
                      ("mixed", SolverOptions(precision=Precision.MIXED, max_ratio=1e4))):
        # warm solve: the first mixed solve builds the BF16 factor planes and
        # tunes the mixed tiles, the first solve of each mode captures its graph
        s.solve_device(d_rho.data_ptr(), d_fwd[name].data_ptr(), d_adj.data_ptr(), opt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = s.solve_device(d_rho.data_ptr(), d_fwd[name].data_ptr(), d_adj.data_ptr(), opt)
        torch.cuda.synchronize()
        out[f"{name}_solve_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
        out[f"{name}_iterations"] = st.iterations
        if gold is not None:  # largest sampled vertex error against the reference's solve
            xy = d_fwd[name].cpu().numpy()
            idx = gold["vidx"]
            out[f"{name}_max_vertex_err_vs_reference"] = float(max(np.abs(xy[idx] - gold["fx"]).max(),
                                                                   np.abs(xy[V + idx] - gold["fy"]).max()))
    if gold is not None:
        out["reference_iterations"] = int(gold["iters"])
    out["max_vertex_diff_mixed_vs_fp64"] = float((d_fwd["mixed"] - d_fwd["fp64"]).abs().max().item())
    out["max_vertex_diff_fp64_gemm_vs_fp64"] = float((d_fwd["fp64_gemm"] - d_fwd["fp64"]).abs().max().item())
    tm = TransportMap(GridMesh.regular(n), GridMesh(n, d_fwd["mixed"].cpu().numpy()), d_adj.cpu().numpy())
    pts = D.uniform_points(points, seed=2000 + n)
    t0 = time.perf_counter()
    tm.init_inverse()
    t_bvh = time.perf_counter() - t0
    pipelines.stippling(tm, rho, pts)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kept = pipelines.stippling(tm, rho, pts)
    torch.cuda.synchronize()
    out.update({"points": points, "kept": int(kept.shape[0]), "bvh_build_ms": round(t_bvh * 1e3, 2),
                "advect_prune_ms": round((time.perf_counter() - t0) * 1e3, 2)})
    return out


def host_cpu():
    """Model name of the host CPU and the cores this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        usable = os.cpu_count()
    return {"model": model, "nproc": usable, "logical_cpus": os.cpu_count()}


def cpu_baseline(n, density, reps=3):
    """The compiled reference (oracle/_ref/libotref.so), 1 thread: init(n)
    once, then `reps` full solves of the same density; the median solve."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        from oracle import Reference, reference_available
    except Exception as e:  # pragma: no cover
        return {"value": None, "unit": "ms/solve", "cores": 1, "kind": "reference", "sample": f"unavailable: {e}"}
    if not reference_available():
        return {"value": None, "unit": "ms/solve", "cores": 1, "kind": "reference",
                "sample": "oracle/_ref/libotref.so not built"}
    t_init, ts = Reference().solve_timed(n, density, reps)
    return {"value": round(float(np.median(ts)) * 1e3, 1), "unit": "ms/solve", "cores": 1, "kind": "reference",
            "solves_ms": [round(t * 1e3, 1) for t in ts], "host": host_cpu(),
            "sample": f"median of {reps} full solves of the same {n}x{n} density (one init, {t_init:.2f} s, "
                      f"excluded), reference built -O3 -DENABLE_SSE_MODE, SimplicialLDLT (no CHOLMOD), 1 thread"}


def fft_supported(n):
    """Sizes the fast-DCT transform takes: n = 2^a 3^b, 8 <= n <= 2048, n % 4 == 0."""
    if n < 8 or n > 2048 or n % 4:
        return False
    while n % 2 == 0:
        n //= 2
    while n % 3 == 0:
        n //= 3
    return n == 1


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); "
              "the batched and single solves are measured per GPU", file=sys.stderr)
        sys.exit(2)
    if args.kernel:
        kernel_only(args)
        return
    import torch

    from otmap_amd import GridBasedTransportSolver, Precision, SolverOptions, Transform
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

    # fp64: the default options (FP64, fast-DCT transform where n allows it);
    # fp64_gemm: the same solve with the DCTs as folded FP64 GEMMs on MFMA;
    # mixed: 3-term BF16-split GEMMs + one FP64 refinement
    fft = fft_supported(n)
    results = {}
    modes = [("fp64", SolverOptions()), ("fp64_gemm", SolverOptions(transform=Transform.GEMM))]
    if n % 64 == 0:
        modes.append(("mixed", SolverOptions(precision=Precision.MIXED)))
    for name, opt in modes:
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
    # the inverse map below runs on the default mode's forward map
    s.solve_device(d_rho.data_ptr(), d_fwd.data_ptr(), d_adj.data_ptr(), SolverOptions())

    # inverse map at 1M uniform query points (device buffers): BVH build on the
    # device (once per map, as init_inverse), then one query per lane
    inv_stats = inverse_throughput(n, d_fwd, d_adj, args.inverse_points, args.steps)
    batch_stats = batched_throughput(args.batch, args.steps, args.warmup) if args.batch > 0 else None
    b1024 = (batched_1024(args.batch_1024, args.steps, args.warmup, results["fp64"]["ms"])
             if args.batch_1024 > 0 and n == 1024 else None)
    bary = None if args.no_pipelines else barycentre_pipeline()
    stipple = None if args.no_pipelines else stippling_pipeline()

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

    # per-kernel device time: back-to-back launches between HIP events recorded
    # on the solver's own stream (the stream every kernel of a solve runs on)
    ms = _lib.C.c_double()
    kernel_ms = {}
    kernels = [("beta", _lib.OTM_KERNEL_BETA), ("dir_coef", _lib.OTM_KERNEL_DIR_COEF),
               ("update", _lib.OTM_KERNEL_UPDATE), ("dir_coef_pass", _lib.OTM_KERNEL_DIR_COEF_PASS),
               ("dct_gemm_f64", _lib.OTM_KERNEL_DCT_GEMM),
               ("dct_solve_f64", _lib.OTM_KERNEL_DCT_SOLVE)]
    if fft:
        kernels += [("fft_solve", _lib.OTM_KERNEL_FFT_SOLVE), ("fft_rows_fwd", _lib.OTM_KERNEL_FFT_ROWS_FWD),
                    ("fft_cols", _lib.OTM_KERNEL_FFT_COLS), ("fft_rows_inv", _lib.OTM_KERNEL_FFT_ROWS_INV)]
    if "mixed" in results:
        kernels += [("mixed_gemm_bf16x3", _lib.OTM_KERNEL_MIXED_GEMM), ("mixed_solve", _lib.OTM_KERNEL_MIXED_SOLVE)]
    for kname, kid in kernels:
        _lib.check(_lib.lib().otm_solver_profile_kernel(s._h, kid, args.profile_reps, _lib.C.byref(ms)), "profile")
        kernel_ms[kname] = ms.value

    head = "fp64"

    def traffic_of(kname):
        t, src = TRAFFIC_BYTES_1024.get(kname, (None, None)) if n == 1024 else (None, None)
        return {"traffic": round(t) if t else None, "traffic_source": src}

    # Roofline of the dominant kernel of the default iteration, HBM-bound: the
    # longer (live) of the line-search pass k_dir_coef and the fast-DCT column
    # pass, the two longest kernels of the iteration; the other is reported
    # beside it.  Algorithmic bytes per launch: k_dir_coef reads d_hat, d_prev,
    # psi and r and writes d (5 n^2 FP64 values = 40 n^2 bytes; tile halos and
    # the workgroup partials are not counted); the column pass reads and
    # writes the n x n FP64 field once (16 n^2 bytes; the twiddle, phase and
    # eigenvalue tables add 64 n bytes and are not counted).
    def hbm_roofline(kname, t_ms, nbytes, label):
        achieved = nbytes / (t_ms * 1e-3) / 1e9
        return {"kernel": label, "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), **traffic_of(kname),
                "ms_per_launch": round(t_ms, 5)}

    roof_stencil = hbm_roofline("k_dir_coef", kernel_ms["dir_coef_pass"], 40.0 * n * n,
                                "k_dir_coef: line-search pass (direction d = d_hat + beta d_prev, quartic "
                                "coefficients, 7 partial dots), reads d_hat, d_prev, psi, r and writes d")
    roof_fft = None
    if fft:
        # the column pass runs in row form for n = 512, 1024 and 2048 (one wave
        # per grid row), in the workgroup form at other sizes
        cname = {512: "k_fftw_cols<4, 4>", 1024: "k_fftw_cols<8, 4>", 2048: "k_fftw_cols<16, 4>"}.get(
            n, "k_fft_cols<1>" if n <= 1024 else "k_fft_cols<2>")
        roof_fft = hbm_roofline("k_fft_cols", kernel_ms["fft_cols"], 16.0 * n * n,
                                f"{cname}: fast-DCT column pass (DCT-II, / lambda, DCT-III of n columns), "
                                "one read + one write of the n x n FP64 field")
    dominant = roof_stencil
    if roof_fft and kernel_ms["fft_cols"] > kernel_ms["dir_coef_pass"]:
        dominant = roof_fft

    # Rooflines of the GEMM routes: one stage of the even/odd-folded DCT
    # (k_mgemm stage M2 for mixed, k_gemm_fold stage F1 for FP64).
    # Algorithmic work per launch is the folded product, n x n outputs over
    # n/2 terms = n^3 flops; the BF16 split multiplicity (6 products) is
    # implementation overhead and is not counted.
    def gemm_roofline(kname, t_ms, peak, label, flops):
        achieved = flops / (t_ms * 1e-3) / 1e12
        return {"kernel": label, "bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), **traffic_of(kname), "ms_per_launch": round(t_ms, 5)}

    tiles, n_tiles = s.gemm_config()
    folded = n_tiles > 0
    stage_flops_f64 = float(n) ** 3 if folded else 2.0 * n ** 3
    roof_f64 = gemm_roofline(f"k_gemm_fold<F1,cfg{tiles[0]}>" if folded else "k_gemm_f64",
                             kernel_ms["dct_gemm_f64"], PEAK_FP64_MFMA_TFLOPS,
                             f"k_gemm_fold<F1> (tile shape {tiles[0]}): one even/odd-folded DCT stage "
                             "(n^3 flops), FP64 MFMA" if folded else
                             "k_gemm_f64: one DCT stage (2 n^3 flops), FP64 MFMA", stage_flops_f64)
    roof_x3 = None
    if "mixed" in results:
        kname = ["k_mgemm<M2,32>", "k_mgemm<M2,64>", "k_mgemm_w<M2,4 waves>", "k_mgemm_w<M2,8 waves>"][
            s.mixed_config()[0][1]]
        roof_x3 = gemm_roofline(kname, kernel_ms["mixed_gemm_bf16x3"], PEAK_BF16_TFLOPS,
                                f"{kname}: one even/odd-folded DCT stage (n^3 flops), 3-term BF16 "
                                "split on MFMA, FP32 accumulate", float(n) ** 3)

    # Whole-iteration roofline: the reference's 26 FP64 field touches per face
    # in the stencil passes (218 MB at 1024^2; this build moves 22: the quartic
    # coefficients are recomputed instead of stored) at HBM peak, plus the
    # Laplacian solve: for the fast-DCT route its three passes' 6 field
    # touches at HBM peak, for the GEMM routes the algorithmic flops of the
    # four transform stages (4 n^3 folded, 8 n^3 unfolded; the BF16 split
    # multiplicity and the refinement transform are not counted) at the MFMA
    # peak of the mode; over the measured device time per iteration.
    def iteration_roofline(name, t_solve_roof):
        t_roof = 26 * 8 * n * n / (PEAK_HBM_GBS * 1e9) + t_solve_roof
        t_iter = results[name]["dev_ms"] * 1e-3 / max(results[name]["iters"], 1)
        return {"t_roof_us": round(t_roof * 1e6, 2), "t_iter_us": round(t_iter * 1e6, 2),
                "frac": round(t_roof / t_iter, 4)}

    t_gemm64 = 4 * stage_flops_f64 / (PEAK_FP64_MFMA_TFLOPS * 1e12)
    iter_roof = {"fp64": iteration_roofline("fp64", 6 * 8 * n * n / (PEAK_HBM_GBS * 1e9) if fft else t_gemm64),
                 "fp64_gemm": iteration_roofline("fp64_gemm", t_gemm64)}
    if "mixed" in results:
        iter_roof["mixed"] = iteration_roofline("mixed", 4.0 * float(n) ** 3 / (PEAK_BF16_TFLOPS * 1e12))
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
        "dtype": "f64",
        "data": f"synthetic: seeded smoothed shapes {n}x{n} (3 discs/rings, blur 5 px, floor 1e-3, mean 1)",
        "config": {"workload": f"single {n}x{n} grid transport solve to |r|^2/A <= 1e-7 (conjugate-Jacobian)",
                   "n": n, "threshold": 1e-7, "precision": "fp64",
                   "transform": "fast DCT (FFT)" if fft else "folded GEMMs"},
        "fp64_ms_per_solve": round(results["fp64"]["ms"], 3),
        "fp64_device_ms_per_solve": round(results["fp64"]["dev_ms"], 3),
        "fp64_gemm_ms_per_solve": round(results["fp64_gemm"]["ms"], 3),
        "mixed_ms_per_solve": round(results["mixed"]["ms"], 3) if "mixed" in results else None,
        "speedup_fp64_vs_fp64_gemm": round(results["fp64_gemm"]["ms"] / results["fp64"]["ms"], 3),
        "speedup_mixed_vs_fp64_gemm": (round(results["fp64_gemm"]["ms"] / results["mixed"]["ms"], 3)
                                       if "mixed" in results else None),
        "iterations": {k: v["iters"] for k, v in results.items()},
        "final_residual": {k: v["res"] for k, v in results.items()},
        "init_s": round(t_init, 4),
        "max_vertex_err_vs_reference": err,
        "inverse_1M": inv_stats,
        "batched": batch_stats,
        "batched_1024": b1024,
        "barycentre_4x512": bary,
        "stippling_2048": stipple,
        "kernel_ms": {k: round(v, 4) for k, v in kernel_ms.items()},
        "roofline": dominant,
        "roofline_dir_coef": roof_stencil,
        "roofline_fft_cols": roof_fft,
        "roofline_fp64_gemm": roof_f64,
        "roofline_mixed_gemm": roof_x3,
        "iteration_roofline": iter_roof,
        "dct_tiles": {"folded": folded, "stage_config": tiles, "candidates": n_tiles},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(n, rho)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
