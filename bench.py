#!/usr/bin/env python3
"""Headline benchmark: wall time to relative residual 1e-12 (factorization + GMRES).

One step = factorize the assembled matrix on the GPU + FP64 GMRES to ||b - A x|| / ||b|| <=
1e-12, with the matrix blocks and b already resident in HBM.  The timed path is FP32 fronts
(batched, one grouped launch per level phase) with learned sparsification ranks (a sketch sized
by the rank predictor, tail-tested, geqp3 fallback) inside FP64 flexible GMRES(50); the plain
batched FP64 path (the reference's left-preconditioned GMRES) is timed on the same GPU in the
same run and reported beside it, as is the oracle (the CPU restatement of the reference path)
on a bounded sample.

Prints one JSON line.  --gpus > 1 is refused: this path runs on one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time to rel. residual 1e-12 (factor+GMRES), FP32/batched/learned vs FP64 plain"
PEAK = {"hbm": 8000.0, "fp64": 78.6, "fp32": 157.3}  # GB/s, TFLOP/s (MI355X dense)
CLASSES = ["geqrt", "larfb", "tpqrt_chain", "tpmqrt_chain", "scale_geqrt", "scale_apply", "rrqr_gather", "geqp3",
           "sparsify_apply_merge", "fill", "precond_solve", "spmv", "gmres_orth", "rank_sketch"]
MFMA_CLASSES = {"geqrt", "larfb", "tpqrt_chain", "tpmqrt_chain", "scale_geqrt", "scale_apply", "sparsify_apply_merge",
                "rank_sketch"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profile(tree):
    from pspaqr_public_amd._lib import lib

    out = np.zeros(5 * len(CLASSES))
    n = lib().spaqr_tree_profile(tree.handle, out.ctypes.data, len(CLASSES))
    rows = out[: 5 * n].reshape(n, 5)
    return {CLASSES[i]: dict(ms=rows[i, 0], launches=int(rows[i, 1]), tasks=int(rows[i, 2]), flops=rows[i, 3],
                             bytes=rows[i, 4]) for i in range(n)}


def run_steps(tree, b_dev, x_dev, n, steps, restart, iters, tag, flexible=False):
    from pspaqr_public_amd._lib import lib

    res = None
    t0 = time.perf_counter()
    fact_s = 0.0
    for s in range(steps):
        f0 = time.perf_counter()
        tree.factorize()
        fact_s += time.perf_counter() - f0
        res = tree.gmres_device(b_dev.ptr, x_dev.ptr, iters=iters, tol=1e-12, restart=restart, true_tol=1e-12,
                                flexible=flexible)
    lib().spaqr_tree_sync(tree.handle)
    t1 = time.perf_counter()
    log("[%s] %d step(s): %.4f s/step (factor %.4f s), gmres iters %d cycles %d true relres %.3e" % (
        tag, steps, (t1 - t0) / max(steps, 1), fact_s / max(steps, 1), res.iters, res.cycles, res.true_res))
    return (t1 - t0) / max(steps, 1), fact_s / max(steps, 1), res


def set_rank_mode(tree, mode):
    if mode:
        tree.set_rank_mode(mode)
        if mode == 1:
            from pspaqr_public_amd.rank_model import load_weights
            tree.set_rank_model(load_weights())


def cpu_baseline(cfg, side):
    """Oracle (CPU FP64 restatement of the reference path, one thread) on a bounded sample."""
    from oracle.oracle import OracleTree
    from pspaqr_public_amd.problems import make_problem

    P = make_problem(cfg, side)
    t0 = time.perf_counter()
    O = OracleTree(P.A, P.n_side, P.dim, P.nlevels, P.tol, P.skip)
    t1 = time.perf_counter()
    O.factorize()
    b = P.b[O.cperm()]
    x, it, _ = O.gmres(b, iters=300, tol=1e-12, restart=P.meta.get("restart", 0), true_tol=1e-12)
    t2 = time.perf_counter()
    r = np.linalg.norm(b - O.spmv(x)) / np.linalg.norm(b)
    log("[cpu oracle] %s: factor+gmres %.2f s (partition+assemble %.2f s), iters %d, relres %.2e" % (
        P.name, t2 - t1, t1 - t0, it, r))
    return P, t2 - t1, it, r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--side", type=int, default=None)
    ap.add_argument("--iters", type=int, default=300)
    ap.add_argument("--cpu-side", type=int, default=32, help="grid side of the CPU-baseline sample")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--rank-mode", type=int, default=1,
                    help="FP32 path's sparsification ranks: 1 learned sketch (default), 2 sketch without the "
                         "predictor, 0 geqp3 as the reference")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device three times and exit: the reference dispatch "
                         "that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        a = torch.ones(args.copy_calibration * (1 << 18), device="cuda")  # float32
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print("copied 3 x %d MiB" % args.copy_calibration, flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this path (one rank's whole share of the tree) runs on a single "
              "MI355X; --gpus %d is not supported (see DESIGN.md for how it would shard)." % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from pspaqr_public_amd._lib import DeviceBuffer, lib
    from pspaqr_public_amd.problems import make_problem
    from pspaqr_public_amd.ptree import build_tree

    P = make_problem(args.config, args.side)
    restart = P.meta.get("restart", 50) or 50
    log("[bench] %s n=%d nnz=%d levels=%d tol=%g" % (P.name, P.A.shape[0], P.A.nnz, P.nlevels, P.tol))

    def prepare(prec, rank_mode=0):
        t = build_tree(P, prec)
        set_rank_mode(t, rank_mode)
        b = P.b[t.cperm()]
        bd = DeviceBuffer.from_array(b)
        xd = DeviceBuffer(8 * t.n)
        return t, b, bd, xd

    # ---- FP32 fronts (the timed path)
    t0 = time.perf_counter()
    T32, b, b32, x32 = prepare(32, args.rank_mode)
    log("[bench] partition+assemble %.2f s" % (time.perf_counter() - t0))
    lib().spaqr_tree_set_profiling(T32.handle, 1)
    if args.warmup:
        run_steps(T32, b32, x32, T32.n, args.warmup, restart, args.iters, "fp32 warmup", True)
    lib().spaqr_tree_profile_reset(T32.handle)
    T32.rank_stats_reset()
    step_s, fact_s, res32 = run_steps(T32, b32, x32, T32.n, args.steps, restart, args.iters, "fp32", True)
    rstats = T32.rank_stats()
    prof = profile(T32)
    stats = T32.level_stats()
    x = x32.download(np.float64, T32.n)
    relres32 = float(np.linalg.norm(b - T32.spmv(x)) / np.linalg.norm(b))

    # ---- plain batched FP64 on the same GPU
    fp64 = None
    if not args.skip_fp64:
        T64, b64h, b64, x64 = prepare(64)
        run_steps(T64, b64, x64, T64.n, 1, 0, args.iters, "fp64 warmup")
        s64, f64, res64 = run_steps(T64, b64, x64, T64.n, 1, 0, args.iters, "fp64")
        fp64 = dict(seconds=s64, factor_s=f64, iters=res64.iters, true_relres=res64.true_res)
        del T64

    # HBM bytes per step of each launch class, from rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    # (calibrated on a device copy) kept beside the kernel-trace summaries; per launch below
    bytes_per_step = {}
    tp = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tp):
        with open(tp) as f:
            bytes_per_step = json.load(f).get("bytes_per_step", {})
    traffic_by_class = {c: int(round(v * max(args.steps, 1) / max(prof[c]["launches"], 1)))
                        for c, v in bytes_per_step.items() if c in prof and prof[c]["launches"]}

    # ---- roofline of the dominant kernel class (by time, over the timed steps)
    cls = max((c for c in prof if c not in ("precond_solve", "gmres_orth")), key=lambda c: prof[c]["ms"])
    p = prof[cls]
    per_launch_ms = p["ms"] / max(p["launches"], 1)
    if cls in MFMA_CLASSES:
        achieved = p["flops"] / (p["ms"] * 1e-3) / 1e12 if p["ms"] > 0 else 0.0
        roof = dict(bound="mfma", achieved=achieved, peak=PEAK["fp32"], unit="TFLOP/s", frac=achieved / PEAK["fp32"],
                    traffic=traffic_by_class.get(cls), kernel=cls, launches=p["launches"], tasks=p["tasks"],
                    avg_launch_ms=per_launch_ms, algorithmic_flops_per_launch=p["flops"] / max(p["launches"], 1))
    else:
        achieved = p["bytes"] / (p["ms"] * 1e-3) / 1e9 if p["ms"] > 0 else 0.0
        roof = dict(bound="hbm", achieved=achieved, peak=PEAK["hbm"], unit="GB/s", frac=achieved / PEAK["hbm"],
                    traffic=traffic_by_class.get(cls), kernel=cls, launches=p["launches"], tasks=p["tasks"],
                    avg_launch_ms=per_launch_ms, algorithmic_bytes_per_launch=p["bytes"] / max(p["launches"], 1))

    # ---- CPU baseline (oracle, bounded sample) and the GPU on the same sample
    cpu = None
    sample_gpu = None
    if not args.skip_cpu:
        Ps, cpu_s, cit, cres = cpu_baseline(args.config, args.cpu_side)
        Ts = build_tree(Ps, 32)
        set_rank_mode(Ts, args.rank_mode)
        bs = Ps.b[Ts.cperm()]
        bsd, xsd = DeviceBuffer.from_array(bs), DeviceBuffer(8 * Ts.n)
        run_steps(Ts, bsd, xsd, Ts.n, 1, restart, args.iters, "fp32 sample warmup", True)
        sg, _, _ = run_steps(Ts, bsd, xsd, Ts.n, 1, restart, args.iters, "fp32 sample", True)
        sample_gpu = sg
        cpu = dict(value=cpu_s, unit="s (factor+GMRES to 1e-12 on the %s sample, n=%d)" % (Ps.name, Ps.A.shape[0]),
                   cores=1, kind="port",
                   sample="%s: same generator and tree parameters at side %d (oracle: FP64, one thread, "
                          "GMRES iters %d, relres %.2e); GPU FP32 path on the same sample: %.4f s"
                          % (Ps.name, args.cpu_side, cit, cres, sg))

    per_class = {c: dict(ms_per_step=round(v["ms"] / max(args.steps, 1), 4), launches=v["launches"],
                         gflop=round(v["flops"] / 1e9, 4), gbytes=round(v["bytes"] / 1e9, 4))
                 for c, v in prof.items() if v["launches"]}
    out = {
        "metric": METRIC,
        "value": round(step_s, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "%s: 3D 7-point upwind convection-diffusion %d^3 (n=%d), tensor-grid ND -n %d -d 3, lvl %d, "
                        "tol %g; FP32 fronts in grouped batched launches, learned sketch ranks, FP64 flexible GMRES(%d) to "
                        "true relres 1e-12"
                        % (P.name, P.n_side, P.A.shape[0], P.n_side, P.nlevels, P.tol, restart)
                        if args.config == 3 else "%s (n=%d, lvl %d, tol %g)" % (P.name, P.A.shape[0], P.nlevels, P.tol),
            "n": int(P.A.shape[0]),
            "levels": int(P.nlevels),
            "tol": P.tol,
            "restart": restart,
        },
        "factor_s": round(fact_s, 6),
        "gmres_iters": res32.iters,
        "gmres_cycles": res32.cycles,
        "true_relres": relres32,
        "fp64_plain": fp64,
        "speedup_vs_fp64": (fp64["seconds"] / step_s) if fp64 else None,
        "factor_gflop": round(float(stats[:, 7].sum()) / 1e9, 3),
        "rank_mode": args.rank_mode,
        "rank_stats_per_step": {k: (v / max(args.steps, 1)) for k, v in rstats.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
        "per_class": per_class,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
