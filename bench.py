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


def run_steps(tree, b_dev, x_dev, n, steps, restart, iters, tag, flexible=False, batch=False):
    """steps x (factorize + GMRES to 1e-12); returns (s/step, factor s/step, [GmresResult per system])."""
    from pspaqr_public_amd._lib import lib

    res = None
    t0 = time.perf_counter()
    fact_s = 0.0
    for s in range(steps):
        f0 = time.perf_counter()
        tree.factorize()
        fact_s += time.perf_counter() - f0
        if batch:
            res = tree.gmres_batch_device(b_dev.ptr, x_dev.ptr, iters=iters, tol=1e-12, restart=restart,
                                          true_tol=1e-12, flexible=flexible)
        else:
            res = [tree.gmres_device(b_dev.ptr, x_dev.ptr, iters=iters, tol=1e-12, restart=restart, true_tol=1e-12,
                                     flexible=flexible)]
    lib().spaqr_tree_sync(tree.handle)
    t1 = time.perf_counter()
    log("[%s] %d step(s): %.4f s/step (factor %.4f s), gmres iters %s, worst true relres %.3e" % (
        tag, steps, (t1 - t0) / max(steps, 1), fact_s / max(steps, 1),
        res[0].iters if len(res) == 1 else "%d..%d" % (min(r.iters for r in res), max(r.iters for r in res)),
        max(r.true_res for r in res)))
    return (t1 - t0) / max(steps, 1), fact_s / max(steps, 1), res


def set_rank_mode(tree, mode):
    if mode:
        tree.set_rank_mode(mode)
        if mode == 1:
            from pspaqr_public_amd.rank_model import load_weights
            tree.set_rank_model(load_weights())


def cpu_baseline(cfg, side):
    """Oracle (CPU FP64 restatement of the reference path) on a bounded sample; returns
    (problem, seconds, GMRES iterations, true relres, threads used)."""
    from oracle.oracle import OracleTree, threads
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
    log("[cpu oracle] %s: factor+gmres %.2f s (partition+assemble %.2f s), iters %d, relres %.2e, %d thread(s)" % (
        P.name, t2 - t1, t1 - t0, it, r, threads()))
    return P, t2 - t1, it, r, threads()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--side", type=int, default=None)
    ap.add_argument("--iters", type=int, default=300)
    ap.add_argument("--cpu-side", type=int, default=None,
                    help="grid side of a smaller CPU-baseline sample (default: the benchmarked workload itself; "
                         "the batch uses the 32^3 config-2 system)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-mode2", action="store_true", help="skip the no-predictor leg")
    ap.add_argument("--graphs", action="store_true", help="also time the FP64 path replayed from hip graphs")
    ap.add_argument("--rank-mode", type=int, default=1,
                    help="FP32 path's sparsification ranks: 1 learned sketch (default), 2 sketch without the "
                         "predictor, 0 geqp3 as the reference")
    ap.add_argument("--tune", default="", metavar="NAME=V,...",
                    help="planner tunables applied after the autotuned table (FP32 leg), e.g. sketch_bf16=2")
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

    from pspaqr_public_amd import autotune
    from pspaqr_public_amd._lib import DeviceBuffer, lib
    from pspaqr_public_amd.problems import make_batch, make_problem
    from pspaqr_public_amd.ptree import build_batch_tree, build_tree

    batch = args.config == 4
    if batch:
        W = make_batch()
        restart, name, wkey = 50, W.name, "batch48"
        n_total, levels, tol = int(W.A.shape[0]), max(W.levels), W.tol
        log("[bench] %s: %d systems, n=%d nnz=%d, levels %s, tol %g" % (W.name, len(W.systems), n_total, W.A.nnz,
                                                                     sorted(set(W.levels)), W.tol))
    else:
        W = make_problem(args.config, args.side)
        restart, name, n_total, levels, tol = W.meta.get("restart", 50) or 50, W.name, int(W.A.shape[0]), W.nlevels, W.tol
        wkey = {1: "poisson2d_64", 2: "diffusion3d_32", 3: "convdiff3d_64", 5: "anisotropic3d_96"}.get(args.config)
        if args.side is not None:
            wkey = None
        log("[bench] %s n=%d nnz=%d levels=%d tol=%g" % (W.name, n_total, W.A.nnz, W.nlevels, W.tol))

    # the per-matrix symbolic phase (partition + assemble, outside the timed step) runs on the
    # device for the FP32 trees and on the host for the FP64 one, both timed
    symbolic = {}

    def prepare(prec, rank_mode=0, tuned=True):
        dev = prec == 32 and not batch
        t = build_batch_tree(W, prec) if batch else build_tree(W, prec, device_symbolic=dev)
        if not batch:
            symbolic.setdefault("device_ms" if dev else "host_ms", [round(v, 2) for v in t.symbolic_ms()])
        set_rank_mode(t, rank_mode)
        used = autotune.apply(t, wkey, prec) if (tuned and wkey) else {}
        if prec == 32 and args.tune:
            extra = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in args.tune.split(",")}
            t.set_tuning(extra)
            used = dict(used, **extra)
        b = W.b[t.cperm()]
        return t, b, DeviceBuffer.from_array(b), DeviceBuffer(8 * t.n), used

    # ---- FP32 fronts with learned ranks (the timed path)
    t0 = time.perf_counter()
    T32, b, b32, x32, tuning = prepare(32, args.rank_mode)
    log("[bench] partition+assemble %.2f s; tuned widths %s" % (time.perf_counter() - t0, tuning or "built-in"))
    lib().spaqr_tree_set_profiling(T32.handle, 1)
    if args.warmup:
        run_steps(T32, b32, x32, T32.n, args.warmup, restart, args.iters, "fp32 warmup", True, batch)
    lib().spaqr_tree_profile_reset(T32.handle)
    T32.rank_stats_reset()
    step_s, fact_s, res32 = run_steps(T32, b32, x32, T32.n, args.steps, restart, args.iters, "fp32", True, batch)
    rstats = T32.rank_stats()
    prof = profile(T32)
    stats = T32.level_stats()
    x = x32.download(np.float64, T32.n)
    r = b - T32.spmv(x)
    if batch:
        relres32 = max(float(np.linalg.norm(r[o:o + m]) / np.linalg.norm(b[o:o + m])) for o, m in T32.systems())
    else:
        relres32 = float(np.linalg.norm(r) / np.linalg.norm(b))
    lib().spaqr_tree_set_profiling(T32.handle, 0)
    T32.close()

    # ---- the same path without the predictor (sketches as wide as min(k, w)): the predictor's gain
    mode2 = None
    if args.rank_mode == 1 and not args.skip_mode2:
        T2, _, b2, x2, _ = prepare(32, 2)
        run_steps(T2, b2, x2, T2.n, 1, restart, args.iters, "fp32 no-predictor warmup", True, batch)
        T2.rank_stats_reset()
        s2, f2, r2 = run_steps(T2, b2, x2, T2.n, max(1, min(args.steps, 3)), restart, args.iters, "fp32 no-predictor",
                               True, batch)
        st2 = T2.rank_stats()
        mode2 = dict(seconds=s2, factor_s=f2, gmres_iters=max(q.iters for q in r2),
                     worst_true_relres=max(q.true_res for q in r2),
                     sketch_cols_per_step=st2["sketch_cols"] / max(1, min(args.steps, 3)))
        T2.close()

    # ---- plain batched FP64 on the same GPU
    fp64 = None
    if not args.skip_fp64:
        T64, _, b64, x64, _ = prepare(64)
        run_steps(T64, b64, x64, T64.n, 1, 0, args.iters, "fp64 warmup", False, batch)
        lib().spaqr_tree_set_profiling(T64.handle, 1)
        lib().spaqr_tree_profile_reset(T64.handle)
        s64, f64, res64 = run_steps(T64, b64, x64, T64.n, 1, 0, args.iters, "fp64", False, batch)
        p64 = profile(T64)
        fp64 = dict(seconds=s64, factor_s=f64, gmres_iters=max(q.iters for q in res64),
                    worst_true_relres=max(q.true_res for q in res64),
                    per_class={c: dict(ms_per_step=round(v["ms"], 4), launches=v["launches"],
                                       gflop=round(v["flops"] * 1e-9, 4), gbytes=round(v["bytes"] * 1e-9, 4))
                               for c, v in p64.items() if v["launches"]})
        T64.close()
        if args.graphs:
            # the same FP64 path with every level's launches replayed from recorded hip graphs
            TG, _, bg, xg, _ = prepare(64)
            TG.set_graphs(True)
            run_steps(TG, bg, xg, TG.n, 2, 0, args.iters, "fp64 graphs record", False, batch)
            sg, fg, rg = run_steps(TG, bg, xg, TG.n, max(1, args.steps), 0, args.iters, "fp64 graphs replay", False,
                                   batch)
            fp64["graph_replay"] = dict(seconds=sg, factor_s=fg, gmres_iters=max(q.iters for q in rg),
                                        worst_true_relres=max(q.true_res for q in rg), **TG.graph_info())
            TG.close()

    # HBM bytes per step of each launch class, from rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    # (calibrated on a device copy) kept beside the kernel-trace summaries; per launch below
    bytes_per_step = {}
    tp = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tp) and not batch:
        with open(tp) as f:
            bytes_per_step = json.load(f).get("bytes_per_step", {})
    traffic_by_class = {c: int(round(v / max(prof[c]["launches"] / max(args.steps, 1), 1e-9)))
                        for c, v in bytes_per_step.items() if c in prof and prof[c]["launches"]}

    # ---- roofline of the dominant launch class (by time, over the timed steps)
    cls = max(prof, key=lambda c: prof[c]["ms"])
    p = prof[cls]
    per_launch_ms = p["ms"] / max(p["launches"], 1)
    common = dict(traffic=traffic_by_class.get(cls), kernel=cls, launches=p["launches"], tasks=p["tasks"],
                  avg_launch_ms=per_launch_ms)
    if cls in MFMA_CLASSES:
        achieved = p["flops"] / (p["ms"] * 1e-3) / 1e12 if p["ms"] > 0 else 0.0
        roof = dict(bound="mfma", achieved=achieved, peak=PEAK["fp32"], unit="TFLOP/s", frac=achieved / PEAK["fp32"],
                    algorithmic_flops_per_launch=p["flops"] / max(p["launches"], 1), **common)
    else:
        achieved = p["bytes"] / (p["ms"] * 1e-3) / 1e9 if p["ms"] > 0 else 0.0
        roof = dict(bound="hbm", achieved=achieved, peak=PEAK["hbm"], unit="GB/s", frac=achieved / PEAK["hbm"],
                    algorithmic_bytes_per_launch=p["bytes"] / max(p["launches"], 1), **common)

    # ---- CPU baseline (oracle, bounded sample) and the GPU on the same sample
    cpu = None
    if not args.skip_cpu:
        cfg_cpu = 2 if batch else args.config
        side_cpu = args.cpu_side if (args.cpu_side or batch) else args.side
        Ps, cpu_s, cit, cres, threads = cpu_baseline(cfg_cpu, side_cpu if not batch else (args.cpu_side or None))
        if not batch and side_cpu == args.side:
            sg = step_s  # the sample is the benchmarked workload: its GPU time is the timed step
            what = "the benchmarked workload itself"
        else:
            Ts = build_tree(Ps, 32)
            set_rank_mode(Ts, args.rank_mode)
            bs = Ps.b[Ts.cperm()]
            bsd, xsd = DeviceBuffer.from_array(bs), DeviceBuffer(8 * Ts.n)
            run_steps(Ts, bsd, xsd, Ts.n, 1, restart, args.iters, "fp32 sample warmup", True)
            sg, _, _ = run_steps(Ts, bsd, xsd, Ts.n, 1, restart, args.iters, "fp32 sample", True)
            Ts.close()
            what = "a smaller system of the same generator"
        cpu = dict(value=cpu_s, unit="s (factor+GMRES to 1e-12 on %s, n=%d)" % (Ps.name, Ps.A.shape[0]),
                   cores=threads, kind="port",
                   sample="%s (%s): the oracle -- the CPU FP64 restatement of the reference path, its per-level "
                          "loops on %d OpenMP threads -- GMRES iters %d, relres %.2e; the GPU FP32 path on the same "
                          "system: %.4f s" % (Ps.name, what, threads, cit, cres, sg))

    per_class = {c: dict(ms_per_step=round(v["ms"] / max(args.steps, 1), 4), launches=v["launches"],
                         gflop=round(v["flops"] / 1e9, 4), gbytes=round(v["bytes"] / 1e9, 4))
                 for c, v in prof.items() if v["launches"]}
    if batch:
        workload = ("%s: %d independent 3D 7-point diffusion systems (sides %s, n=%d in total), per-system tensor-grid "
                    "ND, tol %g; FP32 fronts of all systems in shared per-level grouped launches, learned sketch "
                    "ranks, batched FP64 flexible GMRES(%d) to true relres 1e-12 on every system"
                    % (name, len(W.systems), sorted(set(q.n_side for q in W.systems)), n_total, tol, restart))
    elif args.config == 3:
        workload = ("%s: 3D 7-point upwind convection-diffusion %d^3 (n=%d), tensor-grid ND -n %d -d 3, lvl %d, "
                    "tol %g; FP32 fronts in grouped batched launches, learned sketch ranks, FP64 flexible GMRES(%d) "
                    "to true relres 1e-12" % (name, W.n_side, n_total, W.n_side, levels, tol, restart))
    else:
        workload = "%s (n=%d, lvl %d, tol %g)" % (name, n_total, levels, tol)
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
        "config": {"workload": workload, "n": n_total, "levels": int(levels), "tol": tol, "restart": restart},
        "factor_s": round(fact_s, 6),
        "gmres_iters": res32[0].iters if len(res32) == 1 else [q.iters for q in res32],
        "gmres_cycles": max(q.cycles for q in res32),
        "true_relres": relres32,
        "fp64_plain": fp64,
        "speedup_vs_fp64": (fp64["seconds"] / step_s) if fp64 else None,
        "no_predictor": mode2,
        "predictor_gain": (mode2["seconds"] / step_s) if mode2 else None,
        "tuning": tuning or "built-in",
        "factor_gflop": round(float(stats[:, 7].sum()) / 1e9, 3),
        "rank_mode": args.rank_mode,
        "rank_stats_per_step": {k: (v / max(args.steps, 1)) for k, v in rstats.items()},
        "symbolic_partition_assemble_ms": symbolic or None,
        "roofline": roof,
        "cpu_baseline": cpu,
        "per_class": per_class,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
