"""Summaries of rocprofv3 output for the bench workload: the tables kept as text beside the
measurements (kernel trace, HBM traffic per launch class, MFMA use per kernel).

    python -m pspaqr_public_amd.profsum trace STATS_CSV TRACE_CSV [--steps S] [--bench-log LOG]
    python -m pspaqr_public_amd.profsum traffic --fetch-calib CSV --write-calib CSV --calib-mib 256 \\
        --fetch CSV --write CSV [--steps 1] [--json OUT]
    python -m pspaqr_public_amd.profsum mfma CSV [--steps 1]

STATS_CSV / TRACE_CSV: `rocprofv3 --kernel-trace --stats --output-format csv` results
(`*_kernel_stats.csv`, `*_kernel_trace.csv`).  The traffic and MFMA inputs are the
`*_counter_collection.csv` of `rocprofv3 --pmc ... --output-format csv` runs, one counter set
per run.  Each command prints a Markdown table on stdout; `traffic --json` also writes the
per-class bytes that bench.py reads for its `roofline.traffic`.
"""
import argparse
import collections
import csv
import json
import re
import sys

# Launch classes whose kernels serve that class only (the shared panel, larfb, trsm, copy and
# geqp3 kernels are left out): kernel-name patterns, matched after the namespace.
CLASSES = {
    "precond_solve": r"^(k_vec_|k_convert_)",
    "tpmqrt_chain": r"^k_tpmqrt",
    "tpqrt_chain": r"^k_tpqrt",
    "sparsify_apply_merge": r"^k_spars",
    "rank_sketch": r"^(k_sketch|k_colnorm|k_tail_select|k_rank_features|k_mlp_predict|k_gauss_bf16)",
    "spmv": r"^k_spmv",
    "gmres_orth": r"^(k_mdot|k_cgs)",
}


def short(name):
    """`void spaqr::k_x<float, 64>(args)` -> `k_x<float, 64>`."""
    n = name.strip().replace("(anonymous namespace)::", "")
    if n.startswith("void "):
        n = n[5:]
    depth, cut = 0, len(n)
    for i, ch in enumerate(n):  # the argument list starts at the first '(' outside template brackets
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    n = n[:cut]
    return n.replace("spaqr::", "")


def kclass(name):
    s = short(name)
    for c, pat in CLASSES.items():
        if re.search(pat, s):
            return c
    return None


def read_counters(path):
    """Per dispatch: (kernel, {counter: value}); counters summed over the rows rocprofv3 writes
    for one dispatch (per XCD / per instance)."""
    disp = collections.OrderedDict()
    with open(path) as f:
        for r in csv.DictReader(f):
            key = r["Dispatch_Id"]
            if key not in disp:
                disp[key] = (r["Kernel_Name"], collections.defaultdict(float))
            disp[key][1][r["Counter_Name"]] += float(r["Counter_Value"])
    return list(disp.values())


def cmd_trace(a):
    stats = list(csv.DictReader(open(a.stats)))
    total = sum(int(r["TotalDurationNs"]) for r in stats)
    print("| kernel | calls | total ms | average us | % |")
    print("|---|---:|---:|---:|---:|")
    for r in stats[: a.top]:
        print("| `%s` | %s | %.1f | %.1f | %.2f |" % (short(r["Name"]), r["Calls"], int(r["TotalDurationNs"]) / 1e6,
                                                   float(r["AverageNs"]) / 1e3, float(r["Percentage"])))
    print()
    print("Total kernel time %.1f ms over the run (%.1f ms per step over %d steps)." % (total / 1e6, total / 1e6 / a.steps,
                                                                                       a.steps))
    if a.trace:
        rows = list(csv.DictReader(open(a.trace)))
        solve = [r for r in rows if kclass(r["Kernel_Name"]) == "precond_solve"]
        napp = sum(1 for r in solve if "k_convert_d2s" in r["Kernel_Name"])
        ns = sum(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in solve)
        if napp:
            print("Preconditioner applications: %d; their kernels (`k_vec_*`, `k_convert_*`) total %.1f ms, %.3f ms per "
                  "application." % (napp, ns / 1e6, ns / 1e6 / napp))
    if a.bench_log:
        for line in open(a.bench_log):
            if line.startswith("{"):
                d = json.loads(line)
                rf = d.get("roofline", {})
                print("bench.py's own line in the profiled run: %.1f ms per step (factor %.1f ms, %s GMRES iterations); "
                      "dominant class `%s`: average launch %.3f ms over %s launches (HIP events)." %
                      (d["ms_per_step"], 1e3 * d.get("factor_s", 0.0), d.get("gmres_iters"), rf.get("kernel"),
                       rf.get("avg_launch_ms", 0.0), rf.get("launches")))


def calib_scale(path, counter, mib):
    """Bytes per counter unit: the largest dispatch of the calibration run is the copy of `mib` MiB."""
    disp = read_counters(path)
    kib = max(c.get(counter, 0.0) for _, c in disp)
    return (mib * 1024.0) / kib, kib


def cmd_traffic(a):
    fs, fk = calib_scale(a.fetch_calib, "FETCH_SIZE", a.calib_mib)
    ws, wk = calib_scale(a.write_calib, "WRITE_SIZE", a.calib_mib)
    fetch = read_counters(a.fetch)
    write = read_counters(a.write)
    per = collections.OrderedDict()
    for kind, disp, key, scale in (("r", fetch, "FETCH_SIZE", fs), ("w", write, "WRITE_SIZE", ws)):
        for name, c in disp:
            e = per.setdefault(short(name), {"n_r": 0, "n_w": 0, "r": 0.0, "w": 0.0, "cls": kclass(name)})
            e["n_" + kind] += 1
            e[kind] += c.get(key, 0.0) * scale * 1024.0
    cls = collections.OrderedDict((c, {"bytes": 0.0, "dispatches": 0}) for c in CLASSES)
    for k, e in per.items():
        if e["cls"]:
            cls[e["cls"]]["bytes"] += e["r"] + e["w"]
            cls[e["cls"]]["dispatches"] += e["n_r"]
    print("Calibration: a %d MiB copy counts FETCH_SIZE %.0f KiB and WRITE_SIZE %.0f KiB: FETCH_SIZE x %.3f, "
          "WRITE_SIZE x %.3f." % (a.calib_mib, fk, wk, fs, ws))
    print()
    print("| class | HBM bytes per step (counters) | kernel dispatches per step |")
    print("|---|---:|---:|")
    for c, v in cls.items():
        print("| %s | %.2f GB | %.0f |" % (c, v["bytes"] / a.steps / 1e9, v["dispatches"] / a.steps))
    print()
    print("| kernel | dispatches | read MB / dispatch | written MB / dispatch | total MB / dispatch |")
    print("|---|---:|---:|---:|---:|")
    top = sorted(per.items(), key=lambda kv: -(kv[1]["r"] + kv[1]["w"]))[: a.top]
    for k, e in top:
        n = max(e["n_r"], 1)
        print("| `%s` | %d | %.2f | %.2f | %.2f |" % (k, e["n_r"], e["r"] / n / 1e6, e["w"] / max(e["n_w"], 1) / 1e6,
                                                     (e["r"] / n + e["w"] / max(e["n_w"], 1)) / 1e6))
    if a.json:
        out = {
            "workload": a.workload,
            "fetch_scale": round(fs, 4),
            "write_scale": round(ws, 4),
            "note": "HBM bytes of one step, summed over every dispatch of the kernels that serve only that launch "
                    "class (FETCH_SIZE x fetch_scale + WRITE_SIZE x write_scale, KiB counters)",
            "bytes_per_step": {c: int(round(v["bytes"] / a.steps)) for c, v in cls.items()},
        }
        with open(a.json, "w") as f:
            json.dump(out, f, indent=1)
            f.write("\n")


def cmd_mfma(a):
    disp = read_counters(a.csv)
    agg = collections.OrderedDict()
    for name, c in disp:
        e = agg.setdefault(short(name), collections.defaultdict(float))
        e["n"] += 1
        for k, v in c.items():
            e[k] += v
    rows = []
    for k, e in agg.items():
        fl = 512.0 * (e.get("SQ_INSTS_VALU_MFMA_MOPS_F32", 0.0) + e.get("SQ_INSTS_VALU_MFMA_MOPS_BF16", 0.0))
        if fl <= 0:
            continue
        secs = e.get("GRBM_GUI_ACTIVE", 0.0) / 8.0 / 2.4e9  # GRBM counts on every XCD at 2.4 GHz
        busy = e.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0) / max(4.0 * e.get("SQ_BUSY_CU_CYCLES", 0.0), 1.0)
        bf16 = e.get("SQ_INSTS_VALU_MFMA_MOPS_BF16", 0.0) > e.get("SQ_INSTS_VALU_MFMA_MOPS_F32", 0.0)
        peak = 2500.0 if bf16 else 157.3
        tf = fl / secs / 1e12 if secs > 0 else 0.0
        rows.append((fl, k, int(e["n"]), tf, 100.0 * tf / peak, 100.0 * busy))
    rows.sort(reverse=True)
    print("| kernel | dispatches | MFMA GFLOP | TFLOP/s while running | % of peak | MFMA busy |")
    print("|---|---:|---:|---:|---:|---:|")
    for fl, k, n, tf, pp, busy in rows[: a.top]:
        print("| `%s` | %d | %.1f | %.2f | %.1f%% | %.1f%% |" % (k, n, fl / 1e9 / a.steps, tf, pp, busy))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m pspaqr_public_amd.profsum")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("trace")
    t.add_argument("stats")
    t.add_argument("trace", nargs="?")
    t.add_argument("--steps", type=int, default=1)
    t.add_argument("--top", type=int, default=40)
    t.add_argument("--bench-log")
    h = sub.add_parser("traffic")
    h.add_argument("--fetch-calib", required=True)
    h.add_argument("--write-calib", required=True)
    h.add_argument("--calib-mib", type=int, default=256)
    h.add_argument("--fetch", required=True)
    h.add_argument("--write", required=True)
    h.add_argument("--steps", type=int, default=1)
    h.add_argument("--top", type=int, default=25)
    h.add_argument("--json")
    h.add_argument("--workload", default="convdiff3d_64, FP32 fronts, learned ranks (bf16 sketches): one bench.py step "
                                          "(factorize + flexible GMRES)")
    m = sub.add_parser("mfma")
    m.add_argument("csv")
    m.add_argument("--steps", type=int, default=1)
    m.add_argument("--top", type=int, default=30)
    a = ap.parse_args(argv)
    {"trace": cmd_trace, "traffic": cmd_traffic, "mfma": cmd_mfma}[a.cmd](a)
    return 0


if __name__ == "__main__":
    sys.exit(main())
