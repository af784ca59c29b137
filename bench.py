#!/usr/bin/env python3
"""Benchmark: subproblems/s of the path-decomposition engine on one MI355X.

A step is one complete solve of the workload through the C-ABI (nr_solve):
K10 (45 edges, p = 0.5, 0 -> 9) enumerated in full, 14 866 064 subproblems,
level order.  Graph tables and queues stay resident in HBM between steps (the
library caches its per-device workspace), so a timed step is the solve itself.

Prints one JSON line with the metric, a roofline object for the expand
kernel (HBM bound: algorithmic bytes per subproblem = 2 * (16*W + 8), one
write when the item is created and one read when it is expanded) and the CPU
baseline (the oracle on the host cores, same workload).
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
WORKLOADS = {
    "k10": {"graph": "COMP10-05.txt", "src": 0, "dst": 9, "eps": 0.0, "cap": 1 << 22,
            "desc": "K10 (N=10, E=45, p=0.5) src 0 -> dst 9, full enumeration, level order"},
    "k8p09": {"graph": "COMP8-09.txt", "src": 0, "dst": 7, "eps": 0.0, "cap": 1 << 16,
              "desc": "K8 (N=8, E=28, p=0.9) src 0 -> dst 7, full enumeration, level order"},
}
METRIC = "subproblems/sec; certified gap (ESP) at 60 s, learned vs plain order"


def traffic_from_profiles(kernel: str, workload: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, calibrated against a copy kernel), if a summary exists."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_baseline(wl: dict) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_ffi import ORDER_DFS, OracleGraph

    g = OracleGraph.load(os.path.join(ROOT, "tests", "golden", "graphs", wl["graph"]))
    threads = min(os.cpu_count() or 1, 16)
    runs = []
    nodes = 0
    for _ in range(3):
        r = g.solve(wl["src"], wl["dst"], order=ORDER_DFS, threads=threads,
                    trunc_eps=wl["eps"])
        runs.append(r["elapsed_ms"])
        nodes = r["nodes"]
    r1 = g.solve(wl["src"], wl["dst"], order=ORDER_DFS, threads=1, trunc_eps=wl["eps"])
    med = statistics.median(runs)
    try:
        model = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        model = next((l.split(":", 1)[1].strip() for l in model.splitlines()
                      if l.startswith("Model name")), "unknown")
    except Exception:
        model = "unknown"
    return {
        "value": nodes / (med / 1e3),
        "unit": "subproblems/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"whole workload ({nodes} subproblems) by the oracle (oracle/"
                   f"cpu_reference.cpp, depth-first order, {threads} threads, median of 3 "
                   f"= {med:.1f} ms); 1 thread: {r1['elapsed_ms']:.1f} ms = "
                   f"{nodes / (r1['elapsed_ms'] / 1e3):.4g} subproblems/s; CPU {model}"),
        "single_thread_value": nodes / (r1["elapsed_ms"] / 1e3),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="k10", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        n = args.copy_probe * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n * 4, "launches": 3}))
        return 0
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not part of this "
              "build (the path would shard by subtree, see DESIGN.md)", file=sys.stderr)
        return 2

    import relaibilitygpu_amd as nr

    wl = WORKLOADS[args.workload]
    path = os.path.join(ROOT, "tests", "golden", "graphs", wl["graph"])
    g = nr.Graph.load(path, wl["src"], wl["dst"])
    info = g.info()
    cfg = nr.make_config(queue_capacity=wl["cap"], truncation_eps=wl["eps"], order="level")
    for _ in range(args.warmup):
        nr.solve(g, cfg)
    results = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        results.append(nr.solve(g, cfg))
    t1 = time.perf_counter()
    nodes = results[0]["nodes"]
    assert all(r["nodes"] == nodes and r["paths"] == results[0]["paths"] for r in results)
    wall_s = t1 - t0
    value = nodes * args.steps / wall_s
    # Roofline: the same solve once more with every expand launch bracketed
    # by HIP events on the library's stream (kept out of the timed steps:
    # the events add a few microseconds between launches).
    tcfg = nr.make_config(queue_capacity=wl["cap"], truncation_eps=wl["eps"], order="level",
                          time_launches=True)
    timed = [nr.solve(g, tcfg) for _ in range(3)]
    kernel_ms = sum(r["kernel_ms"] for r in timed)
    launches = sum(r["kernel_launches"] for r in timed)
    bytes_per_sub = 2 * (16 * info.mask_words + 8)
    achieved = nodes * len(timed) * bytes_per_sub / (kernel_ms / 1e3) / 1e9
    traffic = traffic_from_profiles("k_expand_level", args.workload)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "subproblems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_s * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "u64+f64",
        "data": "synthetic",
        "config": {
            "workload": wl["desc"],
            "graph": wl["graph"],
            "N": info.n,
            "E": info.e,
            "subproblems_per_step": nodes,
            "paths_per_step": results[0]["paths"],
            "levels": results[0]["iterations"],
            "order": "level",
            "queue_capacity": wl["cap"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_expand_level",
            "bytes_per_subproblem": bytes_per_sub,
            "avg_launch_ms": kernel_ms / launches,
            "launches_per_step": launches / len(timed),
        },
        "gpu_elapsed_ms_per_step": sum(r["elapsed_ms"] for r in results) / args.steps,
        "kernel_ms_per_step": kernel_ms / len(timed),
        "R": results[-1]["R"],
        "ESP": results[-1]["ESP"],
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(wl)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
