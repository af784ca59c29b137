#!/usr/bin/env python3
"""Benchmark: subproblems/s of the path-decomposition engine on one MI355X.

A step is one complete solve of the workload through the C-ABI (nr_solve):
K10 (45 edges, p = 0.5, 0 -> 9) enumerated in full, 14 866 064 subproblems,
level order.  Graph tables and queues stay resident in HBM between steps (the
library caches its per-device workspace), so a timed step is the solve itself.

Prints one JSON line with the metric, a roofline object for the expand
kernel (HBM bound: algorithmic bytes per subproblem = 2 * (16*W + 8), one
write when the item is created and one read when it is expanded), the CPU
baselines (the oracle on the host cores, and the reference's own CPU heap
program seq/seq.cpp built from source into oracle/_ref, same workload), the
batched workload (4096 Erdos-Renyi instances) with its own roofline, and the
metric's second part, the certified gap ESP at a 60 s budget:
  * K15 and K25 at delta = 1e-12 (the configured workload): the plain path
    rule against the alternating-sweep rule (the rule's gain) and against
    the learned path choice on that rule (the trained model's gain);
  * K15 untruncated (delta = 0) in max-mass order, where the queue capacity
    binds: plain, alternating and learned;
  * a 10x10 grid and a Barabasi-Albert graph (graphs the models were not
    trained on) at a 1 s budget.
Both learned networks are trained here on the GPU by the engine's own
kernels from logs the engine writes (relaibilitygpu_amd/learn.py).
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
    "p2p": {"graph": "p2p-Gn08C.txt", "src": 0, "dst": 17, "eps": 1e-6, "cap": 1 << 22,
            "kernel": "k_expand_big",
            "desc": ("p2p-Gn08C (SNAP, N=6301, 20777 arcs, p=0.5) src 0 -> dst 17, truncation "
                     "1e-6, level order, one wave per subproblem")},
    "er4096": {"batch": 4096, "eps": 1e-12, "cap": 1 << 27,
               "desc": ("4096 seeded Erdos-Renyi instances (n 12..40, mean degree 4..8, "
                        "p 0.80..0.99, terminals at distance >= 3), truncation 1e-12, "
                        "one batched solve")},
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


def copy_peak_child(mib: int = 1024, reps: int = 10) -> float:
    """HBM bandwidth a plain device-to-device copy reaches on this GPU (read
    + write bytes over the median of reps copies, HIP events); runs in a
    process of its own (torch's HIP runtime does not share a process with
    the engine's)."""
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    return 2 * n * 4 / (statistics.median(ms) / 1e3) / 1e9


def copy_peak_gbs():
    """The measured copy peak (copy_peak_child in a child process), or None
    if the child fails."""
    try:
        out = subprocess.run([sys.executable, os.path.abspath(__file__), "--copy-peak"],
                             capture_output=True, text=True, timeout=300, check=True).stdout
        return float(json.loads(out.strip().splitlines()[-1])["copy_peak_gbs"])
    except (subprocess.SubprocessError, ValueError, KeyError, IndexError, OSError):
        return None


def valu_roofline(kernel: str, workload: str, subproblems_per_s: float):
    """Instruction roofline of the expand kernel: VALU wave instructions per
    subproblem (rocprofv3 SQ_INSTS_VALU over whole solves, committed under
    profiles/) times the measured subproblems/s, against the chip's VALU
    issue peak (256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64
    instruction), or None without a counters summary."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        c = json.load(f)
    achieved = c["valu_per_subproblem"] * subproblems_per_s
    return {"bound": "valu", "achieved": achieved, "peak": c["valu_peak_per_s"],
            "unit": "VALU wave instructions/s", "frac": achieved / c["valu_peak_per_s"],
            "valu_per_subproblem": c["valu_per_subproblem"],
            "salu_per_subproblem": c["salu_per_subproblem"], "kernel": kernel,
            "basis": c["valu_peak_basis"]}


def cpu_baseline(wl: dict) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_ffi import ORDER_DFS, OracleGraph

    g = OracleGraph.load(os.path.join(ROOT, "tests", "golden", "graphs", wl["graph"]))
    threads = min(os.cpu_count() or 1, 16)
    runs = []
    nodes = 0
    for _ in range(5):
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
                   f"cpu_reference.cpp, depth-first order, {threads} threads, median of 5 "
                   f"= {med:.1f} ms); 1 thread: {r1['elapsed_ms']:.1f} ms = "
                   f"{nodes / (r1['elapsed_ms'] / 1e3):.4g} subproblems/s; CPU {model}"),
        "single_thread_value": nodes / (r1["elapsed_ms"] / 1e3),
    }


def reference_cpu(wl: dict, runs: int = 5) -> dict:
    """The reference's own CPU program on the workload: seq/seq.cpp (max-mass
    heap, FP64, one thread), compiled from the reference sources into
    oracle/_ref by `make -C oracle ref`.  `runs` copies run at once, one core
    each; the median of their own timers is reported."""
    import re

    exe = os.path.join(ROOT, "oracle", "_ref", "ref_seq_heap")
    if not os.path.exists(exe):
        return {"absent": "oracle/_ref/ref_seq_heap was not built (reference sources absent "
                          "where the build ran)"}
    path = os.path.join(ROOT, "tests", "golden", "graphs", wl["graph"])
    procs = [subprocess.Popen([exe, path, str(wl["src"]), str(wl["dst"])],
                              stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
             for _ in range(runs)]
    secs, rel, paths = [], None, None
    for p in procs:
        out, _ = p.communicate(timeout=600)
        m = re.search(r"^Time:\s+([0-9.eE+-]+) seconds", out, flags=re.M)
        if p.returncode != 0 or not m:
            return {"failed": out[-400:]}
        secs.append(float(m.group(1)))
        rel = float(re.search(r"^Reliability:\s+([0-9.]+)", out, flags=re.M).group(1))
        paths = int(re.search(r"^Paths enumerated:\s+(\d+)", out, flags=re.M).group(1))
    med = statistics.median(secs)
    return {"value": wl["subproblems"] / med, "unit": "subproblems/s", "cores": 1,
            "kind": "reference",
            "sample": (f"whole workload by the reference's seq/seq.cpp (max-mass heap, FP64, one "
                       f"thread) built from source, {runs} concurrent runs one core each, median "
                       f"{med:.2f} s (runs {', '.join(f'{x:.2f}' for x in secs)}); value = the "
                       f"workload's {wl['subproblems']} subproblems / median time"),
            "R": rel, "paths": paths, "median_s": med}


def er_graphs(count: int):
    from relaibilitygpu_amd import workloads

    return [workloads.er_instance(i) for i in range(count)]


def cpu_baseline_batch(insts, eps: float, sample: int) -> dict:
    """The oracle on the first `sample` instances, one instance per host
    thread (single-threaded depth-first solves in a thread pool)."""
    from concurrent.futures import ThreadPoolExecutor

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle_ffi import ORDER_DFS, OracleGraph
    from relaibilitygpu_amd import workloads

    threads = min(os.cpu_count() or 1, 16)
    graphs = []
    for n, edges, probs, s, t in insts[:sample]:
        rp, ci, ei = workloads.to_csr(n, edges)
        graphs.append((OracleGraph.from_csr(rp, ci, ei, probs), s, t))

    def one(item):
        g, s, t = item
        return g.solve(s, t, order=ORDER_DFS, threads=1, trunc_eps=eps)["nodes"]

    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        nodes = sum(ex.map(one, graphs))
    dt = time.perf_counter() - t0
    return {
        "value": nodes / dt,
        "unit": "subproblems/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"first {sample} instances ({nodes} subproblems) by the oracle "
                   f"(oracle/cpu_reference.cpp, depth-first, one instance per thread, "
                   f"{threads} threads) in {dt:.2f} s"),
    }


def run_batch(args, wl: dict, steps: int | None = None, warmup: int | None = None) -> dict:
    import relaibilitygpu_amd as nr
    from relaibilitygpu_amd import workloads

    insts = er_graphs(wl["batch"])
    graphs = []
    wmax = 1
    for n, edges, probs, s, t in insts:
        rp, ci, ei = workloads.to_csr(n, edges)
        g = nr.Graph.from_csr(rp, ci, ei, probs, s, t)
        wmax = max(wmax, g.info().mask_words)
        graphs.append(g)
    cfg = nr.make_config(queue_capacity=wl["cap"], truncation_eps=wl["eps"], order="level")
    batch = nr.Batch(graphs, cfg)  # tables go to the device with the first (warmup) call
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    for w in range(max(1, warmup)):
        r = batch.run()
        print(f"warmup {w}: {r['elapsed_ms'][0]:.1f} ms, reruns {int(r['rerun'].sum())}",
              file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    arrays = []
    for k in range(steps):
        arrays.append(batch.run())
        print(f"step {k}: {arrays[-1]['elapsed_ms'][0]:.1f} ms", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    for res in arrays[1:]:
        assert (res["nodes"] == arrays[0]["nodes"]).all()
    runs = [[{f: res[f][i].item() for f in res.dtype.names} for i in range(len(res))]
            for res in arrays[:1]]
    nodes = sum(r["nodes"] for r in runs[0])
    wall_s = t1 - t0
    tcfg = nr.make_config(queue_capacity=wl["cap"], truncation_eps=wl["eps"], order="level",
                          time_launches=True)
    timed = nr.solve_batch(graphs, tcfg)[0]
    bytes_per_sub = 2 * (16 * wmax + 8)
    achieved = nodes * bytes_per_sub / (timed["kernel_ms"] / 1e3) / 1e9
    r0 = runs[0]
    return {
        "value": nodes * steps / wall_s,
        "ms_per_step": wall_s * 1e3 / steps,
        "config": {
            "workload": wl["desc"],
            "instances": len(graphs),
            "subproblems_per_step": nodes,
            "paths_per_step": sum(r["paths"] for r in r0),
            "rerun_instances": sum(r["rerun"] for r in r0),
            "overflowed_instances": sum(1 for r in r0 if r["overflowed"]),
            "truncation_eps": wl["eps"],
            "queue_capacity": wl["cap"],
            "order": "level",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles("k_expand_batch", "er4096"),
            "kernel": "k_expand_batch",
            "bytes_per_subproblem": bytes_per_sub,
            "avg_launch_ms": timed["kernel_ms"] / max(1, timed["kernel_launches"]),
            "launches_per_step": timed["kernel_launches"],
        },
        "gpu_elapsed_ms_per_step": sum(r[0]["elapsed_ms"] for r in runs) / len(runs),
        "kernel_ms_per_step": timed["kernel_ms"],
        "sum_R": sum(r["R"] for r in r0),
        "max_ESP": max(r["ESP"] for r in r0),
        "_insts": insts,
    }


def train_models():
    """Both learned networks, trained here on the GPU (relaibilitygpu_amd/
    learn.py): the order scorer on children logged from K10, the path
    chooser on choice logs of K10 (full) and K11-K13 (truncated), its margin
    picked on validation graphs (K14, a 7x7 grid, BA(100)) that no leg below
    evaluates.  Returns (Model, training summary)."""
    import relaibilitygpu_amd as nr
    from relaibilitygpu_amd import learn

    order_params, order_info = learn.train_order_scorer(alpha=0.5)
    ch_params, margin, ch_info = learn.train_chooser()
    ch_info = {k: v for k, v in ch_info.items() if k != "trials"}
    model = nr.Model(order_params, alpha=0.5, chooser=ch_params, margin=margin)
    return model, {"order_scorer": order_info, "path_chooser": ch_info}


def _row(r):
    return {"ESP": r["ESP"], "R": r["R"], "nodes": r["nodes"], "elapsed_ms": r["elapsed_ms"],
            "deadline_hit": r["deadline_hit"], "folded": r["folded"],
            "overflowed": r["overflowed"]}


def _ratio(a, b):
    return a / b if b > 0 else None


def esp_leg(model, budget_s: float = 60.0) -> dict:
    """Certified gap ESP at a time budget on K15 and K25 (p = 0.5, capacity
    96 777 216 items), every row with the same budget:
      delta_1e-12 (the configured workload): plain path rule (level, max-mass
        order), alternating-sweep rule (period from pilot solves at 1e-9),
        learned path choice on that rule (level order; and with the learned
        order);
      delta_0 (K15; nothing truncated, the capacity binds in max-mass order):
        plain, alternating and learned choice in max-mass order, and the
        learned choice in the learned order.
    Ratios: plain over alternating is the rule's gain, alternating over
    learned the trained model's own gain, plain over learned both."""
    import relaibilitygpu_amd as nr

    G = os.path.join(ROOT, "tests", "golden", "graphs")
    cap = 96777216
    out = {"budget_s": budget_s, "queue_capacity": cap}
    for name, n in (("K15", 15), ("K25", 25)):
        g = nr.Graph.load(os.path.join(G, f"COMP{n}-05.txt"), 0, n - 1)
        period, pilot = nr.tune_sweep_period(g, pilot_eps=1e-9, max_period=6)
        res = {"sweep_period": period, "pilot_esp": pilot}
        kw = dict(truncation_eps=1e-12, queue_capacity=cap, time_budget_s=budget_s)
        alt = dict(kw, rule="alternate", sweep_period=period)
        lrn = dict(kw, rule="learned", sweep_period=period)
        rows = {"plain_level": nr.solve(g, **kw), "plain_mass": nr.solve(g, order="mass", **kw),
                "alternate_level": nr.solve(g, **alt),
                "alternate_mass": nr.solve(g, order="mass", **alt),
                "learned_choice_level": nr.solve_learned(g, model, order="level", **lrn),
                "learned_choice_mass": nr.solve_learned(g, model, order="mass", **lrn),
                "learned_choice_learned_order": nr.solve_learned(g, model, **lrn)}
        d = {k: _row(v) for k, v in rows.items()}
        res["delta_1e-12"] = d
        res["ESP_plain_over_alternate"] = _ratio(d["plain_level"]["ESP"],
                                                 d["alternate_level"]["ESP"])
        res["ESP_alternate_over_learned_choice"] = _ratio(d["alternate_level"]["ESP"],
                                                          d["learned_choice_level"]["ESP"])
        res["ESP_plain_over_learned_choice"] = _ratio(d["plain_level"]["ESP"],
                                                      d["learned_choice_level"]["ESP"])
        print(f"esp {name} delta 1e-12 (period {period}): "
              + ", ".join(f"{k} {v['ESP']:.3e} ({v['elapsed_ms']:.0f} ms)" for k, v in d.items()),
              file=sys.stderr, flush=True)
        if n == 15:
            kw0 = dict(truncation_eps=0.0, queue_capacity=cap, time_budget_s=budget_s)
            rows0 = {"plain_mass": nr.solve(g, order="mass", **kw0),
                     "alternate_mass": nr.solve(g, order="mass", rule="alternate",
                                                sweep_period=period, **kw0),
                     "learned_choice_mass": nr.solve_learned(g, model, order="mass",
                                                             rule="learned", sweep_period=period,
                                                             **kw0),
                     "learned_choice_learned_order": nr.solve_learned(
                         g, model, rule="learned", sweep_period=period, **kw0)}
            d0 = {k: _row(v) for k, v in rows0.items()}
            res["delta_0"] = d0
            res["delta_0_ESP_plain_over_alternate"] = _ratio(d0["plain_mass"]["ESP"],
                                                             d0["alternate_mass"]["ESP"])
            res["delta_0_ESP_alternate_over_learned_choice"] = _ratio(
                d0["alternate_mass"]["ESP"], d0["learned_choice_mass"]["ESP"])
            print(f"esp {name} delta 0 (period {period}): "
                  + ", ".join(f"{k} {v['ESP']:.3e} ({v['elapsed_ms']:.0f} ms, folded "
                              f"{v['folded']})" for k, v in d0.items()),
                  file=sys.stderr, flush=True)
        out[name] = res
    return out


def sparse_leg(model, budget_s: float = 1.0) -> dict:
    """Graphs the models were not trained on: a 10x10 grid (180 edges,
    p ~ U[0.85, 0.95] seeded, corner to corner) and Barabasi-Albert n = 150,
    m = 2, p = 0.95 (0 -> 149), at delta = 1e-12 and an equal time budget:
    plain path rule, alternating rule (period from pilot solves) and the
    learned path choice on it, in level and max-mass order.  No closed form
    exists; every bracket must contain R, so all brackets must intersect."""
    import relaibilitygpu_amd as nr
    from relaibilitygpu_amd import learn

    graphs = [("grid10x10", learn.grid(10, 10, 0x6121D), 1e-6),
              ("ba150", learn.barabasi_albert(150, 0xBA150), 1e-9)]
    out = {"budget_s": budget_s, "delta": 1e-12, "queue_capacity": 96777216}
    for name, g, pilot_eps in graphs:
        period, _ = nr.tune_sweep_period(g, pilot_eps=pilot_eps, max_period=6)
        kw = dict(truncation_eps=1e-12, queue_capacity=96777216, time_budget_s=budget_s)
        alt = dict(kw, rule="alternate", sweep_period=period)
        lrn = dict(kw, rule="learned", sweep_period=period)
        rows = {"plain_level": nr.solve(g, **kw), "plain_mass": nr.solve(g, order="mass", **kw),
                "alternate_level": nr.solve(g, **alt),
                "alternate_mass": nr.solve(g, order="mass", **alt),
                "learned_choice_level": nr.solve_learned(g, model, order="level", **lrn),
                "learned_choice_mass": nr.solve_learned(g, model, order="mass", **lrn)}
        lo = max(r["R"] for r in rows.values())
        hi = min(r["R"] + r["ESP"] for r in rows.values())
        d = {k: _row(v) for k, v in rows.items()}
        res = {"sweep_period": period, "rows": d,
               "brackets_intersect": bool(lo <= hi * (1 + 1e-12)),
               "ESP_plain_over_alternate": _ratio(d["plain_level"]["ESP"],
                                                  d["alternate_level"]["ESP"]),
               "ESP_alternate_over_learned_choice": _ratio(d["alternate_level"]["ESP"],
                                                           d["learned_choice_level"]["ESP"]),
               "ESP_alternate_mass_over_learned_choice_mass": _ratio(
                   d["alternate_mass"]["ESP"], d["learned_choice_mass"]["ESP"])}
        out[name] = res
        print(f"sparse {name} (period {period}): "
              + ", ".join(f"{k} {v['ESP']:.3e} ({v['elapsed_ms']:.0f} ms)" for k, v in d.items()),
              file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="k10", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample", type=int, default=512,
                    help="batched workloads: instances timed on the CPU")
    ap.add_argument("--no-esp", action="store_true",
                    help="skip the learned-vs-plain certified-gap leg")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched (4096 Erdos-Renyi instances) object")
    ap.add_argument("--esp-budget", type=float, default=60.0)
    ap.add_argument("--sparse-budget", type=float, default=1.0,
                    help="time budget of the grid / Barabasi-Albert certified-gap leg")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration)")
    ap.add_argument("--copy-peak", action="store_true",
                    help="only measure the device copy bandwidth (GB/s)")
    ap.add_argument("--no-copy-peak", action="store_true",
                    help="skip the copy-bandwidth measurement (profiled runs)")
    args = ap.parse_args()
    if args.copy_peak:
        print(json.dumps({"copy_peak_gbs": copy_peak_child()}))
        return 0
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
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not supported "
              "(the solve would shard by subtree: frontier items dealt to the GPUs, five "
              "scalars all-reduced at the end)", file=sys.stderr)
        return 2

    import relaibilitygpu_amd as nr

    wl = WORKLOADS[args.workload]
    if "batch" in wl:
        b = run_batch(args, wl)
        insts = b.pop("_insts")
        cp = None if args.no_copy_peak else copy_peak_gbs()
        b["roofline"]["copy_peak"] = cp
        b["roofline"]["frac_of_copy_peak"] = b["roofline"]["achieved"] / cp if cp else None
        line = {"metric": METRIC, "value": b.pop("value"), "unit": "subproblems/s",
                "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": b.pop("ms_per_step"), "higher_is_better": True,
                "scaling": "strong", "vs_baseline": None, "dtype": "u64+f64",
                "data": "synthetic"}
        line.update(b)
        if not args.no_cpu:
            line["cpu_baseline"] = cpu_baseline_batch(insts, wl["eps"], args.cpu_sample)
        print(json.dumps(line), flush=True)
        return 0
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
    kernel = wl.get("kernel", "k_expand_level")
    traffic = traffic_from_profiles(kernel, args.workload)
    copy_peak = None if args.no_copy_peak else copy_peak_gbs()
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
            "kernel": kernel,
            "bytes_per_subproblem": bytes_per_sub,
            "avg_launch_ms": kernel_ms / launches,
            "launches_per_step": launches / len(timed),
            "copy_peak": copy_peak,
            "frac_of_copy_peak": achieved / copy_peak if copy_peak else None,
        },
        "roofline_valu": valu_roofline(kernel, args.workload, value),
        "gpu_elapsed_ms_per_step": sum(r["elapsed_ms"] for r in results) / args.steps,
        "kernel_ms_per_step": kernel_ms / len(timed),
        "R": results[-1]["R"],
        "ESP": results[-1]["ESP"],
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(wl)
        line["cpu_baseline_reference"] = reference_cpu(dict(wl, subproblems=nodes))
    if not args.no_batch:
        bw = WORKLOADS["er4096"]
        b = run_batch(args, bw, steps=3, warmup=1)
        insts = b.pop("_insts")
        b["roofline"]["copy_peak"] = copy_peak
        b["roofline"]["frac_of_copy_peak"] = (b["roofline"]["achieved"] / copy_peak
                                              if copy_peak else None)
        if not args.no_cpu:
            b["cpu_baseline"] = cpu_baseline_batch(insts, bw["eps"], args.cpu_sample)
        b["unit"] = "subproblems/s"
        line["batch_er4096"] = b
        print(f"batch er4096: {b['value']:.4g} subproblems/s, {b['ms_per_step']:.1f} ms per "
              f"batch", file=sys.stderr, flush=True)
    if not args.no_esp:
        model, info = train_models()
        line["esp_at_budget"] = esp_leg(model, args.esp_budget)
        line["esp_at_budget"]["training"] = info
        line["esp_sparse_graphs"] = sparse_leg(model, args.sparse_budget)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
