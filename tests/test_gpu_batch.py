"""GPU parity of the batched engine (nr_solve_batch) against the CPU oracle.

Every instance of a batch is an independent two-terminal problem; with an
unbounded queue its counts do not depend on the order in which items are
expanded, so per-instance paths / nodes / dead / truncated must equal the
oracle's exactly, and the per-instance tree depth equals the oracle's level
count.  R, ESP and D are exact 128-bit fixed-point sums on the device
(resolution 2^-96 per term), so they agree with the oracle's compensated
FP64 sums to 1e-12 relative.
"""
import numpy as np
import pytest

from conftest import graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from relaibilitygpu_amd import workloads as wl  # noqa: E402
from oracle_ffi import ORDER_DFS, ORDER_LEVEL, OracleGraph  # noqa: E402

REL = 1e-12


def er_pair(i):
    n, edges, probs, s, t = wl.er_instance(i)
    rp, ci, ei = wl.to_csr(n, edges)
    return (nr.Graph.from_csr(rp, ci, ei, probs, s, t), OracleGraph.from_csr(rp, ci, ei, probs),
            s, t)


def close(a, b, rel=REL, abs_=1e-25):
    return abs(a - b) <= rel * max(abs(a), abs(b)) + abs_


def check_inst(r, o, levels=True):
    assert (r["paths"], r["nodes"], r["dead"], r["truncated"]) == (
        o["paths"], o["nodes"], o["dead"], o["truncated"])
    if levels:
        assert r["iterations"] == o["iterations"]
    assert close(r["R"], o["R"]), (r["R"], o["R"])
    assert close(r["ESP"], o["ESP"]), (r["ESP"], o["ESP"])
    assert close(r["D"], o["D"]), (r["D"], o["D"])


@pytest.mark.parametrize("count,eps", [(16, 1e-6), (64, 1e-8)])
def test_er_batch_vs_oracle(count, eps):
    print(f"case ER batch count={count} eps={eps}", flush=True)
    pairs = [er_pair(i) for i in range(count)]
    res = nr.solve_batch([p[0] for p in pairs], truncation_eps=eps, queue_capacity=1 << 22)
    assert len(res) == count
    for (g, og, s, t), r in zip(pairs, res):
        o = og.solve(s, t, order=ORDER_LEVEL, trunc_eps=eps)
        check_inst(r, o)
        assert r["overflowed"] == 0 and r["deadline_hit"] == 0
        assert abs(r["R"] + r["ESP"] + r["D"] - 1.0) <= 1e-12


def test_er_stream_spill_vs_oracle():
    """A queue far smaller than the batch (64 workgroups x 256 slots; the
    batch's widest level is ~40x that): children that do not fit go to the
    spill stack and come back through the refill segment, so levels are
    interleaved across instances and depths.  Every instance still matches
    the oracle exactly, tree depth included."""
    print("case ER stream 512 eps=1e-7 cap=16384 grid=64", flush=True)
    count, eps = 512, 1e-7
    pairs = [er_pair(i) for i in range(count)]
    cfg = nr.make_config(truncation_eps=eps, queue_capacity=1 << 14)
    cfg.grid_blocks = 64
    res = nr.solve_batch([p[0] for p in pairs], cfg)
    for (g, og, s, t), r in zip(pairs, res):
        o = og.solve(s, t, order=ORDER_LEVEL, trunc_eps=eps)
        assert r["overflowed"] == 0
        check_inst(r, o)
    assert res[0]["max_queue_items"] <= (1 << 14) + (1 << 12)
    assert res[0]["kernel_launches"] > 2 * max(r["iterations"] for r in res)


def test_batch_matches_single_solves():
    """Mixed sizes and mask widths in one batch: files (N <= 32, one mask
    word) next to ER graphs up to N = 40 (two-word rows); each result equals
    the same instance solved alone."""
    print("case mixed batch vs single", flush=True)
    gs = [nr.Graph.load(graph_path("COMP8-09.txt"), 0, 7),
          nr.Graph.load(graph_path("test_kite.txt"), 0, 4),
          nr.Graph.load(graph_path("temp4x4.txt"), 0, 3),
          nr.Graph.load(graph_path("COMP9.txt"), 0, 8)]
    gs += [er_pair(i)[0] for i in range(4, 10)]
    eps = 1e-9
    res = nr.solve_batch(gs, truncation_eps=eps, queue_capacity=1 << 22)
    for g, r in zip(gs, res):
        s = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 22)
        assert (r["paths"], r["nodes"], r["dead"], r["truncated"], r["iterations"]) == (
            s["paths"], s["nodes"], s["dead"], s["truncated"], s["iterations"])
        assert close(r["R"], s["R"]) and close(r["ESP"], s["ESP"]) and close(r["D"], s["D"])


def test_batch_of_one_and_repeats():
    """One instance repeated: every copy gets identical results (fixed-point
    sums do not depend on the order of the atomics)."""
    print("case repeated instance", flush=True)
    g = nr.Graph.load(graph_path("COMP8-09.txt"), 0, 7)
    one = nr.solve_batch([g], truncation_eps=0.0, queue_capacity=1 << 20)[0]
    assert (one["paths"], one["nodes"], one["iterations"]) == (26830, 53660, 23)
    many = nr.solve_batch([g] * 37, truncation_eps=0.0, queue_capacity=1 << 22)
    for r in many:
        assert r["R"] == one["R"] and r["D"] == one["D"]
        assert (r["paths"], r["nodes"], r["dead"]) == (one["paths"], one["nodes"], one["dead"])


def test_batch_capacity_bracket():
    """A queue too small for the batch: overflowed children go to ESP and
    every instance still brackets its exact R."""
    print("case batch capacity bracket", flush=True)
    pairs = [er_pair(i) for i in range(8)]
    # 4 workgroups x 64-item segments: each instance overflows even alone
    cfg = nr.make_config(truncation_eps=1e-9, queue_capacity=256)
    cfg.grid_blocks = 4
    res = nr.solve_batch([p[0] for p in pairs], cfg)
    over = 0
    for (g, og, s, t), r in zip(pairs, res):
        # the oracle's own bracket at a finer eps; both contain the exact R
        o = og.solve(s, t, order=ORDER_DFS, threads=4, trunc_eps=1e-12)
        over += r["overflowed"]
        assert r["R"] <= (o["R"] + o["ESP"]) * (1 + 1e-12)
        assert r["R"] + r["ESP"] >= o["R"] * (1 - 1e-12)
        assert r["R"] <= o["R"] * (1 + 1e-12)  # the truncated run can only lose mass
        assert abs(r["R"] + r["ESP"] + r["D"] - 1.0) <= 1e-12
    assert over > 0


def test_batch_time_budget_fold():
    """A deadline that expires mid-run folds every pending item into its own
    instance's ESP; the bracket and R + ESP + D = 1 hold per instance."""
    print("case batch deadline", flush=True)
    gs = [nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9) for _ in range(16)]
    res = nr.solve_batch(gs, truncation_eps=0.0, queue_capacity=1 << 25, time_budget_s=0.001)
    R = 0.99596245255088434
    for r in res:
        assert r["deadline_hit"] == 1 and r["folded"] > 0
        assert r["R"] <= R * (1 + 1e-12) and r["R"] + r["ESP"] >= R * (1 - 1e-12)
        assert abs(r["R"] + r["ESP"] + r["D"] - 1.0) <= 1e-12


def test_batch_rejects():
    print("case batch rejects", flush=True)
    g = nr.Graph.load(graph_path("COMP8-09.txt"), 0, 7)
    with pytest.raises(nr.NRError):
        nr.solve_batch([g], order="mass")
    assert nr.solve_batch([], truncation_eps=0.0) == []


def test_batch_reuses_device_tables_until_a_graph_changes():
    """The library keeps a batch's instance tables on the device between calls
    with the same graphs; changing an instance's terminals (or passing other
    graphs) must rebuild them: every result equals a fresh single solve."""
    print("case batch table cache", flush=True)
    pairs = [er_pair(i) for i in range(40)]
    graphs = [p[0] for p in pairs]
    batch = nr.Batch(graphs, truncation_eps=1e-9, queue_capacity=1 << 22)
    first = batch.run()
    again = batch.run()
    assert (first["nodes"] == again["nodes"]).all() and (first["paths"] == again["paths"]).all()
    g0, o0, s0, t0 = pairs[0]
    n0 = g0.info().n
    g0.set_terminals(t0, (s0 + 1) % n0 if (s0 + 1) % n0 != t0 else (s0 + 2) % n0)
    changed = batch.run()
    single = nr.solve(g0, truncation_eps=1e-9, queue_capacity=1 << 22)
    assert changed["nodes"][0] == single["nodes"] and changed["paths"][0] == single["paths"]
    assert close(changed["R"][0], single["R"])
    assert (changed["nodes"][1:] == first["nodes"][1:]).all()
    other = nr.solve_batch(graphs[5:25], truncation_eps=1e-9, queue_capacity=1 << 22)
    assert [r["nodes"] for r in other] == [int(x) for x in first["nodes"][5:25]]


def test_er4096_full_batch():
    """The batched workload at full size (4096 seeded ER instances, delta
    1e-12, the benchmark's capacity 2^27): every instance closes R + D + ESP
    = 1 with nothing overflowed or re-run, and every 16th instance matches
    the oracle exactly; the batch's bracket for those meets the oracle's from
    a tighter truncation (1e-14)."""
    from concurrent.futures import ThreadPoolExecutor

    print("case ER batch 4096 eps=1e-12", flush=True)
    insts = [wl.er_instance(i) for i in range(4096)]
    graphs = []
    for n, edges, probs, s, t in insts:
        rp, ci, ei = wl.to_csr(n, edges)
        graphs.append(nr.Graph.from_csr(rp, ci, ei, probs, s, t))
    res = nr.solve_batch(graphs, truncation_eps=1e-12, queue_capacity=1 << 27)
    assert len(res) == 4096
    for r in res:
        assert r["overflowed"] == 0 and r["rerun"] == 0 and r["deadline_hit"] == 0
        assert abs(r["R"] + r["ESP"] + r["D"] - 1.0) <= 1e-12
    assert sum(r["nodes"] for r in res) > 1e9

    def oracle(i):
        n, edges, probs, s, t = insts[i]
        rp, ci, ei = wl.to_csr(n, edges)
        og = OracleGraph.from_csr(rp, ci, ei, probs)
        return (og.solve(s, t, order=ORDER_DFS, trunc_eps=1e-12),
                og.solve(s, t, order=ORDER_DFS, trunc_eps=1e-14))

    sample = list(range(0, 4096, 16))
    with ThreadPoolExecutor(16) as ex:
        outs = list(ex.map(oracle, sample))
    for i, (o, tight) in zip(sample, outs):
        check_inst(res[i], o, levels=False)
        lo = max(res[i]["R"], tight["R"])
        hi = min(res[i]["R"] + res[i]["ESP"], tight["R"] + tight["ESP"])
        assert lo <= hi * (1 + 1e-12), i


def test_graph_replays_equal_eager_launches():
    """The chunks of a batch are HIP graphs replayed between host checks;
    with per-launch timing the same launches run eagerly.  Four batches in a
    row (graph, eager, graph after an unrelated eager solve, graph) must give
    identical per-instance results."""
    print("case batch graph replays vs eager launches", flush=True)
    graphs = [er_pair(i)[0] for i in range(96)]
    k8 = nr.Graph.load(graph_path("COMP8-09.txt"), 0, 7)
    kw = dict(truncation_eps=1e-9, queue_capacity=1 << 20)
    runs = [nr.solve_batch(graphs, **kw), nr.solve_batch(graphs, time_launches=True, **kw)]
    nr.solve(k8, truncation_eps=0.0)
    runs.append(nr.solve_batch(graphs, **kw))
    runs.append(nr.solve_batch(graphs, **kw))
    keys = ("paths", "nodes", "dead", "truncated", "iterations", "R", "ESP", "D")
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            assert tuple(a[k] for k in keys) == tuple(b[k] for k in keys)
    assert all(r["kernel_launches"] > 0 for r in runs[0])


def test_tune_batch_picks_a_queue_chunk():
    print("case nr_tune_batch", flush=True)
    graphs = [er_pair(i)[0] for i in range(48)]
    cfg, rate = nr.tune_batch(graphs, truncation_eps=1e-8, queue_capacity=1 << 20)
    assert rate > 0
    assert cfg.block_size in (64, 128, 256)
    assert cfg.chunk_items in (0, (1 << 20) // 8, (1 << 20) // 2)
    res = nr.solve_batch(graphs, cfg)
    ref = nr.solve_batch(graphs, truncation_eps=1e-8, queue_capacity=1 << 20)
    for a, b in zip(res, ref):
        assert (a["paths"], a["nodes"], a["R"]) == (b["paths"], b["nodes"], b["R"])
