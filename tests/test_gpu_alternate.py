"""GPU parity of the alternating-sweep rule (NR_RULE_ALTERNATE) against the
oracle's alternate mode: a path with the fewest UNKNOWN edges (WORKING
edges contracted, lowest-index walk back), children swept from the src end
when (FAILED edges / sweep_period) is even and from the dst end otherwise.

Counts bit-exact (paths, nodes, dead, truncated; levels and widest level in
level order), R and ESP to 1e-12; single expansions bit-identical (planes
and FP64 masses).  In full enumeration R also equals the closed form and
brute force.  The rule has no counterpart in the reference: its counts are
pinned to the oracle, whose path choice and sweep are checked against a 0-1
BFS distance and a component walk in tests/test_oracle.py.
"""
import numpy as np
import pytest

from conftest import closed_form, graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from relaibilitygpu_amd import workloads as wl  # noqa: E402
from oracle_ffi import (MODE_ALTERNATE, ORDER_DFS, ORDER_LEVEL,  # noqa: E402
                        OracleGraph)

REL = 1e-12


def close(a, b, rel=REL):
    return abs(a - b) <= rel * max(abs(a), abs(b)) + 1e-300


def check(r, o, levels=True):
    assert (r["paths"], r["nodes"], r["dead"], r["truncated"]) == (
        o["paths"], o["nodes"], o["dead"], o["truncated"])
    if levels:
        assert r["iterations"] == o["iterations"]
        assert r["max_queue_items"] == o["max_width"]
    assert close(r["R"], o["R"]) and close(r["ESP"], o["ESP"])


@pytest.mark.parametrize("name,s,t", [("test_simple.txt", 0, 3), ("test_kite.txt", 0, 4),
                                      ("test_bridge.txt", 0, 3), ("temp4x4.txt", 0, 3),
                                      ("cycle4.txt", 0, 2), ("COMP8-09.txt", 0, 7),
                                      ("COMP9.txt", 0, 8)])
def test_files_vs_oracle(name, s, t):
    print(f"case alternate {name}", flush=True)
    g = nr.Graph.load(graph_path(name), s, t)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, rule="alternate")
    o = OracleGraph.load(graph_path(name)).solve(s, t, mode=MODE_ALTERNATE, order=ORDER_LEVEL)
    check(r, o)
    assert abs(r["R"] + r["D"] - 1.0) <= 1e-13


@pytest.mark.parametrize("period", [1, 2, 3, 5])
def test_k10_periods(expected, period):
    print(f"case alternate K10 period {period}", flush=True)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, rule="alternate",
                 sweep_period=period)
    o = OracleGraph.load(graph_path("COMP10-05.txt")).solve(
        0, 9, mode=MODE_ALTERNATE, order=ORDER_LEVEL, sweep_period=period)
    check(r, o)
    assert close(r["R"], closed_form(expected, 10, "0.5"))
    assert r["nodes"] < 14866064 // 2  # the plain rule's K10 tree is twice as large


def test_mass_order_same_tree():
    print("case alternate K9 mass order", flush=True)
    g = nr.Graph.load(graph_path("COMP9.txt"), 0, 8)
    lv = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, rule="alternate")
    m = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, rule="alternate", order="mass")
    assert (m["paths"], m["nodes"], m["dead"]) == (lv["paths"], lv["nodes"], lv["dead"])
    assert close(m["R"], lv["R"])


@pytest.mark.parametrize("name,n", [("COMP15-05.txt", 15), ("COMP25-05.txt", 25)])
def test_truncated_vs_oracle_and_tenfold_gap(expected, name, n):
    """delta = 1e-12: counts equal the oracle's, the bracket holds, and the
    certified gap is at least 10x below the plain path rule's."""
    print(f"case alternate {name} eps=1e-12", flush=True)
    g = nr.Graph.load(graph_path(name), 0, n - 1)
    kw = dict(truncation_eps=1e-12, queue_capacity=1 << 26)
    r = nr.solve(g, rule="alternate", **kw)
    o = OracleGraph.load(graph_path(name)).solve(0, n - 1, mode=MODE_ALTERNATE, order=ORDER_DFS,
                                                 threads=16, trunc_eps=1e-12)
    check(r, o, levels=False)
    R = closed_form(expected, n, "0.5")
    assert r["R"] <= R * (1 + 1e-12) <= (r["R"] + r["ESP"]) * (1 + 2e-12)
    plain = nr.solve(g, **kw)
    assert plain["ESP"] >= 10.0 * r["ESP"], (plain["ESP"], r["ESP"])


@pytest.mark.parametrize("i", [3, 17, 40, 77])
def test_er_instances_vs_oracle(i):
    n, edges, probs, s, t = wl.er_instance(i)
    print(f"case alternate ER instance {i} n={n} E={len(edges)}", flush=True)
    rp, ci, ei = wl.to_csr(n, edges)
    g = nr.Graph.from_csr(rp, ci, ei, probs, s, t)
    og = OracleGraph.from_csr(rp, ci, ei, probs)
    for eps in (1e-9, 1e-12):
        r = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 24, rule="alternate")
        o = og.solve(s, t, mode=MODE_ALTERNATE, order=ORDER_LEVEL, trunc_eps=eps)
        check(r, o)


@pytest.mark.parametrize("name,s,t,period", [("COMP10-05.txt", 0, 9, 1),
                                             ("COMP10-05.txt", 0, 9, 2),
                                             ("COMP15-05.txt", 3, 11, 3)])
def test_single_expansions_bit_identical(name, s, t, period):
    print(f"case alternate expansions {name} period {period}", flush=True)
    g = nr.Graph.load(graph_path(name), s, t)
    og = OracleGraph.load(graph_path(name))
    W = og.mask_words
    E = og.dims()["e"]
    rng = np.random.default_rng(7 + period)
    count = 500
    known = np.zeros((count, W), np.uint64)
    up = np.zeros((count, W), np.uint64)
    for i in range(count):
        state = rng.choice(3, size=E, p=[0.55, 0.25, 0.2])  # unknown, failed, working
        for e in range(E):
            if state[e] == 1:
                known[i, e >> 6] |= np.uint64(1 << (e & 63))
            elif state[e] == 2:
                known[i, e >> 6] |= np.uint64(1 << (e & 63))
                up[i, e >> 6] |= np.uint64(1 << (e & 63))
    mass = rng.uniform(0.1, 1.0, count)
    x = nr.expand_items(g, known, up, mass, max_children=16, rule="alternate",
                        sweep_period=period)
    live = 0
    for i in range(count):
        k, succ, ck, cu, cm, _ = og.expand_rule(s, t, known[i], up[i], mass[i], MODE_ALTERNATE,
                                                period)
        assert x.nchild[i] == k
        if k < 0:
            continue
        live += 1
        assert x.succ[i] == succ
        m = min(k, 16)
        assert np.array_equal(x.child_known[i, :m], ck[:m])
        assert np.array_equal(x.child_up[i, :m], cu[:m])
        assert np.array_equal(x.child_mass[i, :m], cm[:m])
    assert live > count // 4


def test_brackets_under_capacity_and_deadline(expected):
    print("case alternate K15 capacity and deadline", flush=True)
    g = nr.Graph.load(graph_path("COMP15-05.txt"), 0, 14)
    R = closed_form(expected, 15, "0.5")
    for kw in (dict(queue_capacity=1 << 14, truncation_eps=1e-15),
               dict(queue_capacity=1 << 14, truncation_eps=1e-15, order="mass"),
               dict(queue_capacity=1 << 24, truncation_eps=0.0, time_budget_s=0.005)):
        r = nr.solve(g, rule="alternate", **kw)
        assert r["R"] <= R * (1 + 1e-12) <= (r["R"] + r["ESP"]) * (1 + 2e-12), kw
        assert abs(r["R"] + r["ESP"] + r["D"] - 1.0) <= 1e-12


def test_checkpoint_chain_ends_like_one_solve(tmp_path):
    print("case alternate K15 mass checkpoint chain", flush=True)
    g = nr.Graph.load(graph_path("COMP15-05.txt"), 0, 14)
    kw = dict(truncation_eps=1e-13, queue_capacity=1 << 24, order="mass", rule="alternate",
              sweep_period=3)
    full = nr.solve(g, **kw)
    ck, links, r = None, 0, None
    while True:
        r, ck = nr.solve_checkpointed(g, resume=ck, time_budget_s=2e-4, **kw)
        links += 1
        if ck is None:
            break
        if links % 3 == 1:
            path = str(tmp_path / "alt.ck")
            ck.save(path)
            ck = nr.Checkpoint.load(path)
        assert links < 4000
    assert links >= 2
    assert (r["paths"], r["nodes"], r["dead"], r["truncated"]) == (
        full["paths"], full["nodes"], full["dead"], full["truncated"])
    assert close(r["R"], full["R"]) and close(r["ESP"], full["ESP"])
    # a checkpoint resumes only under the same sweep period
    r1, ck = nr.solve_checkpointed(g, time_budget_s=1e-4, **kw)
    if ck is not None:
        bad = dict(kw, sweep_period=2)
        with pytest.raises(nr.NRError):
            nr.solve_checkpointed(g, resume=ck, **bad)


def test_learned_order_on_alternate_rule(expected):
    print("case alternate learned order K9 / K15", flush=True)
    g = nr.Graph.load(graph_path("COMP9.txt"), 0, 8)
    X, Y, res = nr.log_subproblems(g, 1 << 22, seed=3, truncation_eps=0.0,
                                   queue_capacity=1 << 22, rule="alternate")
    lv = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, rule="alternate")
    assert res["nodes"] == lv["nodes"] and len(Y) == lv["nodes"] - 1
    assert np.all(np.isfinite(X)) and np.all(Y >= 0)
    rng = np.random.default_rng(5)
    m = nr.Model(rng.normal(0, 0.3, nr.MLP_PARAMS).astype(np.float32), alpha=0.5)
    r = nr.solve_learned(g, m, truncation_eps=0.0, queue_capacity=1 << 22, rule="alternate")
    assert (r["paths"], r["nodes"], r["dead"]) == (lv["paths"], lv["nodes"], lv["dead"])
    assert close(r["R"], closed_form(expected, 9, "0.5"))
    g15 = nr.Graph.load(graph_path("COMP15-05.txt"), 0, 14)
    r = nr.solve_learned(g15, m, truncation_eps=1e-12, queue_capacity=1 << 24, rule="alternate")
    o = nr.solve(g15, truncation_eps=1e-12, queue_capacity=1 << 24, rule="alternate")
    assert (r["paths"], r["nodes"], r["truncated"]) == (o["paths"], o["nodes"], o["truncated"])


def _grid10():
    rng = wl.Xoshiro256ss(0x6121D)
    n, edges, _ = wl.grid_graph(10, 10, [0.0] * 1000)
    probs = [0.85 + 0.10 * rng.uniform() for _ in edges]
    return n, edges, probs


@pytest.mark.parametrize("kind,eps", [("grid", 1e-6), ("grid", 1e-8), ("ba", 1e-12),
                                      ("ba", 1e-15)])
def test_csr_form_vs_oracle(kind, eps):
    """Undirected graphs of more than 64 nodes run the rule on the CSR arcs
    (CsrAltExpander): same counts as the oracle, and a smaller certified gap
    than the plain rule at the same delta."""
    print(f"case alternate CSR {kind} eps={eps}", flush=True)
    if kind == "grid":
        n, edges, probs = _grid10()
        s, t = 0, 99
    else:
        n, edges, probs = wl.barabasi_albert(150, 2, 0.95, 0xBA150)
        s, t = 0, 149
    rp, ci, ei = wl.to_csr(n, edges)
    g = nr.Graph.from_csr(rp, ci, ei, probs, s, t)
    og = OracleGraph.from_csr(rp, ci, ei, probs)
    for period in (1, 2):
        r = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 24, rule="alternate",
                     sweep_period=period)
        o = og.solve(s, t, mode=MODE_ALTERNATE, order=ORDER_LEVEL, trunc_eps=eps,
                     sweep_period=period)
        check(r, o)
    plain = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 24)
    assert plain["ESP"] > r["ESP"]


@pytest.mark.parametrize("kind", ["grid", "ba"])
def test_csr_single_expansions(kind):
    print(f"case alternate CSR expansions {kind}", flush=True)
    if kind == "grid":
        n, edges, probs = _grid10()
        s, t = 0, 99
    else:
        n, edges, probs = wl.barabasi_albert(150, 2, 0.95, 0xBA150)
        s, t = 5, 140
    rp, ci, ei = wl.to_csr(n, edges)
    g = nr.Graph.from_csr(rp, ci, ei, probs, s, t)
    og = OracleGraph.from_csr(rp, ci, ei, probs)
    W = og.mask_words
    E = len(edges)
    rng = np.random.default_rng(31)
    count = 300
    known = np.zeros((count, W), np.uint64)
    up = np.zeros((count, W), np.uint64)
    for i in range(count):
        state = rng.choice(3, size=E, p=[0.6, 0.15, 0.25])
        for e in range(E):
            if state[e]:
                known[i, e >> 6] |= np.uint64(1 << (e & 63))
            if state[e] == 2:
                up[i, e >> 6] |= np.uint64(1 << (e & 63))
    mass = rng.uniform(0.1, 1.0, count)
    x = nr.expand_items(g, known, up, mass, max_children=n, rule="alternate", sweep_period=1)
    live = 0
    for i in range(count):
        k, succ, ck, cu, cm, _ = og.expand_rule(s, t, known[i], up[i], mass[i], MODE_ALTERNATE, 1)
        assert x.nchild[i] == k
        if k < 0:
            continue
        live += 1
        assert x.succ[i] == succ
        assert np.array_equal(x.child_known[i, :k], ck[:k])
        assert np.array_equal(x.child_up[i, :k], cu[:k])
        assert np.array_equal(x.child_mass[i, :k], cm[:k])
    assert live > 10


def test_unsupported_forms():
    print("case alternate unsupported forms", flush=True)
    n, edges, probs = wl.random_digraph(20, 2.5, 0.6, 0.95, 0xD1003)
    rp, ci, ei = wl.to_csr_directed(n, edges)
    d = nr.Graph.from_csr(rp, ci, ei, probs, 0, n - 1)
    with pytest.raises(nr.NRError) as e1:
        nr.solve(d, rule="alternate")
    assert e1.value.status == 5
    g = nr.Graph.load(graph_path("COMP8.txt"), 0, 7)
    g.set_terminal_set([0, 3, 7])
    with pytest.raises(nr.NRError) as e2:
        nr.solve(g, rule="alternate")
    assert e2.value.status == 5
    with pytest.raises(nr.NRError) as e3:
        nr.solve(nr.Graph.load(graph_path("COMP8.txt"), 0, 7), rule="alternate", sweep_period=-1)
    assert e3.value.status == 4
