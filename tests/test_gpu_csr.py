"""GPU parity of the CSR expander (graphs of 65..256 nodes, directed
graphs) against the CPU oracle.

Same bar as the row form: bit-exact paths / nodes / dead / truncated,
levels and widest level in level order, R and ESP within 1e-12 relative,
single expansions bit-identical.  The K10 case embeds the reference's
COMP10-05 graph in a 70-node graph (60 isolated nodes), which routes it to
the CSR form, and must reproduce the logged K10 tree exactly.
"""
import numpy as np
import pytest

from conftest import closed_form, graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from relaibilitygpu_amd import workloads as wl  # noqa: E402
from oracle_ffi import ORDER_LEVEL, OracleGraph  # noqa: E402

REL = 1e-12


def close(a, b, rel=REL):
    return abs(a - b) <= rel * max(abs(a), abs(b)) + 1e-300


def check_same(r, o):
    assert (r["paths"], r["nodes"], r["dead"], r["truncated"]) == (
        o["paths"], o["nodes"], o["dead"], o["truncated"])
    assert r["iterations"] == o["iterations"]
    assert r["max_queue_items"] == o["max_width"]
    assert close(r["R"], o["R"]), (r["R"], o["R"])
    assert close(r["ESP"], o["ESP"]), (r["ESP"], o["ESP"])


def pair_csr(rp, ci, ei, probs, s, t):
    return nr.Graph.from_csr(rp, ci, ei, probs, s, t), OracleGraph.from_csr(rp, ci, ei, probs)


def grid10(seed=0x6121D):
    rng = wl.Xoshiro256ss(seed)
    n, edges, _ = wl.grid_graph(10, 10, [0.0] * 1000)
    probs = [0.85 + 0.10 * rng.uniform() for _ in edges]
    return n, edges, probs


def test_k10_embedded_in_70_nodes(expected):
    print("case K10 in 70 nodes (CSR form)", flush=True)
    og = OracleGraph.load(graph_path("COMP10-05.txt"))
    rp, ci, ei, p = og.arrays()
    rp = list(rp) + [rp[-1]] * 60
    g = nr.Graph.from_csr(rp, ci, ei, p, 0, 9)
    assert g.info().n == 70
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22)
    assert (r["paths"], r["nodes"], r["iterations"], r["max_queue_items"]) == (
        7433032, 14866064, 38, 1596000)
    assert close(r["R"], closed_form(expected, 10, "0.5"))


@pytest.mark.parametrize("eps", [1e-4, 1e-6])
def test_grid_10x10(eps):
    print(f"case grid 10x10 eps={eps}", flush=True)
    n, edges, probs = grid10()
    rp, ci, ei = wl.to_csr(n, edges)
    g, og = pair_csr(rp, ci, ei, probs, 0, 99)
    r = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 22)
    check_same(r, og.solve(0, 99, order=ORDER_LEVEL, trunc_eps=eps))


@pytest.mark.parametrize("eps", [1e-9, 1e-12])
def test_barabasi_albert_150(eps):
    print(f"case BA 150 eps={eps}", flush=True)
    n, edges, probs = wl.barabasi_albert(150, 2, 0.95, 0xBA150)
    rp, ci, ei = wl.to_csr(n, edges)
    g, og = pair_csr(rp, ci, ei, probs, 0, 149)
    assert g.info().n == 150
    r = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 22)
    check_same(r, og.solve(0, 149, order=ORDER_LEVEL, trunc_eps=eps))


@pytest.mark.parametrize("n,mean_out,eps,seed", [(12, 3.0, 0.0, 1), (24, 3.0, 1e-9, 2),
                                                 (60, 2.5, 1e-7, 3), (200, 2.0, 1e-6, 4)])
def test_directed_graphs(n, mean_out, eps, seed):
    print(f"case directed n={n} eps={eps}", flush=True)
    n, arcs, probs = wl.random_digraph(n, mean_out, 0.6, 0.95, 0xD1000 + seed)
    rp, ci, ei = wl.to_csr_directed(n, arcs)
    g, og = pair_csr(rp, ci, ei, probs, 0, n - 1)
    assert g.info().symmetric == 0
    r = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 22)
    o = og.solve(0, n - 1, order=ORDER_LEVEL, trunc_eps=eps)
    check_same(r, o)
    if eps == 0.0:
        assert abs(r["R"] + r["D"] - 1.0) <= 1e-13


def test_directed_one_way_pair_differs():
    """One-way arcs: R(s->t) and R(t->s) differ on a directed graph, and
    each matches the oracle."""
    print("case directed reverse", flush=True)
    n, arcs, probs = wl.random_digraph(16, 3.0, 0.6, 0.95, 0xD2005)
    rp, ci, ei = wl.to_csr_directed(n, arcs)
    og = OracleGraph.from_csr(rp, ci, ei, probs)
    rs = []
    for s, t in [(0, 15), (15, 0)]:
        g = nr.Graph.from_csr(rp, ci, ei, probs, s, t)
        r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22)
        check_same(r, og.solve(s, t, order=ORDER_LEVEL))
        rs.append(r["R"])
    assert rs[0] != rs[1]


@pytest.mark.parametrize("kind", ["grid", "ba", "directed"])
def test_single_expansions_bit_exact(kind):
    print(f"case expansions {kind}", flush=True)
    if kind == "grid":
        n, edges, probs = grid10()
        rp, ci, ei = wl.to_csr(n, edges)
    elif kind == "ba":
        n, edges, probs = wl.barabasi_albert(150, 2, 0.95, 0xBA150)
        rp, ci, ei = wl.to_csr(n, edges)
    else:
        n, edges, probs = wl.random_digraph(90, 3.0, 0.6, 0.95, 0xD3000)
        rp, ci, ei = wl.to_csr_directed(n, edges)
    s, t = 0, n - 1
    g, og = pair_csr(rp, ci, ei, probs, s, t)
    W = g.info().mask_words
    E = len(edges)
    rng = np.random.default_rng(7)
    count = 300
    known = np.zeros((count, W), np.uint64)
    up = np.zeros((count, W), np.uint64)
    for i in range(count):
        st = rng.choice(3, size=E, p=[0.6, 0.25, 0.15])  # unknown, working, failed
        for e in range(E):
            if st[e]:
                known[i, e >> 6] |= np.uint64(1) << np.uint64(e & 63)
            if st[e] == 1:
                up[i, e >> 6] |= np.uint64(1) << np.uint64(e & 63)
    mass = rng.uniform(0.1, 1.0, count)
    ex = nr.expand_items(g, known, up, mass, max_children=n)
    alive = 0
    for i in range(count):
        k, succ, ck, cu, cm, _ = og.expand(s, t, known[i], up[i], float(mass[i]))
        assert ex.nchild[i] == k
        if k < 0:
            continue
        alive += 1
        assert ex.succ[i] == succ
        assert np.array_equal(ex.child_known[i, :k], ck)
        assert np.array_equal(ex.child_up[i, :k], cu)
        assert np.array_equal(ex.child_mass[i, :k], cm)
    assert alive > count // 10


def test_grid_mass_order_and_deadline():
    """Max-mass order on the CSR form: a full run (no trims) visits the level
    order's tree; with a 20 ms budget at eps = 0 the bracket holds against
    the oracle's own bracket at eps = 1e-7."""
    print("case grid mass order", flush=True)
    n, edges, probs = grid10()
    rp, ci, ei = wl.to_csr(n, edges)
    g, og = pair_csr(rp, ci, ei, probs, 0, 99)
    r = nr.solve(g, truncation_eps=1e-5, queue_capacity=1 << 22, order="mass")
    lv = nr.solve(g, truncation_eps=1e-5, queue_capacity=1 << 22, order="level")
    assert (r["paths"], r["nodes"], r["dead"], r["truncated"]) == (
        lv["paths"], lv["nodes"], lv["dead"], lv["truncated"])
    assert close(r["R"], lv["R"])
    o = og.solve(0, 99, order=ORDER_LEVEL, trunc_eps=1e-7)
    for order in ("level", "mass"):
        b = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 24, time_budget_s=0.02,
                     order=order)
        assert b["deadline_hit"] == 1
        assert b["R"] <= (o["R"] + o["ESP"]) * (1 + 1e-12)
        assert b["R"] + b["ESP"] >= o["R"] * (1 - 1e-12)
        assert abs(b["R"] + b["ESP"] + b["D"] - 1.0) <= 1e-12
