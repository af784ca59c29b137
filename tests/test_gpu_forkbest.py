"""GPU parity of the forkBest rule (NR_RULE_FORKBEST) against the oracle's
forkbest mode and the reference's own logged runs: a confirmed WORKING path
is a success, otherwise a two-way branch on the path's UNKNOWN edge of
highest FP32 log p, ties settled as the reference's warp argmax settles them
(lane r mod 32 for hop r from the dst end, then its shuffle tree).

Counts bit-exact, levels and widest level in level order, R to 1e-12; in
full enumeration R also equals the closed form.  K8, K8 at p = 0.3, K9 and
K10 reproduce workingPad/output_maxMem_forBest.txt exactly (paths, nodes,
kernel iterations, max queue items).
"""
import numpy as np
import pytest

from conftest import closed_form, graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from relaibilitygpu_amd import workloads as wl  # noqa: E402
from oracle_ffi import MODE_FORKBEST, ORDER_LEVEL, OracleGraph  # noqa: E402

REL = 1e-12


def close(a, b):
    return abs(a - b) <= REL * max(abs(a), abs(b)) + 1e-300


def check(r, o):
    assert (r["paths"], r["nodes"], r["dead"], r["truncated"]) == (
        o["paths"], o["nodes"], o["dead"], o["truncated"])
    assert r["iterations"] == o["iterations"]
    assert r["max_queue_items"] == o["max_width"]
    assert close(r["R"], o["R"]) and close(r["ESP"], o["ESP"])


@pytest.mark.parametrize("name,s,t", [("test_simple.txt", 0, 3), ("test_kite.txt", 0, 4),
                                      ("temp4x4.txt", 0, 3), ("cycle4.txt", 0, 2),
                                      ("COMP8-09.txt", 0, 7), ("COMP9.txt", 0, 8)])
def test_files_vs_oracle(name, s, t):
    print(f"case forkbest {name}", flush=True)
    g = nr.Graph.load(graph_path(name), s, t)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, rule="forkbest")
    o = OracleGraph.load(graph_path(name)).solve(s, t, mode=MODE_FORKBEST, order=ORDER_LEVEL)
    check(r, o)


@pytest.mark.parametrize("idx", [0, 1, 2, 3])
def test_reference_forkbest_logs(expected, idx):
    """The reference's logged forkBest runs (K8, K8 p = 0.3, K9, K10)."""
    lg = expected["logged_forkbest_counts"][idx]
    print(f"case forkbest log {lg['graph']}", flush=True)
    g = nr.Graph.load(graph_path(lg["graph"]), lg["src"], lg["dst"])
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 23, rule="forkbest")
    assert (r["paths"], r["nodes"], r["iterations"], r["max_queue_items"]) == (
        lg["paths"], lg["nodes"], lg["iterations"], lg["max_width"]), lg
    assert abs(r["R"] + r["D"] - 1.0) <= 1e-13
    if lg["graph"] == "COMP10-05.txt":
        assert close(r["R"], closed_form(expected, 10, "0.5"))


@pytest.mark.parametrize("chords", [0, 3])
def test_forkbest_long_paths(chords):
    """Rings of 80 nodes (CSR form), uniform p, src and dst 40 hops apart:
    paths of more than 32 hops put several hops on one of the reference's
    lanes, which exercises both stages of its tie-break."""
    print(f"case forkbest ring80 chords={chords}", flush=True)
    n = 80
    edges = [(i, (i + 1) % n) for i in range(n)]
    edges = sorted((min(a, b), max(a, b)) for a, b in edges)
    for a, b in [(5, 70), (12, 33), (50, 61)][:chords]:
        edges.append((a, b))
    probs = [0.9] * len(edges)
    rp, ci, ei = wl.to_csr(n, edges)
    g = nr.Graph.from_csr(rp, ci, ei, probs, 0, 40)
    og = OracleGraph.from_csr(rp, ci, ei, probs)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, rule="forkbest")
    check(r, og.solve(0, 40, mode=MODE_FORKBEST, order=ORDER_LEVEL))


def test_forkbest_keys_are_fp32_logs(expected):
    """p = 0.5 plus perturbations below FP32 resolution: the keys tie as in
    the reference (FP32 logf), so the tree is the logged uniform-p K8 one."""
    print("case forkbest fp32 keys", flush=True)
    n, edges = 8, [(a, b) for a in range(8) for b in range(a + 1, 8)]
    probs = [0.5 + 1e-12 * (i % 7) for i in range(len(edges))]
    rp, ci, ei = wl.to_csr(n, edges)
    g = nr.Graph.from_csr(rp, ci, ei, probs, 0, 7)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, rule="forkbest")
    lg = expected["logged_forkbest_counts"][0]
    assert (r["paths"], r["nodes"], r["iterations"]) == (lg["paths"], lg["nodes"], lg["iterations"])


def test_forkbest_mass_order_same_tree(expected):
    print("case forkbest mass order K9", flush=True)
    g = nr.Graph.load(graph_path("COMP9.txt"), 0, 8)
    lv = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, rule="forkbest")
    m = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, rule="forkbest", order="mass")
    assert (m["paths"], m["nodes"], m["dead"]) == (lv["paths"], lv["nodes"], lv["dead"])
    assert close(m["R"], lv["R"])


@pytest.mark.parametrize("kind,eps", [("grid", 1e-5), ("directed", 1e-7), ("er", 1e-9)])
def test_forkbest_csr_and_sparse(kind, eps):
    print(f"case forkbest {kind} eps={eps}", flush=True)
    if kind == "grid":
        rng = wl.Xoshiro256ss(0x6121D)
        n, edges, _ = wl.grid_graph(10, 10, [0.0] * 1000)
        probs = [0.85 + 0.10 * rng.uniform() for _ in edges]
        rp, ci, ei = wl.to_csr(n, edges)
    elif kind == "directed":
        n, edges, probs = wl.random_digraph(60, 2.5, 0.6, 0.95, 0xD1003)
        rp, ci, ei = wl.to_csr_directed(n, edges)
    else:
        n, edges, probs, s0, t0 = wl.er_instance(3)
        rp, ci, ei = wl.to_csr(n, edges)
    s, t = (s0, t0) if kind == "er" else (0, n - 1)
    g = nr.Graph.from_csr(rp, ci, ei, probs, s, t)
    og = OracleGraph.from_csr(rp, ci, ei, probs)
    r = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 22, rule="forkbest")
    check(r, og.solve(s, t, mode=MODE_FORKBEST, order=ORDER_LEVEL, trunc_eps=eps))


def test_forkbest_deadline_bracket(expected):
    print("case forkbest K15 deadline", flush=True)
    g = nr.Graph.load(graph_path("COMP15-05.txt"), 0, 14)
    R = closed_form(expected, 15, "0.5")
    for order in ("level", "mass"):
        r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 24, time_budget_s=0.01,
                     rule="forkbest", order=order)
        assert r["deadline_hit"] == 1
        assert r["R"] <= R * (1 + 1e-12) and r["R"] + r["ESP"] >= R * (1 - 1e-12)
        assert abs(r["R"] + r["ESP"] + r["D"] - 1.0) <= 1e-12


def test_rule_rejects():
    g = nr.Graph.load(graph_path("COMP8-09.txt"), 0, 7)
    with pytest.raises(nr.NRError):
        nr.solve(g, rule=7)
    with pytest.raises(nr.NRError):
        nr.solve_batch([g], rule="forkbest")
