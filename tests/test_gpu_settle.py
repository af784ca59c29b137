"""GPU: children counted at creation (settled) against children queued.

In level order the engine counts a child below the truncation, and in the
row form a dead child (its FAILED hop is a src-dst bridge of the parent's
allowed graph), when the parent creates it, instead of queueing it and
counting it when it is dequeued.  NR_FLAG_QUEUE_ALL queues every child, as
the reference does (workingPad/main.cu:245-252, 537-543, 579-582).  Bar:
identical paths, nodes, dead, truncated, levels and widest level; R, ESP
and D within 1e-12 relative (the same terms summed in another order); and
the same counts as the oracle, which queues every child.
"""
import pytest

from conftest import closed_form, graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from oracle_ffi import OracleGraph  # noqa: E402
from relaibilitygpu_amd.workloads import (er_instance, grid_graph, random_digraph,  # noqa: E402
                                          to_csr, to_csr_directed)

REL = 1e-12
COUNTS = ("paths", "nodes", "dead", "truncated", "iterations", "max_queue_items", "overflowed")


def close(a, b):
    return abs(a - b) <= REL * max(abs(a), abs(b)) + 1e-300


def both(g, **kw):
    settled = nr.solve(g, **kw)
    queued = nr.solve(g, queue_all=True, **kw)
    for key in COUNTS:
        assert settled[key] == queued[key], (key, settled[key], queued[key])
    for key in ("R", "ESP", "D"):
        assert close(settled[key], queued[key]), (key, settled[key], queued[key])
    return settled


def against_oracle(r, og, s, t, eps):
    """The path rule's counts in level order, as the oracle queues them."""
    o = og.solve(s, t, trunc_eps=eps, threads=8)
    assert (r["paths"], r["nodes"], r["dead"], r["truncated"]) == (
        o["paths"], o["nodes"], o["dead"], o["truncated"])
    assert (r["iterations"], r["max_queue_items"]) == (o["iterations"], o["max_width"])
    assert close(r["R"], o["R"]) and close(r["ESP"], o["ESP"])


@pytest.mark.parametrize("name,s,t,eps", [
    ("COMP8-09.txt", 0, 7, 0.0), ("COMP9.txt", 0, 8, 0.0), ("COMP10-05.txt", 0, 9, 0.0),
    ("COMP10-05.txt", 0, 9, 1e-6), ("COMP15-05.txt", 0, 14, 1e-9), ("test_kite.txt", 0, 4, 0.0),
    ("cycle4.txt", 0, 2, 0.0), ("test_bridge.txt", 0, 3, 0.0)])
def test_row_form(name, s, t, eps):
    """Row form: dead and truncated children settled."""
    print(f"case row form {name} {s}->{t} eps={eps}", flush=True)
    g = nr.Graph.load(graph_path(name), s, t)
    r = both(g, truncation_eps=eps, queue_capacity=1 << 24)
    if name != "COMP10-05.txt" or eps > 0:
        against_oracle(r, OracleGraph.load(graph_path(name)), s, t, eps)


def test_k10_closed_form(expected):
    print("case K10 full, settled vs queued, closed form", flush=True)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    r = both(g, truncation_eps=0.0, queue_capacity=4194304)
    assert (r["paths"], r["nodes"], r["iterations"], r["max_queue_items"]) == (
        7433032, 14866064, 38, 1596000)
    assert close(r["R"], closed_form(expected, 10, "0.5"))
    assert abs(r["R"] + r["D"] - 1.0) < 1e-13


@pytest.mark.parametrize("i", [0, 3, 7, 11])
def test_er_instances(i):
    """Erdos-Renyi instances (12-40 nodes: 32- and 64-bit node sets)."""
    n, edges, probs, s, t = er_instance(i)
    print(f"case ER instance {i} n={n} e={len(edges)}", flush=True)
    rp, ci, ei = to_csr(n, edges)
    g = nr.Graph.from_csr(rp, ci, ei, probs, s, t)
    r = both(g, truncation_eps=1e-12, queue_capacity=1 << 24)
    against_oracle(r, OracleGraph.from_csr(rp, ci, ei, probs), s, t, 1e-12)


def test_lean_form_k25():
    """K25: items without rows (lean form), truncated children settled."""
    print("case K25 lean 1e-9", flush=True)
    g = nr.Graph.load(graph_path("COMP25-05.txt"), 0, 24)
    r = both(g, truncation_eps=1e-9, queue_capacity=1 << 24)
    against_oracle(r, OracleGraph.load(graph_path("COMP25-05.txt")), 0, 24, 1e-9)


@pytest.mark.parametrize("name,t", [("COMP15-05.txt", 14), ("COMP25-05.txt", 24)])
def test_alternating_rule(name, t):
    print(f"case alternating {name} 1e-10", flush=True)
    g = nr.Graph.load(graph_path(name), 0, t)
    both(g, truncation_eps=1e-10, queue_capacity=1 << 24, rule="alternate", sweep_period=6)


def test_csr_forms():
    """Graphs past the row form: a 10x10 grid and a directed graph."""
    print("case grid 10x10 and digraph (CSR form)", flush=True)
    n, edges, probs = grid_graph(10, 10, [0.9] * 180)
    rp, ci, ei = to_csr(n, edges)
    g = nr.Graph.from_csr(rp, ci, ei, probs, 0, 99)
    r = both(g, truncation_eps=1e-7, queue_capacity=1 << 24)
    against_oracle(r, OracleGraph.from_csr(rp, ci, ei, probs), 0, 99, 1e-7)
    n, arcs, probs = random_digraph(40, 4.0, 0.7, 0.95, 5)
    rp, ci, ei = to_csr_directed(n, arcs)
    g = nr.Graph.from_csr(rp, ci, ei, probs, 0, 39)
    r = both(g, truncation_eps=1e-9, queue_capacity=1 << 24)
    against_oracle(r, OracleGraph.from_csr(rp, ci, ei, probs), 0, 39, 1e-9)


def test_k_terminal_truncated_only():
    """Terminal sets: truncated children settled, dead ones queued."""
    print("case K9 terminals {0, 3, 8}", flush=True)
    g = nr.Graph.load(graph_path("COMP9.txt"), 0, 8)
    g.set_terminal_set([0, 3, 8])
    both(g, truncation_eps=1e-5, queue_capacity=1 << 22)


def test_overflow_brackets(expected):
    """A capacity that overflows: settled children take no slot, so fewer
    children overflow; both brackets hold around the closed form."""
    print("case K10 capacity 65536", flush=True)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    R = closed_form(expected, 10, "0.5")
    for qa in (False, True):
        r = nr.solve(g, truncation_eps=0.0, queue_capacity=65536, queue_all=qa)
        assert r["overflowed"] > 0
        assert r["R"] <= R + 1e-12 and R <= r["R"] + r["ESP"] + 1e-12
