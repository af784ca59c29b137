"""GPU parity of the edge-pivot rule (NR_RULE_PIVOT; workingPad/cpu_reference.cpp
:107-145): a confirmed WORKING path is a success, an unreachable dst a dead
end, otherwise a two-way branch on the UNKNOWN edge of highest p over the
whole graph (lowest id on ties).

The K8 counts are those of the reference's own program (tests/golden/
pivot_ref.json, written by tests/golden/make_pivot_vectors.py from the
reference's cpu_reference.cpp compiled into oracle/_ref); the other cases
compare with the oracle's pivot mode, which reproduces that program's counts
on every vector (tests/test_oracle.py).  Counts bit-exact, R to 1e-12.
"""
import json
import os

import pytest

from conftest import closed_form, graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from relaibilitygpu_amd import workloads as wl  # noqa: E402
from oracle_ffi import MODE_PIVOT, ORDER_LEVEL, OracleGraph  # noqa: E402

REL = 1e-12


def close(a, b):
    return abs(a - b) <= REL * max(abs(a), abs(b)) + 1e-300


def check(r, o, levels=True):
    assert (r["paths"], r["nodes"], r["dead"], r["truncated"]) == (
        o["paths"], o["nodes"], o["dead"], o["truncated"])
    if levels:
        assert (r["iterations"], r["max_queue_items"]) == (o["iterations"], o["max_width"])
    assert close(r["R"], o["R"]) and close(r["ESP"], o["ESP"])


def test_reference_program_counts(expected):
    """K8 (p = 0.5, 0.3, 0.9): the reference program's node and path counts."""
    vecs = json.load(open(os.path.join(os.path.dirname(graph_path("COMP8.txt")), "..",
                                       "pivot_ref.json")))
    for v in vecs:
        print(f"case pivot {v['graph']} {v['src']}->{v['dst']}", flush=True)
        g = nr.Graph.load(graph_path(v["graph"]), v["src"], v["dst"])
        r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 23, rule="pivot")
        assert (r["paths"], r["nodes"]) == (v["paths"], v["nodes"]), v
        assert abs(r["R"] + r["D"] - 1.0) <= 1e-13
    g = nr.Graph.load(graph_path("COMP8.txt"), 0, 7)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 23, rule="pivot")
    assert close(r["R"], closed_form(expected, 8, "0.5"))


@pytest.mark.parametrize("name,s,t", [("test_kite.txt", 0, 4), ("temp4x4.txt", 0, 3),
                                      ("COMP8-09.txt", 0, 7)])
def test_levels_vs_oracle(name, s, t):
    print(f"case pivot levels {name}", flush=True)
    g = nr.Graph.load(graph_path(name), s, t)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 23, rule="pivot")
    o = OracleGraph.load(graph_path(name)).solve(s, t, mode=MODE_PIVOT, order=ORDER_LEVEL)
    check(r, o)


def test_mass_order_and_truncation(expected):
    print("case pivot K15 truncated, level and mass order", flush=True)
    g = nr.Graph.load(graph_path("COMP15-05.txt"), 0, 14)
    og = OracleGraph.load(graph_path("COMP15-05.txt"))
    o = og.solve(0, 14, mode=MODE_PIVOT, order=ORDER_LEVEL, trunc_eps=1e-5)
    r = nr.solve(g, truncation_eps=1e-5, queue_capacity=1 << 23, rule="pivot")
    check(r, o)
    m = nr.solve(g, truncation_eps=1e-5, queue_capacity=1 << 23, rule="pivot", order="mass")
    assert (m["paths"], m["nodes"], m["dead"], m["truncated"]) == (
        o["paths"], o["nodes"], o["dead"], o["truncated"])
    R = closed_form(expected, 15, "0.5")
    assert r["R"] <= R * (1 + 1e-12) <= (r["R"] + r["ESP"]) * (1 + 2e-12)


def test_csr_form_and_terminal_sets():
    print("case pivot grid (CSR) and K-terminal", flush=True)
    rng = wl.Xoshiro256ss(0x6121D)
    n, edges, _ = wl.grid_graph(10, 10, [0.0] * 1000)
    probs = [0.85 + 0.10 * rng.uniform() for _ in edges]
    rp, ci, ei = wl.to_csr(n, edges)
    g = nr.Graph.from_csr(rp, ci, ei, probs, 0, 99)
    og = OracleGraph.from_csr(rp, ci, ei, probs)
    r = nr.solve(g, truncation_eps=1e-4, queue_capacity=1 << 22, rule="pivot")
    o = og.solve(0, 99, mode=MODE_PIVOT, order=ORDER_LEVEL, trunc_eps=1e-4)
    check(r, o)
    k8 = nr.Graph.load(graph_path("COMP8-09.txt"), 0, 7)
    k8.set_terminal_set([0, 3, 7])
    r = nr.solve(k8, truncation_eps=0.0, queue_capacity=1 << 23, rule="pivot")
    o = OracleGraph.load(graph_path("COMP8-09.txt")).solve_terms([0, 3, 7], mode=MODE_PIVOT,
                                                                 order=ORDER_LEVEL)
    check(r, o)


def test_checkpoint_chain(tmp_path):
    print("case pivot checkpoint chain K8", flush=True)
    g = nr.Graph.load(graph_path("COMP8.txt"), 0, 7)
    kw = dict(truncation_eps=0.0, queue_capacity=1 << 23, rule="pivot")
    full = nr.solve(g, **kw)
    ck, links = None, 0
    while True:
        r, ck = nr.solve_checkpointed(g, resume=ck, time_budget_s=2e-4, **kw)
        links += 1
        if ck is None:
            break
        assert links < 4000
    assert links >= 2
    assert (r["paths"], r["nodes"], r["iterations"]) == (
        full["paths"], full["nodes"], full["iterations"])
