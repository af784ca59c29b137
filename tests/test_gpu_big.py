"""GPU parity of the wave-per-item form (graphs of more than 256 nodes or
512 edges) against the oracle's restatement of the same path rule and level
order for graphs of any size.

Inputs: the reference's SNAP graph p2p-Gn08C (6301 nodes, 20777 arcs,
directed, p = 0.5; workingPad/p2p-Gn08C.txt, copied under tests/golden as
data) and a 40x40 undirected grid.  Counts bit-exact (paths, nodes, dead,
truncated, levels, widest level), R and ESP to 1e-12.
"""
import pytest

from conftest import graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from relaibilitygpu_amd import workloads as wl  # noqa: E402
from oracle_ffi import ORDER_LEVEL, OracleGraph  # noqa: E402

REL = 1e-12


def close(a, b):
    return abs(a - b) <= REL * max(abs(a), abs(b)) + 1e-300


def check(r, o):
    assert (r["paths"], r["nodes"], r["dead"], r["truncated"]) == (
        o["paths"], o["nodes"], o["dead"], o["truncated"])
    assert (r["iterations"], r["max_queue_items"]) == (o["iterations"], o["max_width"])
    assert close(r["R"], o["R"]) and close(r["ESP"], o["ESP"]) and close(r["D"], o["D"])


@pytest.mark.parametrize("dst,eps", [(17, 1e-3), (17, 1e-4), (71, 1e-4), (31, 1e-4)])
def test_snap_graph_vs_oracle(dst, eps):
    print(f"case big p2p-Gn08C 0->{dst} eps={eps}", flush=True)
    path = graph_path("p2p-Gn08C.txt")
    g = nr.Graph.load(path, 0, dst)
    assert g.info().n == 6301
    r = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 20)
    o = OracleGraph.load(path).solve(0, dst, order=ORDER_LEVEL, trunc_eps=eps, capacity=1 << 20)
    check(r, o)
    assert abs(r["R"] + r["D"] + r["ESP"] - 1.0) <= 1e-12


def test_large_undirected_grid_vs_oracle():
    print("case big grid 40x40", flush=True)
    rng = wl.Xoshiro256ss(0x40A40)
    n, edges, _ = wl.grid_graph(40, 40, [0.0] * 4000)
    probs = [0.85 + 0.10 * rng.uniform() for _ in edges]
    rp, ci, ei = wl.to_csr(n, edges)
    og = OracleGraph.from_csr(rp, ci, ei, probs)
    for dst, eps in ((410, 1e-6), (820, 1e-5)):
        g = nr.Graph.from_csr(rp, ci, ei, probs, 0, dst)
        r = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 21)
        o = og.solve(0, dst, order=ORDER_LEVEL, trunc_eps=eps, capacity=1 << 21)
        check(r, o)


def test_brackets_under_capacity_and_deadline():
    print("case big p2p capacity and deadline", flush=True)
    path = graph_path("p2p-Gn08C.txt")
    g = nr.Graph.load(path, 0, 17)
    full = nr.solve(g, truncation_eps=1e-5, queue_capacity=1 << 20)
    for kw in (dict(queue_capacity=1 << 12), dict(queue_capacity=1 << 20, time_budget_s=0.002)):
        r = nr.solve(g, truncation_eps=1e-5, **kw)
        assert abs(r["R"] + r["D"] + r["ESP"] - 1.0) <= 1e-12
        assert r["R"] <= full["R"] + full["ESP"] and full["R"] <= r["R"] + r["ESP"]


def test_big_form_limits():
    path = graph_path("p2p-Gn08C.txt")
    g = nr.Graph.load(path, 0, 17)
    for kw in (dict(order="mass"), dict(rule="alternate"), dict(rule="forkbest")):
        with pytest.raises(nr.NRError) as ei:
            nr.solve(g, truncation_eps=1e-3, **kw)
        assert ei.value.status == 5
