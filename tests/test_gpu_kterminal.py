"""GPU parity of K-terminal reliability (nr_graph_set_terminal_set) against
the oracle's terminal-tree rule (oracle/cpu_reference.cpp, find_tree; the
oracle is pinned by brute force, cycle closed forms and K = {s, t}).

Counts bit-exact, levels and widest level, R to 1e-12, on the row form
(K_n files), the CSR form (grid, directed graph) and with the forkBest
rule, max-mass order and checkpoint/resume."""
import pytest

from conftest import graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from relaibilitygpu_amd import workloads as wl  # noqa: E402
from oracle_ffi import MODE_FORKBEST, MODE_PATH, ORDER_LEVEL, OracleGraph  # noqa: E402


def close(a, b, rel=1e-12):
    return abs(a - b) <= rel * max(abs(a), abs(b)) + 1e-300


def check(r, o):
    assert (r["paths"], r["nodes"], r["dead"], r["truncated"]) == (
        o["paths"], o["nodes"], o["dead"], o["truncated"])
    assert r["iterations"] == o["iterations"]
    assert r["max_queue_items"] == o["max_width"]
    assert close(r["R"], o["R"]) and close(r["ESP"], o["ESP"])


def test_pair_set_is_two_terminal():
    print("case K10 terminal set {0, 9}", flush=True)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    g.set_terminal_set([0, 9])
    assert g.terminal_set() == [0, 9]
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22)
    assert (r["paths"], r["nodes"], r["iterations"], r["max_queue_items"]) == (
        7433032, 14866064, 38, 1596000)
    g.set_terminals(0, 9)  # clears the set
    assert g.terminal_set() == []


@pytest.mark.parametrize("name,K,eps,rule", [
    ("COMP8-09.txt", [0, 3, 7], 0.0, "path"),
    ("COMP9.txt", [0, 3, 8], 0.0, "path"),
    ("COMP9.txt", [0, 3, 8], 0.0, "forkbest"),
    ("COMP10-05.txt", [0, 4, 9], 1e-9, "path"),
    ("COMP10-05.txt", [9, 0, 2, 5, 7], 1e-7, "path"),
])
def test_row_form_vs_oracle(name, K, eps, rule):
    print(f"case kterminal {name} {K} {rule}", flush=True)
    g = nr.Graph.load(graph_path(name), K[0], K[1])
    g.set_terminal_set(K)
    r = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 23, rule=rule)
    o = OracleGraph.load(graph_path(name)).solve_terms(
        K, mode=MODE_FORKBEST if rule == "forkbest" else MODE_PATH, order=ORDER_LEVEL,
        trunc_eps=eps)
    check(r, o)


@pytest.mark.parametrize("n", [5, 9, 16])
def test_all_terminal_cycles(n):
    print(f"case all-terminal C{n}", flush=True)
    p = 0.85
    edges = sorted((min(i, (i + 1) % n), max(i, (i + 1) % n)) for i in range(n))
    rp, ci, ei = wl.to_csr(n, edges)
    g = nr.Graph.from_csr(rp, ci, ei, [p] * n, 0, 1)
    g.set_terminal_set(list(range(n)))
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 16)
    assert close(r["R"], p ** n + n * p ** (n - 1) * (1 - p), 1e-13)


@pytest.mark.parametrize("kind", ["grid", "directed"])
def test_csr_form_vs_oracle(kind):
    print(f"case kterminal {kind}", flush=True)
    if kind == "grid":
        rng = wl.Xoshiro256ss(0x6121D)
        n, edges, _ = wl.grid_graph(10, 10, [0.0] * 1000)
        probs = [0.85 + 0.10 * rng.uniform() for _ in edges]
        rp, ci, ei = wl.to_csr(n, edges)
        K, eps = [0, 9, 90, 99], 1e-4
    else:
        n, edges, probs = wl.random_digraph(60, 2.5, 0.6, 0.95, 0xD1003)
        rp, ci, ei = wl.to_csr_directed(n, edges)
        K, eps = [0, 20, 40, 59], 1e-7
    g = nr.Graph.from_csr(rp, ci, ei, probs, K[0], K[1])
    g.set_terminal_set(K)
    og = OracleGraph.from_csr(rp, ci, ei, probs)
    for rule, mode in (("path", MODE_PATH), ("forkbest", MODE_FORKBEST)):
        r = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 22, rule=rule)
        check(r, og.solve_terms(K, mode=mode, order=ORDER_LEVEL, trunc_eps=eps))


def test_kterminal_mass_order_and_checkpoint(tmp_path):
    print("case kterminal mass order + checkpoint", flush=True)
    g = nr.Graph.load(graph_path("COMP9.txt"), 0, 8)
    g.set_terminal_set([0, 3, 8])
    lv = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22)
    m = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, order="mass")
    assert (m["paths"], m["nodes"], m["dead"]) == (lv["paths"], lv["nodes"], lv["dead"])
    assert close(m["R"], lv["R"])
    ck, links = None, 0
    while True:
        r, ck = nr.solve_checkpointed(g, resume=ck, time_budget_s=1e-4, truncation_eps=0.0,
                                      queue_capacity=1 << 22)
        links += 1
        if ck is None:
            break
        assert links < 2000
    assert (r["paths"], r["nodes"], r["iterations"]) == (lv["paths"], lv["nodes"],
                                                         lv["iterations"])
    assert close(r["R"], lv["R"])


def test_terminal_set_rejects():
    g = nr.Graph.load(graph_path("COMP8-09.txt"), 0, 7)
    with pytest.raises(nr.NRError):
        g.set_terminal_set([3])
    with pytest.raises(nr.NRError):
        g.set_terminal_set([0, 99])
    g.set_terminal_set([0, 3, 7])
    with pytest.raises(nr.NRError):
        nr.solve_batch([g], truncation_eps=1e-6)
