"""GPU parity of the learned path choice (NR_RULE_LEARNED) against the
oracle's learned mode, and what the trained chooser gains.

The chooser sweeps the alternating rule's fewest-UNKNOWN path from the end
it predicts the smaller subtree for (by more than its margin).  Its
inference form is exact (integer features, weights on a 2^-8 grid), so the
GPU's matrix-core scores and the oracle's integer scores pick the same
direction everywhere: counts bit-exact (paths, nodes, dead, truncated;
levels and widest level in level order), R and ESP to 1e-12.  The choice
log's tree equals the oracle's exploring mode.  The rule has no counterpart
in the reference; the oracle's choice is pinned against an independent
restatement of the features in tests/test_oracle.py.
"""
import numpy as np
import pytest

from conftest import closed_form, graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from relaibilitygpu_amd import learn  # noqa: E402
from relaibilitygpu_amd import workloads as wl  # noqa: E402
from oracle_ffi import (MODE_ALTERNATE, MODE_EXPLORE, MODE_LEARNED, ORDER_DFS,  # noqa: E402
                        ORDER_LEVEL, OracleGraph)

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


def random_model(seed, margin=0.0):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.3, nr.MLP_PARAMS).astype(np.float32)
    return nr.Model(None, chooser=p, margin=margin)


@pytest.mark.parametrize("name,s,t,eps,period", [("COMP8-09.txt", 0, 7, 0.0, 2),
                                                 ("COMP10-05.txt", 0, 9, 1e-9, 3),
                                                 ("test_kite.txt", 0, 4, 0.0, 1)])
def test_choice_log_tree_is_the_oracles_exploring_tree(name, s, t, eps, period):
    print(f"case choice log {name} eps={eps} period {period}", flush=True)
    g = nr.Graph.load(graph_path(name), s, t)
    X, Y, r = nr.log_choices(g, 100000, seed=5, truncation_eps=eps, queue_capacity=1 << 22,
                             sweep_period=period)
    o = OracleGraph.load(graph_path(name)).solve(s, t, mode=MODE_EXPLORE, order=ORDER_LEVEL,
                                                 trunc_eps=eps, sweep_period=period)
    check(r, o)
    if len(X):
        assert np.array_equal(X, np.rint(X)) and np.abs(X).max() <= 63
        assert (X[:, 15] == 1).all() and (X[:, 9] == 0).all()
        assert X[:, 10].min() >= 0 and X[:, 10].max() <= 32
        assert (X[:, 0] >= 2).all()          # only items with a choice to make
        assert Y.min() >= np.log2(3) - 1e-6  # self + at least two children


@pytest.mark.parametrize("seed,margin", [(1, 0.0), (2, 0.0), (3, 0.25)])
def test_random_choosers_k10_k15(expected, seed, margin):
    """Random chooser weights: every decision is some decision, and the GPU
    and the oracle must take the same one everywhere."""
    print(f"case random chooser {seed} margin {margin}", flush=True)
    m = random_model(seed, margin)
    ch = m.chooser_weights()
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    r = nr.solve_learned(g, m, rule="learned", order="level", truncation_eps=0.0,
                         queue_capacity=1 << 22, sweep_period=2)
    o = OracleGraph.load(graph_path("COMP10-05.txt")).solve(
        0, 9, mode=MODE_LEARNED, order=ORDER_LEVEL, sweep_period=2, chooser=ch)
    check(r, o)
    assert close(r["R"], closed_form(expected, 10, "0.5"))
    g = nr.Graph.load(graph_path("COMP15-05.txt"), 0, 14)
    r = nr.solve_learned(g, m, rule="learned", order="level", truncation_eps=1e-9,
                         queue_capacity=1 << 26, sweep_period=4)
    o = OracleGraph.load(graph_path("COMP15-05.txt")).solve(
        0, 14, mode=MODE_LEARNED, order=ORDER_DFS, threads=16, trunc_eps=1e-9, sweep_period=4,
        chooser=ch)
    check(r, o, levels=False)


def test_huge_margin_is_the_alternating_rule():
    print("case chooser margin 1e9", flush=True)
    m = random_model(4, margin=1e9)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    a = nr.solve(g, rule="alternate", truncation_eps=0.0, queue_capacity=1 << 22, sweep_period=3)
    b = nr.solve_learned(g, m, rule="learned", order="level", truncation_eps=0.0,
                         queue_capacity=1 << 22, sweep_period=3)
    assert (a["paths"], a["nodes"], a["iterations"], a["max_queue_items"]) == (
        b["paths"], b["nodes"], b["iterations"], b["max_queue_items"])


def test_trained_chooser_vs_oracle_dense(expected, trained_model):
    model, info = trained_model
    ch = model.chooser_weights()
    print(f"case trained chooser dense (margin {info['margin']})", flush=True)
    for name, n, eps, period in [("COMP8-09.txt", 8, 0.0, 6), ("COMP10-05.txt", 10, 0.0, 6),
                                 ("COMP15-05.txt", 15, 1e-10, 6)]:
        g = nr.Graph.load(graph_path(name), 0, n - 1)
        og = OracleGraph.load(graph_path(name))
        r = nr.solve_learned(g, model, rule="learned", order="level", truncation_eps=eps,
                             queue_capacity=1 << 26, sweep_period=period)
        if n <= 10:
            o = og.solve(0, n - 1, mode=MODE_LEARNED, order=ORDER_LEVEL, trunc_eps=eps,
                         sweep_period=period, chooser=ch)
            check(r, o)
        else:
            o = og.solve(0, n - 1, mode=MODE_LEARNED, order=ORDER_DFS, threads=16,
                         trunc_eps=eps, sweep_period=period, chooser=ch)
            check(r, o, levels=False)
        # max-mass and learned order expand the same tree when nothing binds
        for order in ("mass", "learned"):
            q = nr.solve_learned(g, model, rule="learned", order=order, truncation_eps=eps,
                                 queue_capacity=1 << 26, sweep_period=period)
            assert (q["paths"], q["nodes"], q["dead"], q["truncated"]) == (
                r["paths"], r["nodes"], r["dead"], r["truncated"]), order
            assert close(q["R"], r["R"], 1e-11)
    r = nr.solve_learned(nr.Graph.load(graph_path("COMP8-09.txt"), 0, 7), model, rule="learned",
                         order="level", truncation_eps=0.0, queue_capacity=1 << 20)
    assert close(r["R"], closed_form(expected, 8, "0.9"))


def test_trained_chooser_vs_oracle_sparse(trained_model):
    """The CSR form (65-256 nodes) and the lean form on ER instances."""
    model, _ = trained_model
    ch = model.chooser_weights()
    print("case trained chooser sparse", flush=True)
    graphs = [(learn.grid(10, 10, 0x6121D), 1e-6, 1), (learn.barabasi_albert(150, 0xBA150), 1e-9, 1)]
    for i in (3, 40):
        n, edges, probs, s, t = wl.er_instance(i)
        rp, ci, ei = wl.to_csr(n, edges)
        graphs.append((nr.Graph.from_csr(rp, ci, ei, probs, s, t), 1e-10, 2))
    for g, eps, period in graphs:
        gi = g.info()
        rp, ci, ei, p = g.csr()
        og = OracleGraph.from_csr(rp, ci, ei, p)
        r = nr.solve_learned(g, model, rule="learned", order="level", truncation_eps=eps,
                             queue_capacity=1 << 26, sweep_period=period)
        o = og.solve(gi.src, gi.dst, mode=MODE_LEARNED, order=ORDER_LEVEL, trunc_eps=eps,
                     sweep_period=period, chooser=ch)
        check(r, o)


def test_trained_chooser_gains_on_graphs_it_never_saw(trained_model):
    """Graphs used neither for training (K10-K13) nor for the margin's
    validation (K14, 7x7 grid, BA(100)): the learned choice leaves a smaller
    certified gap than the alternating rule it improves on, at the same
    truncation, and the bracket holds."""
    model, info = trained_model
    print(f"case chooser gain (validation gain {info['validation_gain']:.3f})", flush=True)
    cases = [(nr.Graph.load(graph_path("COMP15-05.txt"), 0, 14), 6, 1e-12),
             (learn.barabasi_albert(150, 0xBA150), 1, 1e-12)]
    for g, period, eps in cases:
        a = nr.solve(g, rule="alternate", sweep_period=period, truncation_eps=eps,
                     queue_capacity=96777216)
        b = nr.solve_learned(g, model, rule="learned", order="level", sweep_period=period,
                             truncation_eps=eps, queue_capacity=96777216)
        print(f"  alternate ESP {a['ESP']:.4e} ({a['nodes']} nodes), learned ESP "
              f"{b['ESP']:.4e} ({b['nodes']} nodes)", flush=True)
        assert b["ESP"] < a["ESP"]
        assert max(a["R"], b["R"]) <= min(a["R"] + a["ESP"], b["R"] + b["ESP"]) * (1 + 1e-12)
