"""GPU parity: the HIP engine (through the C-ABI) against the CPU oracle.

Bar: bit-exact paths / nodes / dead / truncated counts, levels and widest
level in level order; R within 1e-12 relative of the oracle and of the closed
form; child items of single expansions bit-identical (planes and FP64 mass).
"""
import os
import random

import numpy as np
import pytest

from conftest import closed_form, graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from oracle_ffi import ORDER_DFS, ORDER_LEVEL, OracleGraph  # noqa: E402

REL = 1e-12


def pair(name, s, t):
    return nr.Graph.load(graph_path(name), s, t), OracleGraph.load(graph_path(name))


def check_same(r, o, levels=True):
    assert r["paths"] == o["paths"]
    assert r["nodes"] == o["nodes"]
    assert r["dead"] == o["dead"]
    assert r["truncated"] == o["truncated"]
    if levels:
        assert r["iterations"] == o["iterations"]
        assert r["max_queue_items"] == o["max_width"]
    assert abs(r["R"] - o["R"]) <= REL * max(abs(o["R"]), 1e-300) + 1e-300
    assert abs(r["ESP"] - o["ESP"]) <= REL * max(abs(o["ESP"]), 1e-300) + 1e-300


@pytest.mark.parametrize("name,s,t", [("test_simple.txt", 0, 3), ("test_bridge.txt", 0, 3),
                                      ("test_kite.txt", 0, 4), ("test_kite.txt", 0, 3),
                                      ("cycle4.txt", 0, 2), ("temp4x4.txt", 0, 3),
                                      ("test_kite.txt", 1, 1)])
def test_small_graphs(name, s, t):
    print(f"case {name} {s}->{t}", flush=True)
    g, og = pair(name, s, t)
    r = nr.solve(g, truncation_eps=0.0)
    o = og.solve(s, t)
    check_same(r, o)


def test_k8_p09_config1(expected):
    """K8, p = 0.9, 0 -> 7, full enumeration."""
    print("case K8 p=0.9 full", flush=True)
    g, og = pair("COMP8-09.txt", 0, 7)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 16)
    o = og.solve(0, 7)
    check_same(r, o)
    assert (r["paths"], r["nodes"], r["iterations"], r["max_queue_items"]) == (
        26830, 53660, 23, 8160)
    R = closed_form(expected, 8, "0.9")
    assert abs(r["R"] - R) <= REL * R
    assert abs(r["R"] + r["D"] - 1.0) <= 1e-13


def test_k10_full_enumeration(expected):
    """K10, p = 0.5: the logged tree and the exact R."""
    print("case K10 p=0.5 full", flush=True)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22)
    assert (r["paths"], r["nodes"], r["iterations"], r["max_queue_items"]) == (
        7433032, 14866064, 38, 1596000)
    assert r["dead"] == 7433032 and r["ESP"] == 0.0
    R = closed_form(expected, 10, "0.5")
    assert abs(r["R"] - R) <= REL * R
    assert abs(r["R"] + r["D"] - 1.0) <= 1e-13


@pytest.mark.parametrize("name,n", [("COMP9.txt", 9), ("COMP8-03.txt", 8)])
def test_complete_graphs_vs_oracle(name, n):
    print(f"case {name} full", flush=True)
    g, og = pair(name, 0, n - 1)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 20)
    o = og.solve(0, n - 1)
    check_same(r, o)


@pytest.mark.parametrize("eps", [1e-7, 1e-9])
def test_truncated_counts_bit_exact(expected, eps):
    """Truncation depends on each item's mass only, so counts are exact."""
    print(f"case K15 trunc eps={eps}", flush=True)
    g, og = pair("COMP15-05.txt", 0, 14)
    r = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 24)
    o = og.solve(0, 14, order=ORDER_LEVEL, trunc_eps=eps)
    check_same(r, o)
    Rt = closed_form(expected, 15, "0.5")
    assert r["R"] <= Rt <= r["R"] + r["ESP"] + 1e-14


def test_k25_truncated_bracket(expected):
    """K25 (300 edges, 5 mask words): counts vs the oracle, bracket vs the
    closed form."""
    print("case K25 trunc eps=1e-8", flush=True)
    g, og = pair("COMP25-05.txt", 0, 24)
    r = nr.solve(g, truncation_eps=1e-8, queue_capacity=1 << 22)
    o = og.solve(0, 24, order=ORDER_LEVEL, trunc_eps=1e-8)
    check_same(r, o)
    Rt = closed_form(expected, 25, "0.5")
    assert r["R"] <= Rt <= r["R"] + r["ESP"] + 1e-14


@pytest.mark.parametrize("cap", [1, 64, 5000])
def test_overflow_bracket(expected, cap):
    print(f"case K9 overflow cap={cap}", flush=True)
    g = nr.Graph.load(graph_path("COMP9.txt"), 0, 8)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=cap)
    Rt = closed_form(expected, 9, "0.5")
    assert r["overflowed"] > 0
    assert r["R"] <= Rt <= r["R"] + r["ESP"] + 1e-14
    assert abs(r["R"] + r["D"] + r["ESP"] - 1.0) <= 1e-13


def _random_items(rng, E, W, count, p_down, p_up):
    known = np.zeros((count, W), np.uint64)
    up = np.zeros((count, W), np.uint64)
    for i in range(count):
        for e in range(E):
            s = rng.random()
            if s < p_down:
                known[i, e >> 6] |= np.uint64(1 << (e & 63))
            elif s < p_down + p_up:
                known[i, e >> 6] |= np.uint64(1 << (e & 63))
                up[i, e >> 6] |= np.uint64(1 << (e & 63))
    mass = np.array([rng.uniform(1e-6, 1.0) for _ in range(count)])
    return known, up, mass


@pytest.mark.parametrize("name,s,t,count", [("COMP10-05.txt", 0, 9, 3000),
                                             ("COMP15-05.txt", 0, 14, 2000),
                                             ("COMP25-05.txt", 0, 24, 1000),
                                             ("COMP25-05.txt", 3, 17, 500),
                                             ("test_kite.txt", 0, 4, 300)])
def test_single_expansions_bit_exact(name, s, t, count):
    """Random edge states: every child (planes + FP64 mass) and success term
    of the GPU expansion equals the oracle's bit for bit."""
    print(f"case expand {name} {s}->{t} x{count}", flush=True)
    g, og = pair(name, s, t)
    n, E = og.dims()["n"], og.dims()["e"]
    W = og.mask_words
    rng = random.Random(sum(name.encode()) * 1000 + s * 31 + t)
    known, up, mass = _random_items(rng, E, W, count, p_down=0.45, p_up=0.15)
    ex = nr.expand_items(g, known, up, mass, max_children=n)
    alive = 0
    for i in range(count):
        k, succ, ck, cu, cm, _ = og.expand(s, t, known[i], up[i], mass[i])
        assert ex.nchild[i] == k, i
        if k < 0:
            continue
        alive += 1
        assert ex.succ[i] == succ
        assert np.array_equal(ex.child_known[i, :k], ck)
        assert np.array_equal(ex.child_up[i, :k], cu)
        assert np.array_equal(ex.child_mass[i, :k], cm)
    assert alive > count // 10


def test_wide_graph_runs_one_wave_per_item(expected):
    """K40 (780 edges, past the per-lane forms) runs in the wave-per-item
    form: counts, levels and widest level equal the oracle's fixed-size
    engine; orders the form lacks fail loudly."""
    print("case K40 wave-per-item form", flush=True)
    g, og = pair("COMP40.txt", 0, 39)
    for eps in (1e-3, 1e-4):
        r = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 20)
        o = og.solve(0, 39, trunc_eps=eps, capacity=1 << 20)
        assert (r["paths"], r["nodes"], r["dead"], r["truncated"]) == (
            o["paths"], o["nodes"], o["dead"], o["truncated"])
        assert (r["iterations"], r["max_queue_items"]) == (o["iterations"], o["max_width"])
        assert abs(r["R"] - o["R"]) <= REL * o["R"] and abs(r["ESP"] - o["ESP"]) <= REL * o["ESP"]
        Rt = closed_form(expected, 40, "0.1")
        assert r["R"] <= Rt <= r["R"] + r["ESP"] + 1e-14
    with pytest.raises(nr.NRError) as ei:
        nr.solve(g, truncation_eps=1e-3, queue_capacity=1024, order="mass")
    assert ei.value.status == 5


# ---------------------------------------------------------------- mass order
@pytest.mark.parametrize("name,n", [("COMP8-09.txt", 8), ("COMP9.txt", 9), ("COMP10-05.txt", 10)])
def test_mass_order_full_enumeration_counts(expected, name, n):
    """Max-mass class rounds visit the same tree: counts bit-exact, R to 1e-12."""
    print(f"case mass-order full {name}", flush=True)
    g = nr.Graph.load(graph_path(name), 0, n - 1)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 23, order="mass")
    lv = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 23, order="level")
    assert (r["paths"], r["nodes"], r["dead"]) == (lv["paths"], lv["nodes"], lv["dead"])
    assert r["ESP"] == 0.0 and r["overflowed"] == 0 and r["folded"] == 0
    assert abs(r["R"] - lv["R"]) <= REL * lv["R"]
    assert abs(r["R"] + r["D"] - 1.0) <= 1e-13


@pytest.mark.parametrize("eps", [1e-7, 1e-9])
def test_mass_order_truncated_counts(eps):
    print(f"case mass-order K15 eps={eps}", flush=True)
    g, og = pair("COMP15-05.txt", 0, 14)
    r = nr.solve(g, truncation_eps=eps, queue_capacity=1 << 24, order="mass")
    o = og.solve(0, 14, order=ORDER_LEVEL, trunc_eps=eps)
    check_same(r, o, levels=False)


@pytest.mark.parametrize("cap", [64, 2000, 50000])
def test_mass_order_capacity_bracket(expected, cap):
    print(f"case mass-order K10 cap={cap}", flush=True)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=cap, order="mass")
    Rt = closed_form(expected, 10, "0.5")
    assert r["overflowed"] + r["folded"] > 0
    assert r["R"] <= Rt <= r["R"] + r["ESP"] + 1e-14
    assert abs(r["R"] + r["D"] + r["ESP"] - 1.0) <= 1e-13


@pytest.mark.parametrize("budget", [0.01, 1e-6])
@pytest.mark.parametrize("order", ["level", "mass"])
def test_time_budget_bracket(expected, order, budget):
    """A 10 ms budget on K15 (delta 1e-15, a run of several hundred ms), and
    a 1 us one that expires before the first level: the queue is folded
    into ESP and the bracket R <= R_true <= R + ESP holds."""
    print(f"case deadline K15 {order} budget {budget}", flush=True)
    g = nr.Graph.load(graph_path("COMP15-05.txt"), 0, 14)
    r = nr.solve(g, truncation_eps=1e-15, queue_capacity=1 << 24, order=order,
                 time_budget_s=budget)
    Rt = closed_form(expected, 15, "0.5")
    assert r["deadline_hit"] == 1
    assert r["R"] <= Rt <= r["R"] + r["ESP"] + 1e-14
    assert abs(r["R"] + r["D"] + r["ESP"] - 1.0) <= 1e-12


def test_mass_order_beats_level_order_under_time_pressure(expected):
    """With a budget far below the run time (K15, delta 1e-15), max-mass
    order resolves heavier subproblems first and leaves a smaller gap."""
    print("case mass vs level K15 20 ms", flush=True)
    g = nr.Graph.load(graph_path("COMP15-05.txt"), 0, 14)
    kw = dict(truncation_eps=1e-15, queue_capacity=1 << 24, time_budget_s=0.02)
    lv = nr.solve(g, order="level", **kw)
    ms = nr.solve(g, order="mass", **kw)
    print(f"ESP level {lv['ESP']:.3e} mass {ms['ESP']:.3e}", flush=True)
    assert ms["ESP"] < lv["ESP"]
