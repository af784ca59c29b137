"""GPU: max-mass class rounds that split the pool first (k_class_split).

A round that carries or gives up a quarter or more of the pool runs a split
pass first: items given up go to ESP, carried items to the output, and the
items to expand are compacted to the front of their segments, so the round's
waves expand only.  The tree is the one level order builds (seq/seq.cpp's
heap order made bulk-synchronous, every order an exact decomposition), so
without a binding capacity paths, nodes, dead and truncated equal level
order's and R, ESP and D agree within 1e-12 relative; with a capacity that
binds, the brackets hold.  Forms: rows carried by items (K10, K15), the lean
row form (K25), CSR (10x10 grid), the alternating rule.
"""
import pytest

from conftest import closed_form, graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from relaibilitygpu_amd import learn  # noqa: E402

REL = 1e-12


def close(a, b):
    return abs(a - b) <= REL * max(abs(a), abs(b)) + 1e-300


def same_tree(g, **kw):
    lv = nr.solve(g, order="level", **kw)
    ms = nr.solve(g, order="mass", **kw)
    for key in ("paths", "nodes", "dead", "truncated"):
        assert lv[key] == ms[key], (key, lv[key], ms[key])
    for key in ("R", "ESP", "D"):
        assert close(lv[key], ms[key]), (key, lv[key], ms[key])
    assert ms["overflowed"] == 0 and ms["folded"] == 0
    return ms


def test_k10_rows_full(expected):
    print("case K10 mass order, full enumeration (items carry rows)", flush=True)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    r = same_tree(g, truncation_eps=0.0, queue_capacity=1 << 23)
    assert (r["paths"], r["nodes"]) == (7433032, 14866064)
    assert close(r["R"], closed_form(expected, 10, "0.5"))


@pytest.mark.parametrize("name,t,eps", [("COMP15-05.txt", 14, 1e-10), ("COMP25-05.txt", 24, 1e-10)])
def test_complete_truncated(name, t, eps):
    print(f"case {name} mass order eps={eps}", flush=True)
    g = nr.Graph.load(graph_path(name), 0, t)
    same_tree(g, truncation_eps=eps, queue_capacity=1 << 26)


@pytest.mark.parametrize("eps", [1e-9, 1e-10])
def test_grid_csr(eps):
    print(f"case grid 10x10 mass order eps={eps}", flush=True)
    same_tree(learn.grid(10, 10, 0x6121D), truncation_eps=eps, queue_capacity=96777216)


def test_alternating_rule():
    print("case K15 alternating rule mass order 1e-13", flush=True)
    g = nr.Graph.load(graph_path("COMP15-05.txt"), 0, 14)
    same_tree(g, truncation_eps=1e-13, queue_capacity=1 << 26, rule="alternate", sweep_period=6)


@pytest.mark.parametrize("cap", [1 << 20, 1 << 22])
def test_capacity_binding_brackets(cap):
    """Rounds that give up the lightest classes (split role 'given up'):
    R + D + ESP = 1 and the bracket holds around level order's without a
    binding capacity."""
    print(f"case grid 10x10 mass order eps=1e-11 capacity {cap}", flush=True)
    g = learn.grid(10, 10, 0x6121D)
    ref = nr.solve(g, truncation_eps=1e-9, queue_capacity=96777216)
    m = nr.solve(g, truncation_eps=1e-11, queue_capacity=cap, order="mass")
    assert m["folded"] > 0
    assert abs(m["R"] + m["D"] + m["ESP"] - 1.0) < 1e-12
    assert m["R"] <= ref["R"] + ref["ESP"] + 1e-12
    assert ref["R"] <= m["R"] + m["ESP"] + 1e-12



@pytest.mark.parametrize("case", ["K10", "grid"])
def test_learned_order_rounds(case):
    """Learned order: class rounds on the scorer's classes (the items' tags),
    split by tag; a random scorer still builds level order's tree."""
    import numpy as np

    print(f"case learned order, random scorer, {case}", flush=True)
    rng = np.random.default_rng(5)
    m = nr.Model(rng.normal(0, 0.3, nr.MLP_PARAMS).astype(np.float32), alpha=1.0)
    if case == "K10":
        g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
        kw = dict(truncation_eps=0.0, queue_capacity=1 << 23)
    else:
        g = learn.grid(10, 10, 0x6121D)
        kw = dict(truncation_eps=1e-9, queue_capacity=96777216)
    lv = nr.solve(g, order="level", **kw)
    r = nr.solve_learned(g, m, **kw)
    for key in ("paths", "nodes", "dead", "truncated"):
        assert lv[key] == r[key], (key, lv[key], r[key])
    for key in ("R", "ESP", "D"):
        assert close(lv[key], r[key]), (key, lv[key], r[key])


@pytest.mark.parametrize("grid,block", [(1, 256), (7, 64), (300, 128), (5000, 256)])
def test_launch_shapes(grid, block):
    """Split and round launches at any grid and block size: one workgroup
    (one segment), odd grids, 64-thread blocks (one wave), more workgroups
    than the device holds at once."""
    print(f"case launch shape grid {grid} block {block}", flush=True)
    g = nr.Graph.load(graph_path("COMP9.txt"), 0, 8)
    same_tree(g, truncation_eps=1e-7, queue_capacity=1 << 22, grid_blocks=grid, block_size=block)
