"""GPU tests of checkpoint / resume (nr_solve_checkpointed).

A chain of time-budgeted solves, each resuming the previous one's pending
subproblems (through a file for some links), must end with the paths,
nodes, dead and truncated counts, levels and widest level of one
uninterrupted solve, and R within 1e-12; every intermediate result must be
a certified bracket R <= R_true <= R + ESP with R + ESP + D = 1.
"""
import os

import pytest

from conftest import closed_form, graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from relaibilitygpu_amd import workloads as wl  # noqa: E402


def chain(g, tmp_path, budget, max_links=4000, save_every=3, **kw):
    """Run the chain; returns (final result, links, intermediate results)."""
    ck = None
    mids = []
    for link in range(max_links):
        r, ck = nr.solve_checkpointed(g, resume=ck, time_budget_s=budget, **kw)
        if ck is None:
            return r, link + 1, mids
        mids.append(r)
        if link % save_every == 0:  # through a file
            path = os.path.join(str(tmp_path), "ck.bin")
            ck.save(path)
            ck = nr.Checkpoint.load(path)
    raise AssertionError("chain did not finish")


def same(a, b, levels=True):
    assert (a["paths"], a["nodes"], a["dead"], a["truncated"]) == (
        b["paths"], b["nodes"], b["dead"], b["truncated"])
    if levels:
        assert (a["iterations"], a["max_queue_items"]) == (b["iterations"], b["max_queue_items"])
    assert abs(a["R"] - b["R"]) <= 1e-12 * abs(b["R"])
    assert abs(a["ESP"] - b["ESP"]) <= 1e-12 * abs(b["ESP"]) + 1e-300


def test_k10_level_chain(tmp_path, expected):
    print("case checkpoint K10 level chain", flush=True)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    full = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22)
    r, links, mids = chain(g, tmp_path, 2e-4, truncation_eps=0.0, queue_capacity=1 << 22)
    assert links >= 3
    same(r, full)
    R = closed_form(expected, 10, "0.5")
    assert (r["paths"], r["nodes"], r["iterations"], r["max_queue_items"]) == (
        7433032, 14866064, 38, 1596000)
    for m in mids:
        assert m["deadline_hit"] == 1
        assert m["R"] <= R * (1 + 1e-12) <= (m["R"] + m["ESP"]) * (1 + 2e-12)
        assert abs(m["R"] + m["ESP"] + m["D"] - 1.0) <= 1e-12
    # the nodes expanded so far only grow along the chain
    done = [m["nodes"] for m in mids]
    assert done == sorted(done)


def test_k15_mass_chain(tmp_path, expected):
    print("case checkpoint K15 mass chain", flush=True)
    g = nr.Graph.load(graph_path("COMP15-05.txt"), 0, 14)
    kw = dict(truncation_eps=1e-10, queue_capacity=1 << 26, order="mass")
    full = nr.solve(g, **kw)
    r, links, mids = chain(g, tmp_path, 3e-4, **kw)
    assert links >= 2
    same(r, full, levels=False)
    R = closed_form(expected, 15, "0.5")
    for m in mids:
        assert m["R"] <= R * (1 + 1e-12) and m["R"] + m["ESP"] >= R * (1 - 1e-12)


def test_forkbest_and_csr_chains(tmp_path):
    print("case checkpoint forkbest K9 and grid", flush=True)
    g = nr.Graph.load(graph_path("COMP9.txt"), 0, 8)
    kw = dict(truncation_eps=0.0, queue_capacity=1 << 22, rule="forkbest")
    r, links, _ = chain(g, tmp_path, 1e-4, **kw)
    assert links >= 2
    same(r, nr.solve(g, **kw))
    rng = wl.Xoshiro256ss(0x6121D)
    n, edges, _ = wl.grid_graph(10, 10, [0.0] * 1000)
    probs = [0.85 + 0.10 * rng.uniform() for _ in edges]
    rp, ci, ei = wl.to_csr(n, edges)
    grid = nr.Graph.from_csr(rp, ci, ei, probs, 0, 99)
    kw = dict(truncation_eps=1e-6, queue_capacity=1 << 22)
    r, links, _ = chain(grid, tmp_path, 1e-4, **kw)
    assert links >= 2
    same(r, nr.solve(grid, **kw))


def test_checkpoint_info_and_rejects(tmp_path):
    print("case checkpoint info / rejects", flush=True)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    r, ck = nr.solve_checkpointed(g, time_budget_s=2e-4, truncation_eps=0.0,
                                  queue_capacity=1 << 22)
    assert ck is not None
    info = ck.info()
    assert info["pending"] > 0 and info["nodes_done"] == r["nodes"]
    assert r["ESP"] >= info["pending_mass"] > 0
    assert abs(r["R"] + r["ESP"] + r["D"] - 1.0) <= 1e-12
    other = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 8)   # other terminal
    with pytest.raises(nr.NRError):
        nr.solve_checkpointed(other, resume=ck, truncation_eps=0.0, queue_capacity=1 << 22)
    with pytest.raises(nr.NRError):                              # other order
        nr.solve_checkpointed(g, resume=ck, truncation_eps=0.0, queue_capacity=1 << 22,
                              order="mass")
    with pytest.raises(nr.NRError) as ei:                        # other truncation eps
        nr.solve_checkpointed(g, resume=ck, truncation_eps=1e-9, queue_capacity=1 << 22)
    assert ei.value.status == 4
    # without a budget nothing is kept
    r2, ck2 = nr.solve_checkpointed(g, resume=ck, truncation_eps=0.0, queue_capacity=1 << 22)
    assert ck2 is None and r2["nodes"] == 14866064
