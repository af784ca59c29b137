"""Edge cases and the boundaries between the engine's forms, against the
oracle: node-set widths 32/33 and 64/65 (row form u32 / u64 / CSR), 256/257
nodes (CSR / one wave per item), 512/513 edges, 8/9 plane words; graphs with
no edges, isolated terminals, pre-failed (p = 0) and pre-contracted (p = 1)
edges, self-loops and degree-0 nodes.  Counts bit-exact, R to 1e-12.
"""
import random

import pytest

from conftest import graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from relaibilitygpu_amd import workloads as wl  # noqa: E402
from oracle_ffi import ORDER_LEVEL, OracleGraph  # noqa: E402

REL = 1e-12


def close(a, b):
    return abs(a - b) <= REL * max(abs(a), abs(b)) + 1e-300


def both(n, edges, probs, s, t):
    rp, ci, ei = wl.to_csr(n, edges)
    return nr.Graph.from_csr(rp, ci, ei, probs, s, t), OracleGraph.from_csr(rp, ci, ei, probs)


def check(g, og, s, t, eps, cap=1 << 22):
    r = nr.solve(g, truncation_eps=eps, queue_capacity=cap)
    o = og.solve(s, t, order=ORDER_LEVEL, trunc_eps=eps, capacity=cap)
    assert (r["paths"], r["nodes"], r["dead"], r["truncated"]) == (
        o["paths"], o["nodes"], o["dead"], o["truncated"])
    assert (r["iterations"], r["max_queue_items"]) == (o["iterations"], o["max_width"])
    assert close(r["R"], o["R"]) and close(r["ESP"], o["ESP"]) and close(r["D"], o["D"])
    assert abs(r["R"] + r["D"] + r["ESP"] - 1.0) <= 1e-12
    return r


def random_connected(n, extra, seed, plo=0.6, phi=0.95):
    rng = random.Random(seed)
    edges = set()
    for v in range(1, n):  # a random spanning tree, then extra edges
        edges.add((rng.randrange(v), v))
    while len(edges) < n - 1 + extra:
        a, b = rng.randrange(n), rng.randrange(n)
        if a != b:
            edges.add((min(a, b), max(a, b)))
    edges = sorted(edges)
    return edges, [rng.uniform(plo, phi) for _ in edges]


@pytest.mark.parametrize("n,extra,eps", [(32, 40, 1e-6), (33, 40, 1e-6), (64, 60, 1e-5),
                                         (65, 60, 1e-5), (256, 200, 1e-4), (257, 200, 1e-4)])
def test_node_set_width_boundaries(n, extra, eps):
    print(f"case boundary n={n}", flush=True)
    edges, probs = random_connected(n, extra, 1000 + n)
    g, og = both(n, edges, probs, 0, n - 1)
    check(g, og, 0, n - 1, eps)


@pytest.mark.parametrize("n,e", [(40, 512), (40, 513), (60, 576), (60, 577)])
def test_edge_count_boundaries(n, e):
    """512 edges = 8 plane words (per-lane forms); 513 = 9 words (one wave per
    item); 576/577 likewise around 9 words."""
    print(f"case boundary E={e}", flush=True)
    edges, probs = random_connected(n, e - (n - 1), 2000 + e, 0.2, 0.5)
    assert len(edges) == e
    g, og = both(n, edges, probs, 0, n - 1)
    check(g, og, 0, n - 1, 1e-3, cap=1 << 20)


def test_degenerate_graphs():
    print("case degenerate graphs", flush=True)
    # no edges at all: one dead root
    g, og = both(5, [], [], 0, 4)
    r = check(g, og, 0, 4, 0.0)
    assert r["R"] == 0.0 and r["nodes"] == 1 and r["dead"] == 1
    # isolated destination, degree-0 nodes
    edges = [(0, 1), (1, 2), (0, 2)]
    g, og = both(6, edges, [0.9, 0.8, 0.7], 0, 5)
    r = check(g, og, 0, 5, 0.0)
    assert r["R"] == 0.0
    # src == dst
    g, og = both(4, [(0, 1), (1, 2)], [0.5, 0.5], 2, 2)
    r = check(g, og, 2, 2, 0.0)
    assert r["R"] == 1.0


def test_prefailed_precontracted_and_self_loops():
    """p = 0 edges start FAILED, p = 1 edges WORKING (main.cu:717-748); a
    self-loop never lies on a path."""
    print("case p=0 / p=1 / self-loops", flush=True)
    n = 12
    edges, probs = random_connected(n, 20, 77)
    for i in range(0, len(edges), 5):
        probs[i] = 0.0
    for i in range(2, len(edges), 7):
        probs[i] = 1.0
    g, og = both(n, edges, probs, 0, n - 1)
    check(g, og, 0, n - 1, 0.0)
    # self-loops through nr_graph_create (the file format stores simple graphs)
    base = list(edges) + [(3, 3), (7, 7)]
    p2 = probs + [0.5, 0.9]
    rows = [[] for _ in range(n)]
    for i, (a, b) in enumerate(base):
        rows[a].append((b, i))
        if a != b:
            rows[b].append((a, i))
    rp, ci, ei = [0], [], []
    for u in range(n):
        for v, i in sorted(rows[u]):
            ci.append(v)
            ei.append(i)
        rp.append(len(ci))
    g2 = nr.Graph.from_csr(rp, ci, ei, p2, 0, n - 1)
    og2 = OracleGraph.from_csr(rp, ci, ei, p2)
    r2 = check(g2, og2, 0, n - 1, 0.0)
    r1 = nr.solve(g, truncation_eps=0.0)
    assert close(r2["R"], r1["R"])


def test_directed_graph_boundaries():
    print("case directed 65 and 300 nodes", flush=True)
    for n, seed in ((65, 0xD65), (300, 0xD300)):
        nn, edges, probs = wl.random_digraph(n, 2.5, 0.6, 0.95, seed)
        rp, ci, ei = wl.to_csr_directed(nn, edges)
        g = nr.Graph.from_csr(rp, ci, ei, probs, 0, nn - 1)
        og = OracleGraph.from_csr(rp, ci, ei, probs)
        check(g, og, 0, nn - 1, 1e-6)


def test_device_tables_follow_the_graph():
    """Repeated solves reuse the uploaded graph tables only while the graph
    (and its terminals) is the same: new terminals, a terminal set, another
    graph in between and a graph made at a freed graph's address all give the
    fresh graph's result."""
    print("case table reuse", flush=True)
    k8 = graph_path("COMP8.txt")
    k9 = graph_path("COMP9.txt")

    def fresh(path, s, t, terms=None):
        g = nr.Graph.load(path, s, t)
        if terms:
            g.set_terminal_set(terms)
        r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22)
        g.close()
        return r

    def same(r, o):
        assert (r["paths"], r["nodes"], r["iterations"]) == (o["paths"], o["nodes"], o["iterations"])
        assert r["R"] == o["R"]

    g = nr.Graph.load(k8, 0, 7)
    same(nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22), fresh(k8, 0, 7))
    same(nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22), fresh(k8, 0, 7))
    g.set_terminals(2, 5)
    same(nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22), fresh(k8, 2, 5))
    h = nr.Graph.load(k9, 0, 8)
    same(nr.solve(h, truncation_eps=0.0, queue_capacity=1 << 22), fresh(k9, 0, 8))
    same(nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22), fresh(k8, 2, 5))
    g.set_terminal_set([0, 3, 7])
    same(nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22), fresh(k8, 0, 7, [0, 3, 7]))
    g.close()
    h.close()
    for _ in range(3):  # new graphs, possibly at the freed addresses
        g2 = nr.Graph.load(k9, 1, 6)
        same(nr.solve(g2, truncation_eps=0.0, queue_capacity=1 << 22), fresh(k9, 1, 6))
        g2.close()


@pytest.mark.parametrize("block,grid", [(64, 37), (128, 100), (256, 1), (256, 5000), (64, 4097)])
def test_launch_shapes_same_tree(expected, block, grid):
    """Level launches at other workgroup sizes and grids (one workgroup, a
    grid past 4096 segments, more than eight segment counts per thread in
    the prefix, blocks of one and two waves) build the same tree: the
    reference log's K9 counts, and R equal to the default launch's."""
    print(f"case launch shape block={block} grid={grid}", flush=True)
    g = nr.Graph.load(graph_path("COMP9.txt"), 0, 8)
    base = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, block_size=block, grid_blocks=grid)
    assert (r["block_size"], r["grid_blocks"]) == (block, grid)
    lg = [x for x in expected["logged_counts"] if x["graph"] == "COMP9.txt"][0]
    assert (r["paths"], r["nodes"], r["iterations"], r["max_queue_items"]) == (
        lg["paths"], lg["nodes"], lg["iterations"], lg["max_width"])
    assert r["overflowed"] == 0
    assert close(r["R"], base["R"])


@pytest.mark.parametrize("form", ["csr", "alternate", "forkbest"])
@pytest.mark.parametrize("block,grid", [(64, 4097), (128, 77)])
def test_launch_shapes_other_forms(form, block, grid):
    """The CSR form (a directed 60-node graph), the alternating rule and
    forkBest (K9) at other launch shapes: counts equal the default launch's,
    R to 1e-12."""
    print(f"case launch shape {form} block={block} grid={grid}", flush=True)
    if form == "csr":
        n, edges, probs = wl.random_digraph(60, 2.5, 0.6, 0.95, 0xD1003)
        rp, ci, ei = wl.to_csr_directed(n, edges)
        g = nr.Graph.from_csr(rp, ci, ei, probs, 0, n - 1)
        kw = dict(truncation_eps=1e-7, queue_capacity=1 << 22)
    else:
        g = nr.Graph.load(graph_path("COMP9.txt"), 0, 8)
        kw = dict(truncation_eps=0.0, queue_capacity=1 << 22, rule=form)
    base = nr.solve(g, **kw)
    r = nr.solve(g, block_size=block, grid_blocks=grid, **kw)
    assert (r["block_size"], r["grid_blocks"]) == (block, grid)
    assert (r["paths"], r["nodes"], r["dead"], r["truncated"], r["iterations"],
            r["max_queue_items"]) == (base["paths"], base["nodes"], base["dead"],
                                      base["truncated"], base["iterations"], base["max_queue_items"])
    assert r["overflowed"] == 0 == base["overflowed"]
    assert close(r["R"], base["R"]) and close(r["ESP"], base["ESP"])
