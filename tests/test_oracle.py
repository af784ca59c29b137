"""The CPU oracle against known answers (CPU only).

Pins, each chosen so that a plausible mistake in the oracle fails one:
  * closed form R(K_n) (exact rationals)            -> wrong mass products
  * brute force over all 2^E states                 -> wrong child construction
  * reference-logged paths/nodes/levels/width       -> wrong parent rule / counting
  * R + D + ESP = 1 in every mode                   -> lost or double-counted mass
  * counts equal across orders without limits       -> order-dependent tree
  * series / parallel / bridge graphs               -> textbook formulas
"""
import math
import random

import numpy as np
import pytest

from conftest import closed_form, graph_path
from oracle_ffi import (MODE_FORKBEST, MODE_PATH, MODE_PIVOT, ORDER_DFS, ORDER_HEAP,
                        ORDER_LEVEL, OracleGraph)


def load(name):
    return OracleGraph.load(graph_path(name))


@pytest.mark.parametrize("idx", range(7))
def test_brute_force_small_graphs(expected, idx):
    bf = expected["brute_force"][idx]
    g = load(bf["graph"])
    for order in (ORDER_LEVEL, ORDER_HEAP, ORDER_DFS):
        r = g.solve(bf["src"], bf["dst"], order=order)
        assert abs(r["R"] - bf["R"]) <= 1e-15, (bf, order, r)
        assert r["ESP"] == 0
        assert abs(r["R"] + r["D"] - 1.0) <= 1e-15


@pytest.mark.parametrize("idx", range(5))
def test_small_graph_counts(expected, idx):
    c = expected["derived_counts"][idx]
    r = load(c["graph"]).solve(c["src"], c["dst"])
    assert (r["paths"], r["nodes"], r["iterations"], r["max_width"]) == (
        c["paths"], c["nodes"], c["iterations"], c["max_width"])


@pytest.mark.parametrize("name,n,p", [("COMP8.txt", 8, "0.5"), ("COMP8-03.txt", 8, "0.3"),
                                      ("COMP8-09.txt", 8, "0.9"), ("temp4x4.txt", 4, "0.5")])
def test_closed_form_small(expected, name, n, p):
    r = load(name).solve(0, n - 1)
    R = closed_form(expected, n, p)
    assert abs(r["R"] - R) <= 1e-14 * R
    assert abs(r["R"] + r["D"] - 1.0) <= 1e-14


@pytest.mark.parametrize("name", ["COMP8.txt", "COMP9.txt"])
def test_logged_counts_level_order(expected, name):
    c = next(x for x in expected["logged_counts"] if x["graph"] == name)
    r = load(name).solve(c["src"], c["dst"])
    assert (r["paths"], r["nodes"], r["iterations"], r["max_width"]) == (
        c["paths"], c["nodes"], c["iterations"], c["max_width"]), c["source"]


def test_k10_counts_and_closed_form(expected):
    """K10 p=0.5: the logged tree (7433032 paths / 14866064 nodes) and the
    exact R, via the multithreaded depth-first order (order-independent)."""
    c = next(x for x in expected["logged_counts"] if x["graph"] == "COMP10-05.txt")
    r = load("COMP10-05.txt").solve(0, 9, order=ORDER_DFS, threads=8)
    assert (r["paths"], r["nodes"]) == (c["paths"], c["nodes"])
    R = closed_form(expected, 10, "0.5")
    assert abs(r["R"] - R) <= 1e-13 * R
    assert r["dead"] == r["nodes"] - r["paths"]


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_counts_independent_of_order_and_threads(threads):
    g = load("COMP8-03.txt")
    lv = g.solve(0, 7, order=ORDER_LEVEL)
    hp = g.solve(0, 7, order=ORDER_HEAP)
    df = g.solve(0, 7, order=ORDER_DFS, threads=threads)
    for r in (hp, df):
        assert (r["paths"], r["nodes"], r["dead"]) == (lv["paths"], lv["nodes"], lv["dead"])
        assert abs(r["R"] - lv["R"]) <= 1e-15


@pytest.mark.parametrize("eps", [1e-3, 1e-5, 1e-8])
def test_truncation_is_order_independent_and_brackets(expected, eps):
    g = load("COMP9.txt")
    Rt = closed_form(expected, 9, "0.5")
    rs = [g.solve(0, 8, order=o, trunc_eps=eps, threads=4)
          for o in (ORDER_LEVEL, ORDER_HEAP, ORDER_DFS)]
    for r in rs:
        assert (r["paths"], r["nodes"], r["truncated"]) == (
            rs[0]["paths"], rs[0]["nodes"], rs[0]["truncated"])
        assert r["R"] <= Rt + 1e-15 and Rt <= r["R"] + r["ESP"] + 1e-15
        assert abs(r["R"] + r["D"] + r["ESP"] - 1.0) <= 1e-14
    assert rs[0]["truncated"] > 0


@pytest.mark.parametrize("order", [ORDER_LEVEL, ORDER_HEAP])
@pytest.mark.parametrize("cap", [1, 7, 100, 1000])
def test_capacity_overflow_brackets(expected, order, cap):
    g = load("COMP9.txt")
    Rt = closed_form(expected, 9, "0.5")
    r = g.solve(0, 8, order=order, capacity=cap)
    assert r["R"] <= Rt + 1e-15 <= r["R"] + r["ESP"] + 2e-15
    assert abs(r["R"] + r["D"] + r["ESP"] - 1.0) <= 1e-14
    assert r["overflowed"] > 0


def test_heap_order_beats_level_order_under_capacity(expected):
    """Max-mass order resolves more mass than FIFO for the same capacity."""
    g = load("COMP10-05.txt")
    lv = g.solve(0, 9, order=ORDER_LEVEL, capacity=2000)
    hp = g.solve(0, 9, order=ORDER_HEAP, capacity=2000)
    assert hp["ESP"] < lv["ESP"]


@pytest.mark.parametrize("order", [ORDER_LEVEL, ORDER_HEAP, ORDER_DFS])
def test_time_budget_fold_keeps_bracket(expected, order):
    g = load("COMP15-05.txt")
    Rt = closed_form(expected, 15, "0.5")
    r = g.solve(0, 14, order=order, time_budget_s=0.2, threads=2)
    assert r["deadline_hit"] == 1
    assert r["R"] <= Rt <= r["R"] + r["ESP"] + 1e-14
    assert abs(r["R"] + r["D"] + r["ESP"] - 1.0) <= 1e-13


def test_src_equals_dst():
    r = load("test_kite.txt").solve(2, 2)
    assert r["R"] == 1.0 and r["paths"] == 1 and r["nodes"] == 1


def _csr(n, edges):
    rows = [[] for _ in range(n)]
    for i, (a, b) in enumerate(edges):
        rows[a].append((b, i))
        rows[b].append((a, i))
    rp, ci, ei = [0], [], []
    for u in range(n):
        for v, i in sorted(rows[u]):
            ci.append(v)
            ei.append(i)
        rp.append(len(ci))
    return rp, ci, ei


def test_series_graph_is_product():
    rnd = random.Random(1)
    n = 12
    edges = [(i, i + 1) for i in range(n - 1)]
    p = [rnd.uniform(0.5, 0.99) for _ in edges]
    g = OracleGraph.from_csr(*_csr(n, edges), p)
    r = g.solve(0, n - 1)
    assert abs(r["R"] - math.prod(p)) <= 1e-15
    assert r["nodes"] == n and r["paths"] == 1  # root + (n-1) dead children


def test_parallel_bundle():
    """Paths of length 2 through k middle nodes: 1 - prod(1 - p_a p_b)."""
    rnd = random.Random(2)
    k = 9
    edges = []
    for m in range(1, k + 1):
        edges += [(0, m), (m, k + 1)]
    p = [rnd.uniform(0.3, 0.9) for _ in edges]
    g = OracleGraph.from_csr(*_csr(k + 2, edges), p)
    r = g.solve(0, k + 1)
    want = 1.0 - math.prod(1 - p[2 * i] * p[2 * i + 1] for i in range(k))
    assert abs(r["R"] - want) <= 1e-14


def test_bridge_factors_on_bridge():
    """Two K4 blocks joined by one bridge edge: R = R1 * p_bridge * R2."""
    edges = [(a, b) for a in range(4) for b in range(a + 1, 4)]
    edges += [(a + 4, b + 4) for a in range(4) for b in range(a + 1, 4)]
    edges.append((3, 4))
    p = [0.5] * 12 + [0.8]
    g = OracleGraph.from_csr(*_csr(8, edges), p)
    r = g.solve(0, 7)
    # R(K4, 0->3) = 0.75 (closed form); inside the second block 4->7 also 0.75
    assert abs(r["R"] - 0.75 * 0.8 * 0.75) <= 1e-15


def test_pivot_mode_matches_closed_form(expected):
    """Edge-pivot factoring (workingPad/cpu_reference.cpp semantics)."""
    for name, n, p in [("COMP8.txt", 8, "0.5"), ("temp4x4.txt", 4, "0.5")]:
        r = load(name).solve(0, n - 1, mode=MODE_PIVOT, order=ORDER_DFS)
        R = closed_form(expected, n, p)
        assert abs(r["R"] - R) <= 1e-14


def test_pivot_mode_brute_force(expected):
    for bf in expected["brute_force"]:
        r = load(bf["graph"]).solve(bf["src"], bf["dst"], mode=MODE_PIVOT)
        assert abs(r["R"] - bf["R"]) <= 1e-15


def test_expand_children_partition_mass():
    """One expansion: children + success mass = item mass (telescoping sum),
    each child differs from the item on the path edges only."""
    g = load("COMP10-05.txt")
    W = g.mask_words
    rnd = random.Random(3)
    for _ in range(200):
        known = up = 0
        for e in range(45):
            s = rnd.random()
            if s < 0.3:
                known |= 1 << e  # down
            elif s < 0.5:
                known |= 1 << e
                up |= 1 << e
        m = rnd.uniform(0.1, 1.0)
        k, succ, ck, cu, cm, path = g.expand(0, 9, [known], [up], m)
        if k < 0:
            continue
        assert abs(sum(cm) + succ - m) <= 1e-15 * (k + 1)
        for i in range(k):
            assert int(ck[i][0]) & known == known and int(cu[i][0]) & up == up
        # failed hop of child i is the i-th unknown hop on the path
        unknown_hops = [int(e) for e in path if not (known >> int(e)) & 1]
        assert len(unknown_hops) == k
        for i, e in enumerate(unknown_hops):
            assert (int(ck[i][0]) >> e) & 1 and not (int(cu[i][0]) >> e) & 1
        assert W == 1


MODE_FORKBEST, MODE_MINUNK = 2, 3


@pytest.mark.parametrize("mode", [MODE_FORKBEST, MODE_MINUNK])
def test_alternative_rules_exact(expected, mode):
    """forkBest (main.cu:304-430) and the min-unknown path rule decompose
    exactly: R equals brute force and the closed form."""
    for bf in expected["brute_force"]:
        r = load(bf["graph"]).solve(bf["src"], bf["dst"], mode=mode)
        assert abs(r["R"] - bf["R"]) <= 1e-15, bf
        assert abs(r["R"] + r["D"] - 1.0) <= 1e-15
    for name, n, p in [("COMP8.txt", 8, "0.5"), ("COMP8-09.txt", 8, "0.9")]:
        r = load(name).solve(0, n - 1, mode=mode)
        assert abs(r["R"] - closed_form(expected, n, p)) <= 1e-14


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_forkbest_matches_logs(expected, idx):
    """forkBest reproduces the reference's logged runs exactly: paths, nodes,
    levels and widest level (workingPad/output_maxMem_forBest.txt).  With a
    uniform p every hop ties, so these counts pin the tie-break of the
    reference's warp argmax (lane r mod 32, shuffle tree); taking the first
    UNKNOWN hop from either end instead gives other counts (K8: 26830 paths,
    107319 nodes).  The tree is binary: nodes = 2 * leaves - 1."""
    lg = expected["logged_forkbest_counts"][idx]
    r = load(lg["graph"]).solve(lg["src"], lg["dst"], mode=MODE_FORKBEST, threads=4)
    assert (r["paths"], r["nodes"], r["iterations"], r["max_width"]) == (
        lg["paths"], lg["nodes"], lg["iterations"], lg["max_width"]), lg
    assert r["nodes"] == 2 * (r["paths"] + r["dead"]) - 1


def test_forkbest_keys_are_fp32_logs(expected):
    """Keys are logf of p in FP32 (graph.h:138): K8 with p = 0.5 plus
    perturbations below FP32 resolution ties exactly as the uniform K8 does
    and reproduces its logged tree; comparing the FP64 p would not."""
    from relaibilitygpu_amd import workloads as wl
    n, edges = 8, [(a, b) for a in range(8) for b in range(a + 1, 8)]
    probs = [0.5 + 1e-12 * (i % 7) for i in range(len(edges))]
    rp, ci, ei = wl.to_csr(n, edges)
    r = OracleGraph.from_csr(rp, ci, ei, probs).solve(0, 7, mode=MODE_FORKBEST)
    lg = expected["logged_forkbest_counts"][0]
    assert (r["paths"], r["nodes"], r["iterations"], r["max_width"]) == (
        lg["paths"], lg["nodes"], lg["iterations"], lg["max_width"])


def test_minunk_never_longer_in_unknowns():
    """On the kite graph the min-unknown rule needs fewer subproblems."""
    g = load("test_kite.txt")
    a = g.solve(0, 4)
    b = g.solve(0, 4, mode=MODE_MINUNK)
    assert b["nodes"] <= a["nodes"] and abs(a["R"] - b["R"]) <= 1e-15


# ------------------------------------------------------------ K-terminal
def _brute_k(n, edges, probs, K):
    """Probability that every terminal of K is connected (undirected), by
    enumerating all 2^E edge states with a union-find."""
    import itertools
    R = 0.0
    for st in itertools.product([0, 1], repeat=len(edges)):
        par = list(range(n))

        def f(x):
            while par[x] != x:
                par[x] = par[par[x]]
                x = par[x]
            return x
        for (a, b), s in zip(edges, st):
            if s:
                par[f(a)] = f(b)
        if len({f(k) for k in K}) == 1:
            R += float(np.prod([p if s else 1 - p for p, s in zip(probs, st)]))
    return R


def test_kterminal_pair_is_two_terminal():
    """K = {s, t} builds exactly the two-terminal tree (same counts, R)."""
    g = load("COMP10-05.txt")
    a = g.solve(0, 9)
    b = g.solve_terms([0, 9])
    assert (a["paths"], a["nodes"], a["iterations"], a["max_width"]) == (
        b["paths"], b["nodes"], b["iterations"], b["max_width"])
    assert a["R"] == b["R"]


@pytest.mark.parametrize("n", [4, 5, 6, 8])
def test_all_terminal_cycle_closed_form(n):
    """All-terminal reliability of the cycle C_n: p^n + n p^(n-1) q."""
    from relaibilitygpu_amd import workloads as wl
    p = 0.85
    edges = sorted((min(i, (i + 1) % n), max(i, (i + 1) % n)) for i in range(n))
    rp, ci, ei = wl.to_csr(n, edges)
    og = OracleGraph.from_csr(rp, ci, ei, [p] * n)
    r = og.solve_terms(list(range(n)))
    assert abs(r["R"] - (p ** n + n * p ** (n - 1) * (1 - p))) <= 1e-14
    assert abs(r["R"] + r["D"] - 1.0) <= 1e-14


@pytest.mark.parametrize("seed,K", [(3, [0, 3, 5]), (4, [6, 1, 2, 0]), (5, [2, 4]),
                                    (6, [0, 1, 2, 3, 4, 5, 6])])
def test_kterminal_brute_force(seed, K):
    """Random graphs on 7 nodes: path rule, forkBest rule, a permuted
    terminal list and depth-first order all give the brute-force R."""
    from relaibilitygpu_amd import workloads as wl
    rng = np.random.default_rng(seed)
    n = 7
    edges = sorted({(int(min(x, y)), int(max(x, y))) for x, y in rng.integers(0, n, (14, 2))
                    if x != y})
    probs = list(rng.uniform(0.3, 0.9, len(edges)))
    rp, ci, ei = wl.to_csr(n, edges)
    og = OracleGraph.from_csr(rp, ci, ei, probs)
    R = _brute_k(n, edges, probs, K)
    for r in (og.solve_terms(K), og.solve_terms(K[::-1]), og.solve_terms(K, mode=MODE_FORKBEST),
              og.solve_terms(K, order=ORDER_DFS, threads=2)):
        assert abs(r["R"] - R) <= 1e-13
        assert abs(r["R"] + r["D"] - 1.0) <= 1e-13


def test_kterminal_detects_a_wrong_tree():
    """A plausible slip, branching only on the path to the last terminal,
    changes R: on C4 with K = all nodes it would give two-terminal R."""
    from relaibilitygpu_amd import workloads as wl
    p = 0.85
    edges = [(0, 1), (0, 3), (1, 2), (2, 3)]
    rp, ci, ei = wl.to_csr(4, edges)
    og = OracleGraph.from_csr(rp, ci, ei, [p] * 4)
    assert abs(og.solve_terms([0, 1, 2, 3])["R"] - og.solve(0, 3)["R"]) > 1e-3


# ------------------------------------------------- alternating-sweep rule
from oracle_ffi import MODE_ALTERNATE  # noqa: E402


def _states(rnd, e, pf=0.3, pw=0.2):
    known = up = 0
    for i in range(e):
        s = rnd.random()
        if s < pf:
            known |= 1 << i
        elif s < pf + pw:
            known |= 1 << i
            up |= 1 << i
    return known, up


def _zero_one_distance(n, ends, src, dst, known, up):
    """Fewest UNKNOWN edges on a src->dst walk that avoids FAILED edges
    (WORKING edges cost 0), by 0-1 BFS; None when dst is unreachable."""
    import collections
    adj = [[] for _ in range(n)]
    for i, (a, b) in enumerate(ends):
        if (known >> i) & 1 and not (up >> i) & 1:
            continue
        w = 0 if (known >> i) & 1 else 1
        adj[a].append((b, w))
        adj[b].append((a, w))
    dist = [None] * n
    dist[src] = 0
    dq = collections.deque([src])
    while dq:
        u = dq.popleft()
        for v, w in adj[u]:
            if dist[v] is None or dist[u] + w < dist[v]:
                dist[v] = dist[u] + w
                (dq.appendleft if w == 0 else dq.append)(v)
    return dist[dst]


def _components(n, ends, up):
    lab = list(range(n))

    def find(x):
        while lab[x] != x:
            x = lab[x]
        return x
    for i, (a, b) in enumerate(ends):
        if (up >> i) & 1:
            lab[find(a)] = find(b)
    return [find(v) for v in range(n)]


@pytest.mark.parametrize("period", [1, 2, 3])
def test_alternate_rule_exact(expected, period):
    """Any path and any hop order split an item exactly: R equals brute
    force and the closed form, R + D = 1, for several sweep periods."""
    for bf in expected["brute_force"]:
        r = load(bf["graph"]).solve(bf["src"], bf["dst"], mode=MODE_ALTERNATE,
                                    sweep_period=period)
        assert abs(r["R"] - bf["R"]) <= 1e-15, bf
        assert abs(r["R"] + r["D"] - 1.0) <= 1e-15
    for name, n, p in [("COMP8.txt", 8, "0.5"), ("COMP8-09.txt", 8, "0.9"),
                       ("COMP8-03.txt", 8, "0.3")]:
        r = load(name).solve(0, n - 1, mode=MODE_ALTERNATE, sweep_period=period)
        assert abs(r["R"] - closed_form(expected, n, p)) <= 1e-14


def test_alternate_k10_closed_form_and_fewer_nodes(expected):
    g = load("COMP10-05.txt")
    a = g.solve(0, 9, mode=MODE_ALTERNATE, order=ORDER_DFS, threads=8)
    assert abs(a["R"] - closed_form(expected, 10, "0.5")) <= 1e-13
    assert a["ESP"] == 0 and abs(a["R"] + a["D"] - 1.0) <= 1e-14
    # fewer than half the plain rule's 14 866 064 subproblems
    assert a["nodes"] < 14866064 // 2


def test_alternate_counts_independent_of_order_and_threads():
    g = load("COMP8.txt")
    ref = g.solve(0, 7, mode=MODE_ALTERNATE)
    for order, threads in [(ORDER_HEAP, 1), (ORDER_DFS, 1), (ORDER_DFS, 4)]:
        r = g.solve(0, 7, mode=MODE_ALTERNATE, order=order, threads=threads)
        assert (r["paths"], r["nodes"], r["dead"]) == (ref["paths"], ref["nodes"], ref["dead"])
        assert abs(r["R"] - ref["R"]) <= 1e-15
    # leaves (successes + dead) and internal nodes: every live item adds
    # exactly one success, so paths = live items and nodes = paths + dead
    assert ref["nodes"] == ref["paths"] + ref["dead"]


@pytest.mark.parametrize("eps", [1e-6, 1e-9])
def test_alternate_brackets_under_truncation(expected, eps):
    g = load("COMP15-05.txt")
    r = g.solve(0, 14, mode=MODE_ALTERNATE, order=ORDER_DFS, threads=8, trunc_eps=eps)
    R = closed_form(expected, 15, "0.5")
    assert r["R"] <= R <= r["R"] + r["ESP"] * (1 + 1e-12)
    assert abs(r["R"] + r["D"] + r["ESP"] - 1.0) <= 1e-13


@pytest.mark.parametrize("name,src,dst,period", [("COMP10-05.txt", 0, 9, 1),
                                                  ("COMP10-05.txt", 0, 9, 2),
                                                  ("test_kite.txt", 0, 4, 1),
                                                  ("temp4x4.txt", 0, 3, 3)])
def test_alternate_single_expansion(name, src, dst, period):
    """One expansion on random edge states: the path has the fewest UNKNOWN
    hops (0-1 BFS distance, WORKING edges free), its hops chain src's
    WORKING component to dst's, the sweep starts at the src end exactly
    when (FAILED count / period) is even, and child i has hops 0..i-1
    WORKING, hop i FAILED and mass m * p_0 ... p_{i-1} * q_i."""
    g = load(name)
    rp, ci, ei, p = g.arrays()
    n, E = len(rp) - 1, len(p)
    ends = [None] * E
    for u in range(n):
        for j in range(rp[u], rp[u + 1]):
            ends[ei[j]] = (min(u, ci[j]), max(u, ci[j]))
    rnd = random.Random(11)
    seen_dirs = set()
    for _ in range(400):
        known, up = _states(rnd, E)
        m = rnd.uniform(0.1, 1.0)
        kw = [(known >> (64 * w)) & ((1 << 64) - 1) for w in range(g.mask_words)]
        uw = [(up >> (64 * w)) & ((1 << 64) - 1) for w in range(g.mask_words)]
        k, succ, ck, cu, cm, hops = g.expand_rule(src, dst, kw, uw, m, MODE_ALTERNATE, period)
        d = _zero_one_distance(n, ends, src, dst, known, up)
        if d is None:
            assert k == -1
            continue
        assert k == d
        lab = _components(n, ends, up & known)
        hops = [int(h) for h in hops]
        # orientation: the sweep starts at src's component or at dst's
        f = bin(known & ~up).count("1")
        from_src = (f // period) % 2 == 0
        chain = hops if from_src else hops[::-1]
        reached = {lab[src]}
        for e in chain:
            a, b = ends[e]
            assert not (known >> e) & 1  # every hop is UNKNOWN
            assert (lab[a] in reached) != (lab[b] in reached)
            reached |= {lab[a], lab[b]}
        assert lab[dst] in reached
        if k:
            seen_dirs.add(from_src)
        run = m
        pre = 0
        for i, e in enumerate(hops):
            ckv = sum(int(x) << (64 * w) for w, x in enumerate(ck[i]))
            cuv = sum(int(x) << (64 * w) for w, x in enumerate(cu[i]))
            assert ckv == known | pre | (1 << e) and cuv == up | pre
            assert cm[i] == run * (1.0 - p[e])
            run = run * p[e]
            pre |= 1 << e
        assert succ == run
        assert abs(sum(cm) + succ - m) <= 1e-15 * (k + 1)
    assert seen_dirs == {True, False}


def test_alternate_detects_a_longer_path():
    """Planted mistake: the plain rule's path (hop-shortest, WORKING edges
    counted) is not the fewest-UNKNOWN path, and the 0-1 distance check
    above tells them apart on the kite graph."""
    g = load("test_kite.txt")
    rp, ci, ei, p = g.arrays()
    n, E = len(rp) - 1, len(p)
    ends = [None] * E
    for u in range(n):
        for j in range(rp[u], rp[u + 1]):
            ends[ei[j]] = (min(u, ci[j]), max(u, ci[j]))
    rnd = random.Random(5)
    differs = 0
    for _ in range(300):
        known, up = _states(rnd, E, 0.1, 0.5)
        d = _zero_one_distance(n, ends, 0, 4, known, up)
        k, *_ = g.expand(0, 4, [known], [up], 1.0)
        if d is not None and k >= 0 and k != d:
            differs += 1
    assert differs > 0


# ------------------------------------------- graphs past the fixed sizes
def test_big_graph_oracle_matches_the_fixed_size_engine(expected):
    """K8 embedded in a 300-node graph (292 isolated extra nodes, so the
    any-size restatement runs): the same counts, levels, widest level and R
    as the fixed-size engine on K8, whose counts the reference logs pin."""
    small = load("COMP8.txt")
    rp, ci, ei, p = small.arrays()
    n = 300
    rp_big = list(rp) + [rp[-1]] * (n - 8)
    g = OracleGraph.from_csr(rp_big, ci, ei, p)
    assert g.dims()["n"] == n
    a = g.solve(0, 7)
    b = small.solve(0, 7)
    for k in ("paths", "nodes", "dead", "truncated", "iterations", "max_width"):
        assert a[k] == b[k], k
    assert abs(a["R"] - closed_form(expected, 8, "0.5")) <= 1e-14
    # a renumbering that puts the terminals far apart in index order
    perm = list(range(n))
    perm[0], perm[250] = 250, 0
    perm[7], perm[299] = 299, 7
    edges = []
    for u in range(8):
        for j in range(rp[u], rp[u + 1]):
            v = ci[j]
            if u < v:
                edges.append((perm[u], perm[v], p[ei[j]]))
    rows = [[] for _ in range(n)]
    for i, (x, y, _) in enumerate(edges):
        rows[x].append((y, i))
        rows[y].append((x, i))
    rp2, ci2, ei2 = [0], [], []
    for u in range(n):
        for v, i in sorted(rows[u]):
            ci2.append(v)
            ei2.append(i)
        rp2.append(len(ci2))
    g2 = OracleGraph.from_csr(rp2, ci2, ei2, [e[2] for e in edges])
    c = g2.solve(250, 299)
    assert abs(c["R"] - b["R"]) <= 1e-14 and c["paths"] + c["dead"] == c["nodes"]


def test_big_graph_oracle_series_and_bundle():
    """Closed forms on 300-node graphs: a series path (R = prod p, one
    success, n nodes) and a parallel bundle of ten 2-hop paths with isolated
    padding (R = 1 - prod(1 - p_a p_b))."""
    rnd = random.Random(9)
    n = 300
    edges = [(i, i + 1) for i in range(n - 1)]
    p = [rnd.uniform(0.97, 0.999) for _ in edges]
    g = OracleGraph.from_csr(*_csr(n, edges), p)
    r = g.solve(0, n - 1)
    assert abs(r["R"] - math.prod(p)) <= 1e-14 and r["paths"] == 1 and r["nodes"] == n
    k = 10  # ten 2-hop paths; 288 isolated nodes make the graph big
    edges = []
    for m in range(1, k + 1):
        edges += [(0, m), (m, k + 1)]
    p = [rnd.uniform(0.3, 0.6) for _ in edges]
    g = OracleGraph.from_csr(*_csr(n, edges), p)
    r = g.solve(0, k + 1)
    want = 1.0 - math.prod(1 - p[2 * i] * p[2 * i + 1] for i in range(k))
    assert abs(r["R"] - want) <= 1e-14 and r["ESP"] == 0


# ------------------------------------ edge pivot against the reference program
def test_pivot_mode_matches_the_reference_program():
    """tests/golden/pivot_ref.json holds the reference's own edge-pivot program
    (workingPad/cpu_reference.cpp, built from its source into oracle/_ref)
    run on these graphs: the oracle's pivot mode must give the same node and
    path counts; R agrees to the float32 parsing of p that program uses
    (exactly where p is a float, e.g. 0.5)."""
    import json
    import os
    vecs = json.load(open(os.path.join(os.path.dirname(graph_path("COMP8.txt")), "..",
                                       "pivot_ref.json")))
    assert len(vecs) >= 9
    for v in vecs:
        r = load(v["graph"]).solve(v["src"], v["dst"], mode=MODE_PIVOT, order=ORDER_DFS,
                                   threads=4)
        assert (r["paths"], r["nodes"]) == (v["paths"], v["nodes"]), v
        tol = 1e-12 if v["graph"] in ("COMP8.txt", "cycle4.txt", "temp4x4.txt") else 1e-7
        assert abs(r["R"] - v["R_float_p"]) <= tol * max(r["R"], 1e-3) + 5e-13, v


# ------------------------------------------------- learned sweep choice
from oracle_ffi import MODE_EXPLORE, MODE_LEARNED  # noqa: E402


def _random_chooser(seed):
    rng = np.random.default_rng(seed)
    return (rng.integers(-255, 256, size=(16, 64)).astype(np.int32),
            rng.integers(0, 2, size=64).astype(np.int32))


@pytest.mark.parametrize("mode", [MODE_LEARNED, MODE_EXPLORE])
def test_learned_and_exploring_sweeps_exact(expected, mode):
    """Either sweep direction splits an item exactly, so whatever the
    chooser (random integer weights here) or the exploring hash picks, R
    equals brute force and the closed forms and R + D = 1."""
    ch = _random_chooser(5)
    for bf in expected["brute_force"]:
        r = load(bf["graph"]).solve(bf["src"], bf["dst"], mode=mode, chooser=ch)
        assert abs(r["R"] - bf["R"]) <= 1e-15, bf
        assert abs(r["R"] + r["D"] - 1.0) <= 1e-15
    for name, n, p in [("COMP8.txt", 8, "0.5"), ("COMP8-09.txt", 8, "0.9"),
                       ("COMP8-03.txt", 8, "0.3")]:
        r = load(name).solve(0, n - 1, mode=mode, chooser=ch)
        assert abs(r["R"] - closed_form(expected, n, p)) <= 1e-14


def test_zero_chooser_is_the_alternating_rule():
    """Equal scores fall back to the alternating direction: an all-zero
    chooser builds the alternating rule's tree exactly."""
    zero = (np.zeros((16, 64), np.int32), np.zeros(64, np.int32))
    for name, n, period, eps in [("COMP8.txt", 8, 2, 0.0), ("COMP9.txt", 9, 3, 0.0),
                                 ("COMP15-05.txt", 15, 4, 1e-7)]:
        g = load(name)
        a = g.solve(0, n - 1, mode=MODE_ALTERNATE, sweep_period=period, trunc_eps=eps)
        b = g.solve(0, n - 1, mode=MODE_LEARNED, sweep_period=period, trunc_eps=eps,
                    chooser=zero)
        assert (a["paths"], a["nodes"], a["dead"], a["truncated"], a["iterations"],
                a["max_width"]) == (b["paths"], b["nodes"], b["dead"], b["truncated"],
                                    b["iterations"], b["max_width"])


def test_learned_counts_independent_of_order_and_threads():
    g = load("COMP9.txt")
    ch = _random_chooser(9)
    ref = g.solve(0, 8, mode=MODE_LEARNED, chooser=ch)
    for order, threads in [(ORDER_HEAP, 1), (ORDER_DFS, 1), (ORDER_DFS, 4)]:
        r = g.solve(0, 8, mode=MODE_LEARNED, order=order, threads=threads, chooser=ch)
        assert (r["paths"], r["nodes"], r["dead"]) == (ref["paths"], ref["nodes"], ref["dead"])
        assert abs(r["R"] - ref["R"]) <= 1e-15


def _lq(x):
    """round(-4 log2 x) in 0..63 (63 for x <= 0): the choice's probability step."""
    if not x > 0:
        return 63
    return int(min(63.0, max(0.0, np.rint(-4.0 * math.log2(x)))))


def _floor_log2(m):
    return math.frexp(m)[1] - 1


def _fdiv2(v):
    return v // 2  # Python floors


def _choice_x(n, ends, p, src, dst, known, up, mass, eps, hops_fwd, from_dst):
    """The 16 choice features, restated from their definition."""
    E = len(ends)
    lab = _components(n, ends, up & known)
    S = {v for v in range(n) if lab[v] == lab[src]}
    T = {v for v in range(n) if lab[v] == lab[dst]}

    def leaving(C):
        cut = failed = 0
        for i, (a, b) in enumerate(ends):
            if (a in C) == (b in C):
                continue
            if (known >> i) & 1 and not (up >> i) & 1:
                failed += 1
            else:
                cut += 1
        return len(C), cut, failed
    A, B = (T, S) if from_dst else (S, T)
    nA, cA, fA = leaving(A)
    nB, cB, fB = leaving(B)
    f = bin(known & ~up).count("1")
    u = E - bin(known).count("1")
    lim = lambda v: max(-63, min(63, v))  # noqa: E731
    lm = _floor_log2(mass)
    le = _floor_log2(eps) if eps > 2.0 ** -64 else -64
    head = max(0, min(32, _fdiv2(lm - le)))
    first, last = (hops_fwd[-1], hops_fwd[0]) if from_dst else (hops_fwd[0], hops_fwd[-1])
    lp = sum(_lq(p[e]) for e in hops_fwd)
    return [lim(len(hops_fwd)), lim(nA), lim(cA), lim(fA), lim(nB), lim(cB), lim(fB),
            (32 * f) // E, (32 * u) // E, 0, head, -lim(lp), -_lq(1.0 - p[first]),
            -_lq(1.0 - p[last]), lim(cA - cB), 1]


def _explore_bit(words_known, words_up):
    M = (1 << 64) - 1

    def fin(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)
    h = 0x9E3779B97F4A7C15
    for k, u in zip(words_known, words_up):
        h = fin(h ^ k)
        h = fin(h ^ u)
    return (h >> 62) == 0


@pytest.mark.parametrize("name,src,dst,eps,pf,pw", [("COMP10-05.txt", 0, 9, 0.0, 0.25, 0.15),
                                                    ("COMP10-05.txt", 2, 7, 1e-9, 0.4, 0.05),
                                                    ("test_kite.txt", 0, 4, 0.0, 0.25, 0.15),
                                                    ("COMP25-05.txt", 0, 24, 1e-12, 0.6, 0.003)])
def test_learned_single_expansion_choice(name, src, dst, eps, pf, pw):
    """One expansion on random edge states under a random integer chooser:
    the path is the alternating rule's, swept from the end whose features
    (restated here from their definition) score lower, the alternating
    direction on ties; the exploring mode flips the alternating direction
    where the hash's top two bits are zero."""
    g = load(name)
    rp, ci, ei, p = g.arrays()
    n, E = len(rp) - 1, len(p)
    ends = [None] * E
    for u in range(n):
        for j in range(rp[u], rp[u + 1]):
            ends[ei[j]] = (min(u, ci[j]), max(u, ci[j]))
    w, sg = _random_chooser(17)
    margin = 5000 if name == "test_kite.txt" else 0
    rnd = random.Random(23)
    dirs = set()
    for _ in range(300):
        known, up = _states(rnd, E, pf, pw)
        m = rnd.choice([1.0, 0.5, 3e-7, 2.5e-11, 7e-13, 1e-19, 3e-22])
        kw = [(known >> (64 * i)) & ((1 << 64) - 1) for i in range(g.mask_words)]
        uw = [(up >> (64 * i)) & ((1 << 64) - 1) for i in range(g.mask_words)]
        k0, fwd = g.expand_choice(src, dst, kw, uw, m, MODE_ALTERNATE, period=1000)
        k1, got = g.expand_choice(src, dst, kw, uw, m, MODE_LEARNED, period=3,
                                  chooser=(w, sg, margin), eps=eps)
        k2, exp_hops = g.expand_choice(src, dst, kw, uw, m, MODE_EXPLORE, period=3)
        assert k0 == k1 == k2
        if k0 < 2:
            continue
        score = []
        for from_dst in (False, True):
            x = np.array(_choice_x(n, ends, p, src, dst, known, up, m, eps, fwd, from_dst),
                         np.int64)
            h = x @ w.astype(np.int64)
            score.append(int(np.where(sg == 1, -h, h)[h > 0].sum()))
        f = bin(known & ~up).count("1")
        alt = (f // 3) % 2
        rev = bool(alt ^ 1) if score[alt] - score[alt ^ 1] > margin else bool(alt)
        assert got == (fwd[::-1] if rev else fwd)
        dirs.add(rev)
        alt_rev = (f // 3) % 2 == 1
        assert exp_hops == (fwd[::-1] if alt_rev != _explore_bit(kw, uw) else fwd)
    assert dirs == {True, False}


# ------------------------------------------------- static prefix partition
@pytest.mark.parametrize("t", [1, 3, 6])
@pytest.mark.parametrize("order", [ORDER_LEVEL, ORDER_HEAP, ORDER_DFS])
def test_prefix_partition_exact(expected, t, order):
    """The 2^t seeds over the root's first t UNKNOWN edges partition the
    states (original/Rel_16M.cu:97-101), so R is the closed form and
    R + D = 1 whatever t and order."""
    r = load("COMP8-09.txt").solve(0, 7, order=order, prefix_edges=t, threads=2)
    assert abs(r["R"] - closed_form(expected, 8, "0.9")) <= 1e-14
    assert abs(r["R"] + r["D"] - 1.0) <= 1e-14
    assert r["nodes"] >= 2 ** t


def test_prefix_over_every_edge_is_brute_force(expected):
    """Fixing every edge makes each seed a full state: it is a success (a
    WORKING path) or dead, so the solve is the brute-force enumeration."""
    for bf in expected["brute_force"]:
        g = load(bf["graph"])
        E = g.dims()["e"]
        if E > 16:
            continue
        r = g.solve(bf["src"], bf["dst"], prefix_edges=E)
        assert r["nodes"] == 2 ** E and r["paths"] + r["dead"] == 2 ** E
        assert r["iterations"] == 1
        assert abs(r["R"] - bf["R"]) <= 1e-15, bf


def test_prefix_shards_add_up():
    g = load("COMP9.txt")
    whole = g.solve(0, 8, prefix_edges=5)
    parts = [g.solve(0, 8, prefix_edges=5, shard=s, shards=3) for s in range(3)]
    for f in ("paths", "nodes", "dead", "truncated"):
        assert sum(p[f] for p in parts) == whole[f], f
    assert abs(sum(p["R"] for p in parts) - whole["R"]) <= 1e-14
    assert abs(sum(p["D"] for p in parts) - whole["D"]) <= 1e-14


@pytest.mark.parametrize("name,s,t,eps", [("test_kite.txt", 0, 4, 0.0), ("COMP8-03.txt", 0, 7, 1e-7)])
def test_child_feature_restatement_identities(name, s, t, eps):
    """The learned order's child features restated on the CPU
    (oracle/features_oracle.py) obey what their definition fixes."""
    from features_oracle import child_feature_rows

    og = OracleGraph.load(graph_path(name))
    X, Y, nodes = child_feature_rows(og, s, t, eps)
    assert len(X) == nodes - 1  # every node but the root is some item's child
    # each edge is UNKNOWN, WORKING or FAILED in the child
    assert np.abs(X[:, 1] + X[:, 2] + X[:, 3] - 1.0).max() < 1e-6
    # a child's mass is its parent's times a probability: f0 <= f13
    assert np.all(X[:, 0] <= X[:, 13] + 1e-6)
    # the last child of an item: f4 = 1 and f14 = 0, and only then
    last = np.abs(X[:, 4] - 1.0) < 1e-6
    assert np.array_equal(last, np.abs(X[:, 14]) < 1e-9)
    assert np.all((X[:, 12] == 0.0) | (X[:, 12] == 1.0)) and np.all(X[:, 15] == 1.0)
    assert np.all((X[:, 9] >= 0.0) & (X[:, 9] <= 1.0))
    # labels are log2 of integer subtree sizes
    sizes = np.round(np.exp2(Y)).astype(np.int64)
    assert np.allclose(np.exp2(Y), sizes, rtol=1e-12) and sizes.min() >= 1


def _brute_conditional(rp, ci, ei, p, src, dst, known, up):
    """P(src reaches dst | the WORKING and FAILED edges of (known, up)):
    every assignment of the UNKNOWN edges, vectorised over assignments,
    reachability by relaxing every edge n times."""
    n, E = len(rp) - 1, len(p)
    bit = lambda planes, e: (int(planes[e >> 6]) >> (e & 63)) & 1  # noqa: E731
    ends = {}
    for u in range(n):
        for j in range(rp[u], rp[u + 1]):
            ends[int(ei[j])] = (u, int(ci[j]))
    unknown = [e for e in range(E) if not bit(known, e)]
    A = 1 << len(unknown)
    on = np.zeros((A, E), bool)
    prob = np.ones(A)
    assign = (np.arange(A)[:, None] >> np.arange(len(unknown))[None, :]) & 1
    for e in range(E):
        if bit(known, e):
            on[:, e] = bool(bit(up, e))
    for i, e in enumerate(unknown):
        on[:, e] = assign[:, i] == 1
        prob *= np.where(on[:, e], p[e], 1.0 - p[e])
    reach = np.zeros((A, n), bool)
    reach[:, src] = True
    for _ in range(n):
        for e, (a, b) in ends.items():
            w = on[:, e]
            ra, rb = reach[:, a].copy(), reach[:, b].copy()
            reach[:, b] |= ra & w
            reach[:, a] |= rb & w
    return float(prob[reach[:, dst]].sum())


@pytest.mark.parametrize("name,s,t,max_unknown,limit", [
    ("test_kite.txt", 0, 4, 64, None), ("cycle4.txt", 0, 2, 64, None),
    ("test_bridge.txt", 0, 3, 64, None), ("COMP8-09.txt", 0, 7, 11, 400)])
def test_child_conditional_reliability_brute_force(name, s, t, max_unknown, limit):
    """The tree log's second label (a child's subtree success mass over its
    own mass) equals brute force over the child's UNKNOWN edges."""
    from features_oracle import child_rows_with_reliability

    og = load(name)
    rp, ci, ei, p = og.arrays()
    X, Y, C, planes, nodes = child_rows_with_reliability(og, s, t, 0.0)
    assert len(C) == nodes - 1 and np.all((C >= 0.0) & (C <= 1.0 + 1e-15))
    E = len(p)
    picked = [i for i in range(len(C))
              if E - sum(bin(int(w)).count("1") for w in planes[i, 0]) <= max_unknown]
    if limit:
        picked = picked[:: max(1, len(picked) // limit)]
    assert len(picked) >= min(len(C), 20)
    for i in picked:
        bf = _brute_conditional(rp, ci, ei, p, s, t, planes[i, 0], planes[i, 1])
        assert abs(C[i] - bf) <= 1e-12, (i, C[i], bf)
    # the root: its own success mass plus its children's conditional
    # reliabilities weighted by their masses is R
    z = np.zeros(og.mask_words, np.uint64)
    k, succ, ck, cu, cm, _ = og.expand(s, t, z, z, 1.0)
    first = [i for i in range(len(C)) if np.array_equal(planes[i, 0], ck[0])][0]
    total = succ + sum(C[first + j] * cm[j] for j in range(k))
    assert np.array_equal(planes[first:first + k, 0], ck)
    assert abs(total - og.solve(s, t)["R"]) <= 1e-13
