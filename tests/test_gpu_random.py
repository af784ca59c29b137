"""GPU: seeded random graphs against the CPU oracle.

Erdos-Renyi graphs (8-40 nodes, mean degree 2.5-6), grids and dense graphs,
with random edge probabilities, truncation, capacity, rule (path or
alternating, several sweep periods) and order (level or max-mass).  Bar:
without overflow or folding, paths / nodes / dead / truncated equal the
oracle's and R, ESP within 1e-11 relative (the oracle runs depth-first: the
same tree in another order); with a capacity that binds, the brackets
intersect; R + D + ESP = 1 always.
"""
import random

import pytest

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from oracle_ffi import MODE_ALTERNATE, MODE_PATH, ORDER_DFS, OracleGraph  # noqa: E402
from relaibilitygpu_amd import workloads as wl  # noqa: E402


def close(a, b, tol):
    return abs(a - b) <= tol * max(abs(a), abs(b), 1e-300) + 1e-300


def random_case(rng):
    kind = rng.choice(["er", "er", "grid", "dense"])
    if kind == "er":
        n = rng.randint(8, 40)
        d = rng.uniform(2.5, 6.0)
        edges = {(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < d / (n - 1)}
        edges |= {(u, u + 1) for u in range(n - 1) if rng.random() < 0.5}
        edges = sorted(edges)
    elif kind == "grid":
        r, c = rng.randint(3, 7), rng.randint(3, 8)
        n = r * c
        edges = [(i * c + j, i * c + j + 1) for i in range(r) for j in range(c - 1)]
        edges += [(i * c + j, (i + 1) * c + j) for i in range(r - 1) for j in range(c)]
    else:
        n = rng.randint(6, 11)
        edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < 0.8]
    probs = [rng.choice([rng.uniform(0.3, 0.99), 0.5, 0.9]) for _ in edges]
    eps = rng.choice([0.0, 1e-5, 1e-7, 1e-9]) if n <= 12 else rng.choice([1e-5, 1e-7, 1e-8])
    rule = rng.choice(["path", "path", "alternate"])
    kw = dict(truncation_eps=eps, queue_capacity=rng.choice([1 << 12, 1 << 16, 1 << 20, 1 << 23]),
              order=rng.choice(["level", "mass"]), rule=rule)
    if rule == "alternate":
        kw["sweep_period"] = rng.choice([1, 2, 3, 6])
    return kind, n, edges, probs, kw


@pytest.mark.parametrize("seed", [11, 12, 13, 14])
def test_random_graphs_against_oracle(seed):
    rng = random.Random(seed)
    done = 0
    while done < 60:
        kind, n, edges, probs, kw = random_case(rng)
        if not edges:
            continue
        rp, ci, ei = wl.to_csr(n, edges)
        mode = MODE_ALTERNATE if kw["rule"] == "alternate" else MODE_PATH
        o = OracleGraph.from_csr(rp, ci, ei, probs).solve(
            0, n - 1, mode=mode, order=ORDER_DFS, threads=8, trunc_eps=kw["truncation_eps"],
            sweep_period=kw.get("sweep_period", 0))
        if o["nodes"] > 3e6:
            continue
        print(f"case seed {seed} #{done} {kind} n={n} e={len(edges)} {kw}", flush=True)
        r = nr.solve(nr.Graph.from_csr(rp, ci, ei, probs, 0, n - 1), **kw)
        if r["overflowed"] == 0 and r["folded"] == 0:
            for key in ("paths", "nodes", "dead", "truncated"):
                assert r[key] == o[key], (key, r[key], o[key])
            assert close(r["R"], o["R"], 1e-11) and close(r["ESP"], o["ESP"], 1e-10)
        else:
            assert r["R"] <= o["R"] + o["ESP"] + 1e-12 and o["R"] <= r["R"] + r["ESP"] + 1e-12
        assert abs(r["R"] + r["D"] + r["ESP"] - 1.0) <= 1e-12
        done += 1


@pytest.mark.parametrize("seed", [21, 22])
def test_random_csr_digraphs_and_choosers(seed):
    """Directed graphs (10-120 nodes), undirected graphs of 65-250 nodes (the
    CSR form) and the learned path choice with random chooser weights, in
    level and max-mass order, against the oracle (its learned mode takes the
    same integer chooser)."""
    import numpy as np

    from oracle_ffi import MODE_LEARNED

    rng = random.Random(seed)
    done = 0
    while done < 50:
        kind = rng.choice(["digraph", "big_und", "choice_small", "choice_csr"])
        eps = rng.choice([1e-5, 1e-6, 1e-7])
        kw = dict(truncation_eps=eps, queue_capacity=rng.choice([1 << 18, 1 << 22]),
                  order=rng.choice(["level", "mass"]))
        model = None
        if kind == "digraph":
            n, arcs, probs = wl.random_digraph(rng.randint(10, 120), rng.uniform(2.0, 4.0), 0.6,
                                               0.97, rng.randint(0, 10**6))
            rp, ci, ei = wl.to_csr_directed(n, arcs)
            if len(arcs) > 512:
                continue
        else:
            n = rng.randint(65, 250) if kind != "choice_small" else rng.randint(8, 40)
            d = rng.uniform(2.2, 4.0)
            edges = {(u, v) for u in range(n) for v in range(u + 1, n) if rng.random() < d / (n - 1)}
            edges |= {(u, u + 1) for u in range(n - 1) if rng.random() < 0.7}
            edges = sorted(edges)
            if len(edges) > 512:
                continue
            probs = [rng.uniform(0.7, 0.99) for _ in edges]
            rp, ci, ei = wl.to_csr(n, edges)
            if kind.startswith("choice"):
                p = np.random.default_rng(rng.randint(0, 10**6)).normal(
                    0.0, 0.3, nr.MLP_PARAMS).astype(np.float32)
                model = nr.Model(None, chooser=p, margin=rng.choice([0.0, 0.5]))
                kw.update(rule="learned", sweep_period=rng.choice([1, 2, 3]))
        og = OracleGraph.from_csr(rp, ci, ei, probs)
        if model is not None:
            o = og.solve(0, n - 1, mode=MODE_LEARNED, order=ORDER_DFS, threads=8, trunc_eps=eps,
                         sweep_period=kw["sweep_period"], chooser=model.chooser_weights(),
                         time_budget_s=5.0)
        else:
            o = og.solve(0, n - 1, order=ORDER_DFS, threads=8, trunc_eps=eps, time_budget_s=5.0)
        if o["deadline_hit"] or o["nodes"] > 3e6:
            continue
        print(f"case seed {seed} #{done} {kind} n={n} {kw}", flush=True)
        g = nr.Graph.from_csr(rp, ci, ei, probs, 0, n - 1)
        r = nr.solve_learned(g, model, **kw) if model is not None else nr.solve(g, **kw)
        if r["overflowed"] == 0 and r["folded"] == 0:
            for key in ("paths", "nodes", "dead", "truncated"):
                assert r[key] == o[key], (key, r[key], o[key])
            assert close(r["R"], o["R"], 1e-11) and close(r["ESP"], o["ESP"], 1e-10)
        else:
            assert r["R"] <= o["R"] + o["ESP"] + 1e-12 and o["R"] <= r["R"] + r["ESP"] + 1e-12
        assert abs(r["R"] + r["D"] + r["ESP"] - 1.0) <= 1e-12
        done += 1
