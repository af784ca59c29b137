"""The learned configuration on graphs it was not trained on: a 10x10 grid
(180 edges, p ~ U[0.85, 0.95] seeded, corner to corner) and Barabasi-Albert
n = 150, m = 2, p = 0.95 (0 -> 149).  The order scorer is trained on the GPU
from children logged on K10 and the path chooser from choice logs of
complete graphs (relaibilitygpu_amd/learn.py, the benchmark's training).

* learned order alone (alternating rule): the same tree as the oracle's
  alternating rule (order never changes counts when nothing binds);
* learned order with the learned path choice: the oracle's learned mode
  with the model's exact chooser;
* a binding 0.2 s budget: the bracket R <= R_true <= R + ESP intersects the
  oracle's bracket from a truncated run.
"""
import pytest

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from relaibilitygpu_amd import learn  # noqa: E402
from oracle_ffi import MODE_ALTERNATE, MODE_LEARNED, ORDER_DFS, ORDER_LEVEL, OracleGraph  # noqa: E402


def close(a, b, rel=1e-12):
    return abs(a - b) <= rel * max(abs(a), abs(b)) + 1e-300


def counts(r):
    return (r["paths"], r["nodes"], r["dead"], r["truncated"])


GRAPHS = {"grid10x10": (lambda: learn.grid(10, 10, 0x6121D), (1e-6, 1e-8)),
          "ba150": (lambda: learn.barabasi_albert(150, 0xBA150), (1e-12,))}


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_learned_configuration_vs_oracle(trained_model, name):
    model, _ = trained_model
    ch = model.chooser_weights()
    make, deltas = GRAPHS[name]
    g = make()
    gi = g.info()
    rp, ci, ei, p = g.csr()
    og = OracleGraph.from_csr(rp, ci, ei, p)
    for eps in deltas:
        print(f"case c5 {name} eps={eps}", flush=True)
        kw = dict(truncation_eps=eps, queue_capacity=1 << 26, sweep_period=1)
        r = nr.solve_learned(g, model, rule="alternate", **kw)  # learned order
        o = og.solve(gi.src, gi.dst, mode=MODE_ALTERNATE, order=ORDER_LEVEL, trunc_eps=eps,
                     sweep_period=1)
        assert r["overflowed"] == 0 and counts(r) == counts(o)
        assert close(r["R"], o["R"]) and close(r["ESP"], o["ESP"])
        r = nr.solve_learned(g, model, rule="learned", **kw)  # learned order + choice
        o = og.solve(gi.src, gi.dst, mode=MODE_LEARNED, order=ORDER_LEVEL, trunc_eps=eps,
                     sweep_period=1, chooser=ch)
        assert r["overflowed"] == 0 and counts(r) == counts(o)
        assert close(r["R"], o["R"]) and close(r["ESP"], o["ESP"])


def test_binding_budget_brackets_intersect_the_oracles(trained_model):
    """0.2 s on the grid at delta 1e-12 (the alternating rule needs ~0.85 s
    there): the deadline folds the queue into ESP and the bracket must still
    contain R, hence meet the oracle's bracket from a delta = 1e-8 run."""
    model, _ = trained_model
    g = learn.grid(10, 10, 0x6121D)
    gi = g.info()
    rp, ci, ei, p = g.csr()
    o = OracleGraph.from_csr(rp, ci, ei, p).solve(gi.src, gi.dst, mode=MODE_ALTERNATE,
                                                  order=ORDER_DFS, threads=16, trunc_eps=1e-8,
                                                  sweep_period=1)
    for rule in ("alternate", "learned"):
        print(f"case c5 budget grid rule {rule}", flush=True)
        r = nr.solve_learned(g, model, rule=rule, truncation_eps=1e-12, queue_capacity=1 << 26,
                             sweep_period=1, time_budget_s=0.2)
        assert r["deadline_hit"] == 1 and r["folded"] > 0
        assert abs(r["R"] + r["D"] + r["ESP"] - 1.0) <= 1e-12
        assert max(r["R"], o["R"]) <= min(r["R"] + r["ESP"], o["R"] + o["ESP"]) * (1 + 1e-12)
