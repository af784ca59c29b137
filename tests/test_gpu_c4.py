"""The plain GPU method at the certified-gap configuration (K15 and K25,
p = 0.5, path rule, truncation 1e-12, queue capacity 96 777 216 items as
workingPad/README.md:95,160 runs them) against the oracle: paths, nodes,
dead and truncated bit-exact, R and ESP within 1e-12, no overflow, and the
bracket R <= R_true <= R + ESP around the closed form.  This leg is the
baseline of every certified-gap ratio, so it is checked on its own, in level
and in max-mass order (the same tree: order only matters under a limit).
"""
import pytest

from conftest import closed_form, graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from oracle_ffi import ORDER_DFS, OracleGraph  # noqa: E402

CAP = 96777216
_oracle = {}


def oracle_run(name, n):
    if name not in _oracle:
        _oracle[name] = OracleGraph.load(graph_path(name)).solve(
            0, n - 1, order=ORDER_DFS, threads=16, trunc_eps=1e-12)
    return _oracle[name]


def close(a, b, rel=1e-12):
    return abs(a - b) <= rel * max(abs(a), abs(b))


@pytest.mark.parametrize("name,n", [("COMP15-05.txt", 15), ("COMP25-05.txt", 25)])
@pytest.mark.parametrize("order", ["level", "mass"])
def test_plain_rule_at_1e_12_vs_oracle(expected, name, n, order):
    print(f"case c4 plain {name} {order}", flush=True)
    g = nr.Graph.load(graph_path(name), 0, n - 1)
    r = nr.solve(g, truncation_eps=1e-12, queue_capacity=CAP, order=order)
    o = oracle_run(name, n)
    assert r["overflowed"] == 0 and r["folded"] == 0 and r["deadline_hit"] == 0
    assert (r["paths"], r["nodes"], r["dead"], r["truncated"]) == (
        o["paths"], o["nodes"], o["dead"], o["truncated"])
    assert close(r["R"], o["R"]) and close(r["ESP"], o["ESP"]) and close(r["D"], o["D"])
    R = closed_form(expected, n, "0.5")
    assert r["R"] <= R * (1 + 1e-12) <= (r["R"] + r["ESP"]) * (1 + 2e-12)
    assert abs(r["R"] + r["D"] + r["ESP"] - 1.0) <= 1e-12
