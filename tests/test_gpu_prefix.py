"""The static prefix partition on the GPU (nr_config.prefix_edges, the
reference's TID_TO_MASK of original/Rel_16M.cu:97-101): the root split into
2^t seeds over its first t UNKNOWN edges, optionally one shard of them.
Counts, levels and widest level equal the oracle's from the same seeds, R
the closed form; the shards of a split add up to the whole, which is how the
solve would be spread over several GPUs without communication."""
import pytest

from conftest import closed_form, graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from oracle_ffi import MODE_ALTERNATE, ORDER_LEVEL, OracleGraph  # noqa: E402


def close(a, b, rel=1e-12):
    return abs(a - b) <= rel * max(abs(a), abs(b)) + 1e-300


@pytest.mark.parametrize("name,n,p,t", [("COMP8-09.txt", 8, "0.9", 4),
                                        ("COMP10-05.txt", 10, "0.5", 10),
                                        ("COMP9.txt", 9, "0.5", 1)])
def test_prefix_seeds_vs_oracle(expected, name, n, p, t):
    print(f"case prefix {name} t={t}", flush=True)
    g = nr.Graph.load(graph_path(name), 0, n - 1)
    r = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, prefix_edges=t)
    o = OracleGraph.load(graph_path(name)).solve(0, n - 1, order=ORDER_LEVEL, prefix_edges=t)
    assert (r["paths"], r["nodes"], r["dead"], r["iterations"], r["max_queue_items"]) == (
        o["paths"], o["nodes"], o["dead"], o["iterations"], o["max_width"])
    assert close(r["R"], o["R"]) and close(r["R"], closed_form(expected, n, p))
    m = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, prefix_edges=t, order="mass")
    assert (m["paths"], m["nodes"], m["dead"]) == (r["paths"], r["nodes"], r["dead"])
    a = nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 22, prefix_edges=t,
                 rule="alternate", sweep_period=2)
    oa = OracleGraph.load(graph_path(name)).solve(0, n - 1, mode=MODE_ALTERNATE,
                                                  order=ORDER_LEVEL, prefix_edges=t,
                                                  sweep_period=2)
    assert (a["paths"], a["nodes"], a["iterations"]) == (oa["paths"], oa["nodes"],
                                                        oa["iterations"])


def test_prefix_shards_add_up_to_the_whole():
    print("case prefix shards K10 t=8 over 5 shards", flush=True)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    kw = dict(truncation_eps=0.0, queue_capacity=1 << 22, prefix_edges=8)
    whole = nr.solve(g, **kw)
    parts = [nr.solve(g, shard=s, shards=5, **kw) for s in range(5)]
    for f in ("paths", "nodes", "dead"):
        assert sum(x[f] for x in parts) == whole[f], f
    assert close(sum(x["R"] for x in parts), whole["R"], 1e-13)
    assert close(sum(x["D"] for x in parts), whole["D"], 1e-11)


def test_prefix_rejects_bad_arguments():
    g = nr.Graph.load(graph_path("COMP8-09.txt"), 0, 7)
    for kw in (dict(prefix_edges=25), dict(prefix_edges=3, shard=4, shards=4),
               dict(prefix_edges=3, shard=-1, shards=2)):
        with pytest.raises(nr.NRError) as ei:
            nr.solve(g, truncation_eps=0.0, queue_capacity=1 << 16, **kw)
        assert ei.value.status == 4, kw
