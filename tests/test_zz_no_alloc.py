"""nr_solve_static (no device allocation: every buffer in the library's
fixed scratch array) against nr_solve on the same inputs: counts, levels and
widest level identical, R, ESP and D equal.  Two calls in a row, the second
with a smaller input (the array keeps the first call's contents, so what the
second needs must be cleared by the call itself), then a request beyond the
array, which must fail with NR_ERR_SCRATCH and leave the next call intact."""
import pytest

from conftest import graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")

FIELDS = ("paths", "nodes", "dead", "truncated", "overflowed", "folded", "iterations",
          "max_queue_items", "item_bytes")


def same(a, b):
    # no overflow: which children overflow a full segment depends on the
    # order waves append, so only runs without overflow are reproducible
    assert a["overflowed"] == 0 and b["overflowed"] == 0
    assert tuple(a[f] for f in FIELDS) == tuple(b[f] for f in FIELDS)
    for f in ("R", "ESP", "D"):
        assert abs(a[f] - b[f]) <= 1e-15 * max(abs(b[f]), 1e-300), f


@pytest.mark.parametrize("rule,order", [("path", "level"), ("alternate", "mass")])
def test_static_solve_equals_nr_solve(rule, order):
    print(f"case no-alloc K9 then K8 rule {rule} order {order}", flush=True)
    k9 = nr.Graph.load(graph_path("COMP9.txt"), 0, 8)
    k8 = nr.Graph.load(graph_path("COMP8-09.txt"), 0, 7)
    kw9 = dict(truncation_eps=0.0, queue_capacity=1 << 18, rule=rule, order=order)
    kw8 = dict(truncation_eps=0.0, queue_capacity=1 << 16, rule=rule, order=order)
    a9 = nr.solve_static(k9, **kw9)
    a8 = nr.solve_static(k8, **kw8)   # smaller input after a larger one
    same(a9, nr.solve(k9, **kw9))
    same(a8, nr.solve(k8, **kw8))
    assert a9["nodes"] > a8["nodes"] > 0
    if rule == "path":  # the logged trees (output_maxMem_forkPath.txt)
        assert (a9["paths"], a9["nodes"]) == (412015, 824030)
        assert (a8["paths"], a8["nodes"]) == (26830, 53660)


def test_static_solve_rejects_sizes_beyond_the_scratch_array():
    print("case no-alloc K10 beyond the scratch array", flush=True)
    k10 = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    with pytest.raises(nr.NRError) as ei:
        nr.solve_static(k10, truncation_eps=0.0, queue_capacity=1 << 22)
    assert ei.value.status == 8
    k8 = nr.Graph.load(graph_path("COMP8-09.txt"), 0, 7)
    kw = dict(truncation_eps=0.0, queue_capacity=1 << 16)
    same(nr.solve_static(k8, **kw), nr.solve(k8, **kw))
