"""The ramp (k_ramp_levels): after the levels one workgroup runs alone, every
workgroup expands its own slice's subtrees level after level, with no launch
per level, until the levels hold about nr_config.ramp_items items.  The tree
is the one the level launches build, so every count, the number of levels and
the widest level equal the oracle's (workingPad/main.cu:963-1008 counts one
level per launch) and those of a solve without the ramp (ramp_items < 0).  A
ramp whose workgroup region overflows is discarded and the solve runs again
without it (nr_result.rerun), so the counts never depend on it."""
import pytest

from conftest import graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from oracle_ffi import MODE_ALTERNATE, MODE_PATH, ORDER_LEVEL, OracleGraph  # noqa: E402

COUNTS = ("paths", "nodes", "dead", "truncated", "overflowed", "iterations", "max_queue_items")


def close(a, b, rel=1e-12):
    return abs(a - b) <= rel * max(abs(a), abs(b)) + 1e-300


def same(a, b, rel=1e-12):
    assert tuple(a[f] for f in COUNTS) == tuple(b[f] for f in COUNTS)
    assert close(a["R"], b["R"], rel) and close(a["ESP"], b["ESP"], rel)


@pytest.mark.parametrize("name,n,rule,eps", [("COMP10-05.txt", 10, "path", 0.0),
                                             ("COMP10-05.txt", 10, "alternate", 0.0),
                                             ("COMP9.txt", 9, "path", 0.0),
                                             ("COMP15-05.txt", 15, "path", 1e-9),
                                             ("COMP15-05.txt", 15, "alternate", 1e-9)])
def test_ramp_vs_oracle_and_launches(name, n, rule, eps):
    print(f"case ramp {name} {rule} eps={eps}", flush=True)
    g = nr.Graph.load(graph_path(name), 0, n - 1)
    kw = dict(truncation_eps=eps, queue_capacity=1 << 22, rule=rule, sweep_period=2)
    on = nr.solve(g, **kw)
    off = nr.solve(g, ramp_items=-1, **kw)
    assert on["rerun"] == 0 and off["rerun"] == 0, (on["grid_blocks"], on["block_size"])
    same(on, off)
    mode = MODE_ALTERNATE if rule == "alternate" else MODE_PATH
    o = OracleGraph.load(graph_path(name)).solve(0, n - 1, mode=mode, order=ORDER_LEVEL,
                                                 trunc_eps=eps, sweep_period=2)
    assert (on["paths"], on["nodes"], on["dead"], on["iterations"], on["max_queue_items"]) == (
        o["paths"], o["nodes"], o["dead"], o["iterations"], o["max_width"])
    assert close(on["R"], o["R"])


@pytest.mark.parametrize("target", [1 << 12, 1 << 16, 1 << 20])
def test_ramp_lengths(target):
    """Other ramp lengths (4, 8 and 12 levels on K10) build the same tree."""
    print(f"case ramp K10 target={target}", flush=True)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    kw = dict(truncation_eps=0.0, queue_capacity=1 << 23)
    on = nr.solve(g, ramp_items=target, **kw)
    assert on["rerun"] == 0
    same(on, nr.solve(g, ramp_items=-1, **kw))


def test_ramp_overflow_runs_again():
    """A ramp of about 16 levels on K10 overflows the 4096-item region of the
    workgroup holding the fastest-growing subtree: the solve runs again
    without it and reports the same tree."""
    print("case ramp overflow K10", flush=True)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    kw = dict(truncation_eps=0.0, queue_capacity=1 << 22)
    on = nr.solve(g, ramp_items=1 << 23, **kw)
    assert on["rerun"] == 1 and on["overflowed"] == 0
    same(on, nr.solve(g, ramp_items=-1, **kw))
    again = nr.solve(g, **kw)  # the next solve tries the default ramp afresh
    assert again["rerun"] == 0
    same(again, on)


@pytest.mark.parametrize("kind", ["grid10x10", "ba120", "lean", "two_blocks"])
def test_ramp_forms(kind):
    """The CSR form (100- and 120-node graphs), the lean row form and a grid of
    two workgroups (one ramp workgroup)."""
    print(f"case ramp form {kind}", flush=True)
    from relaibilitygpu_amd import learn

    if kind == "grid10x10":
        g, kw = learn.grid(10, 10, 0x51), dict(truncation_eps=1e-7)
    elif kind == "ba120":
        g, kw = learn.barabasi_albert(120, 0xBA), dict(truncation_eps=1e-7)
    elif kind == "lean":
        g, kw = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9), dict(truncation_eps=0.0,
                                                                      lean=True)
    else:
        g, kw = nr.Graph.load(graph_path("COMP9.txt"), 0, 8), dict(truncation_eps=0.0,
                                                                  grid_blocks=2)
    kw["queue_capacity"] = 1 << 22
    on = nr.solve(g, **kw)
    off = nr.solve(g, ramp_items=-1, **kw)
    assert on["rerun"] == 0
    same(on, off)


def test_ramp_static_entry_point():
    """nr_solve_static runs the ramp from its fixed scratch array too."""
    print("case ramp static K9", flush=True)
    g = nr.Graph.load(graph_path("COMP9.txt"), 0, 8)
    kw = dict(truncation_eps=0.0, queue_capacity=1 << 18)
    same(nr.solve_static(g, **kw), nr.solve(g, ramp_items=-1, **kw))
