"""Host-side checks that need no GPU: the product loader against the oracle's
independent loader, file-format edge cases, the C-ABI export list, and the
command lines' argument handling."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GRAPHS, ROOT, graph_path
import relaibilitygpu_amd as nr
from oracle_ffi import CLI_PATH as ORACLE_CLI
from oracle_ffi import OracleGraph

ALL_GRAPHS = sorted(f for f in os.listdir(GRAPHS) if f.endswith(".txt"))


@pytest.mark.parametrize("name", ALL_GRAPHS)
def test_loader_matches_oracle_loader(name):
    g = nr.Graph.load(graph_path(name), 0, 1)
    o = OracleGraph.load(graph_path(name))
    gi, od = g.info(), o.dims()
    assert (gi.n, gi.e, gi.nnz, gi.header_e, gi.type) == (
        od["n"], od["e"], od["nnz"], od["header_e"], od["type"])
    for a, b in zip(g.csr(), o.arrays()):
        assert np.array_equal(a, b)


def test_edge_ids_follow_first_appearance():
    """graph.h:76-90: ids in order of first appearance of (min,max)."""
    g = nr.Graph.load(graph_path("test_kite.txt"), 0, 4)
    rp, ci, ei, p = g.csr()
    seen, order = set(), []
    for u in range(len(rp) - 1):
        for j in range(rp[u], rp[u + 1]):
            key = (min(u, ci[j]), max(u, ci[j]))
            if key not in seen:
                seen.add(key)
                order.append(key)
                assert ei[j] == len(order) - 1
    assert list(p) == [0.95, 0.9, 0.8, 0.85, 0.9, 0.7, 0.95]


def _write(tmp_path, text, name="g.txt"):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


def test_header_without_type_and_crlf(tmp_path):
    a = _write(tmp_path, "4 3\r\n1 1 0.9\r\n2 0 0.9 2 0.8\r\n2 1 0.8 3 0.7\r\n1 2 0.7\r\n")
    g = nr.Graph.load(a, 0, 3)
    gi = g.info()
    assert (gi.n, gi.e, gi.type, gi.header_e) == (4, 3, 0, 3)
    b = _write(tmp_path, "4 6 U\n1 1 0.9\n2 0 0.9 2 0.8\n2 1 0.8 3 0.7\n1 2 0.7\n", "b.txt")
    gi = nr.Graph.load(b, 0, 3).info()
    assert (gi.type, gi.header_e) == (ord("U"), 3)


def test_fp64_parse():
    g = nr.Graph.load(graph_path("COMP8-03.txt"), 0, 7)
    assert g.csr()[3][0] == 0.3  # not the float32 0.30000001192...


@pytest.mark.parametrize("text,status", [
    ("", 2), ("3", 2), ("2 1 U\n1 1 0.5\n", 2), ("2 1 U\n1 5 0.5\n1 0 0.5\n", 3),
    ("2 1 U\n1 1 zz\n1 0 0.5\n", 2), ("-1 0\n", 2)])
def test_malformed_files(tmp_path, text, status):
    with pytest.raises(nr.NRError) as ei:
        nr.Graph.load(_write(tmp_path, text), 0, 0)
    assert ei.value.status == status


def test_missing_file_and_bad_terminals():
    with pytest.raises(nr.NRError) as ei:
        nr.Graph.load("/nonexistent/graph.txt", 0, 1)
    assert ei.value.status == 1
    with pytest.raises(nr.NRError) as ei:
        nr.Graph.load(graph_path("test_simple.txt"), 0, 4)
    assert ei.value.status == 3


def test_precontracted_and_deleted_counts(tmp_path):
    a = _write(tmp_path, "3 3\n2 1 1.0 2 0.0\n2 0 1.0 2 0.5\n2 0 0.0 1 0.5\n")
    gi = nr.Graph.load(a, 0, 2).info()
    assert (gi.precontracted, gi.predeleted) == (1, 1)


def test_asymmetric_csr_detected(tmp_path):
    a = _write(tmp_path, "3 2 D\n1 1 0.5\n1 2 0.5\n0\n")
    assert nr.Graph.load(a, 0, 2).info().symmetric == 0


def test_from_edges_roundtrip():
    g = nr.Graph.from_edges(4, [(0, 1), (1, 2), (2, 3)], [0.9, 0.8, 0.7], 0, 3)
    gi = g.info()
    assert (gi.n, gi.e, gi.nnz, gi.symmetric) == (4, 3, 6, 1)


def test_cabi_exports_every_declared_symbol():
    """Every function declared in include/*.h is exported by libnrgpu.so."""
    names = set()
    inc = os.path.join(ROOT, "include")
    for h in os.listdir(inc):
        text = open(os.path.join(inc, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(nr_[a-z_0-9]+)\s*\(", text))
    assert names, "no declarations found"
    out = subprocess.run(["nm", "-D", "--defined-only", nr.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (nr_[a-z_0-9]+)$", out, flags=re.M))
    assert names <= exported, names - exported
    lib = nr.lib()
    for n in names:
        assert hasattr(lib, n)


def test_status_strings():
    lib = nr.lib()
    assert lib.nr_status_str(0) == b"ok"
    assert lib.nr_status_str(5).startswith(b"graph beyond")


def test_cli_usage_exit_code():
    r = subprocess.run([nr.CLI_PATH], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage:" in r.stderr
    r = subprocess.run([nr.CLI_PATH, "/nonexistent", "0", "1"], capture_output=True, text=True)
    assert r.returncode == 1


def test_oracle_cli_report_layout():
    r = subprocess.run([ORACLE_CLI, graph_path("test_simple.txt"), "0", "3"],
                       capture_output=True, text=True, check=True)
    assert "Reliability:      R(G) = 0.504000000000" in r.stdout
    assert "Paths enumerated: 1" in r.stdout and "Nodes processed:  4" in r.stdout


def test_checkpoint_files_fail_cleanly(tmp_path):
    """Loading needs no GPU: a missing file is an I/O error, a file that is
    not a checkpoint (or is cut short) a parse error."""
    with pytest.raises(nr.NRError) as ei:
        nr.Checkpoint.load(str(tmp_path / "missing.bin"))
    assert ei.value.status == 1
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"NRCKPT01" + b"\x00" * 20)
    with pytest.raises(nr.NRError) as ei:
        nr.Checkpoint.load(str(bad))
    assert ei.value.status == 2
    bad.write_bytes(b"not a checkpoint at all, just text")
    with pytest.raises(nr.NRError) as ei:
        nr.Checkpoint.load(str(bad))
    assert ei.value.status == 2


def _ckpt_header(w: int, count: int) -> bytes:
    """A checkpoint file header (layout of nr_checkpoint_save): magic, eight
    ints, eps, graph hash, three sums and their low parts, eight counters,
    iterations, widest level and the item count."""
    import struct

    return (b"NRCKPT01" + struct.pack("<8i", 10, 45, w, 0, 9, 0, 0, 0)
            + struct.pack("<2d", 0.0, 0.0) + b"\x00" * 8 * 3 * 2 + b"\x00" * 8 * 8
            + struct.pack("<3Q", 0, 0, count))


def test_checkpoint_oversized_headers_fail_cleanly(tmp_path):
    """A header whose item count the file cannot hold is a parse error, never
    an allocation of that size (nor an abort of the caller)."""
    f = tmp_path / "big.bin"
    for w, count, items in ((16, (1 << 34) - 1, b""), (1, 1000, b"\x00" * 24 * 999),
                            (1, 2, b"\x00" * 24 * 3)):
        f.write_bytes(_ckpt_header(w, count) + items)
        with pytest.raises(nr.NRError) as ei:
            nr.Checkpoint.load(str(f))
        assert ei.value.status == 2, (w, count)
    f.write_bytes(_ckpt_header(1, 2) + b"\x00" * 24 * 2)
    ck = nr.Checkpoint.load(str(f))
    assert ck.info()["pending"] == 2


def test_graph_degree_beyond_the_text_is_a_parse_error(tmp_path):
    f = tmp_path / "g.txt"
    f.write_text("2 2 U\n100000000000000 1 0.5\n1 0 0.5\n")
    with pytest.raises(nr.NRError) as ei:
        nr.Graph.load(str(f), 0, 1)
    assert ei.value.status == 2


def test_cli_rejects_unknown_rule_or_order():
    for flag, value in (("--rule", "forkBest"), ("--rule", "heap"), ("--order", "heap")):
        r = subprocess.run([nr.CLI_PATH, graph_path("test_simple.txt"), "0", "3", flag, value],
                           capture_output=True, text=True)
        assert r.returncode == 1 and "Usage:" in r.stderr and value in r.stderr, (flag, value)


def test_log_rel_argument_checks():
    """nr_log_subproblems_rel rejects missing outputs and bad configs
    before it touches a device (NR_ERR_ARG = 4, NR_ERR_UNSUPPORTED for a
    rule whose tree is not logged)."""
    import ctypes

    lib = nr.lib()
    g = nr.Graph.load(graph_path("COMP8-09.txt"), 0, 7)
    cfg = nr.make_config(truncation_eps=0.0, queue_capacity=1 << 16)
    X = (ctypes.c_float * 16)()
    Y = (ctypes.c_float * 1)()
    C = (ctypes.c_double * 1)()
    n = ctypes.c_longlong()
    res = nr.Result()
    f = lib.nr_log_subproblems_rel
    assert f(None, ctypes.byref(cfg), 1, 1, X, Y, C, ctypes.byref(n), ctypes.byref(res)) == 4
    assert f(g.handle, ctypes.byref(cfg), 1, 1, X, Y, None, ctypes.byref(n), ctypes.byref(res)) == 4
    assert f(g.handle, ctypes.byref(cfg), -1, 1, X, Y, C, ctypes.byref(n), ctypes.byref(res)) == 4
    assert f(g.handle, ctypes.byref(cfg), 1, 1, X, Y, C, None, ctypes.byref(res)) == 4
    bad = nr.make_config(truncation_eps=-1.0)
    assert f(g.handle, ctypes.byref(bad), 1, 1, X, Y, C, ctypes.byref(n), ctypes.byref(res)) == 4
    fork = nr.make_config(rule="forkbest")
    assert f(g.handle, ctypes.byref(fork), 1, 1, X, Y, C, ctypes.byref(n), ctypes.byref(res)) != 0
