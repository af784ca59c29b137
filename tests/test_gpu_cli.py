"""The product command line (relaibilitygpu_amd/bin/network_reliability) on
the GPU: its report block keeps the layout of workingPad/main.cu:1037-1056
line for line, and its values equal the oracle's (counts bit-exact, R and
ESP to 1e-12).  The pre-run lines follow graph.h:142-143 and main.cu:944-948.
"""
import re
import subprocess

import pytest

from conftest import graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
from oracle_ffi import OracleGraph  # noqa: E402

SEP = "=" * 44
DASH = "-" * 44
# (label regex, value regex): the printf formats of main.cu:1037-1056
F12 = r"(-?\d+\.\d{12})"
E6 = r"(-?\d\.\d{6}e[+-]\d{2})"
LINES = [
    (r"Graph:            N = (\d+), E = (\d+)", "graph"),
    (r"Terminals:        src = (\d+), dst = (\d+)", "terms"),
    (DASH, None),
    (r"Reliability:      R\(G\) = " + F12, "R"),
    (r"R \+ ESP:                 " + F12, "RE"),
    (r"ESP \(truncated\):         " + E6, "ESP"),
    (r"Truncation bound:        " + E6, "bound"),
    (DASH, None),
    (r"Paths enumerated: (\d+)", "paths"),
    (r"Nodes processed:  (\d+)", "nodes"),
    (r"Kernel iterations:(\d+)", "iters"),
    (r"Max queue RSS:    (\d+\.\d{3}) MB  \((\d+) items\)", "rss"),
    (r"Total time:       (\d+\.\d{3}) ms", "time"),
    (SEP, None),
]


def parse_report(out: str) -> dict:
    lines = out.splitlines()
    i = lines.index("  Two-Terminal Network Reliability Results")
    assert lines[i - 1] == SEP and lines[i + 1] == SEP
    got = {}
    for (pat, key), line in zip(LINES, lines[i + 2:]):
        if key is None:
            assert line == pat
            continue
        m = re.fullmatch(pat, line)
        assert m, (pat, line)
        got[key] = m.groups()
    assert len(lines) == i + 2 + len(LINES), "lines after the report block"
    return got


@pytest.mark.parametrize("name,s,t,args", [
    ("test_simple.txt", 0, 3, []),
    ("test_kite.txt", 0, 4, ["--truncation_eps", "0"]),
    ("COMP8-09.txt", 0, 7, ["--queue_capacity", "65536", "--truncation_eps", "0"]),
    ("COMP10-05.txt", 0, 9, ["--queue_capacity", "4194304"]),
])
def test_report_matches_oracle(name, s, t, args):
    print(f"case cli {name} {' '.join(args)}", flush=True)
    r = subprocess.run([nr.CLI_PATH, graph_path(name), str(s), str(t)] + args,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert re.search(r"^Graph loaded: N=\d+  E=\d+", r.stdout, flags=re.M)
    assert re.search(r"^Launch config: \d+ blocks x \d+ threads", r.stdout, flags=re.M)
    got = parse_report(r.stdout)
    og = OracleGraph.load(graph_path(name))
    eps = float(args[args.index("--truncation_eps") + 1]) if "--truncation_eps" in args else 1e-15
    o = og.solve(s, t, trunc_eps=eps)
    d = og.dims()
    assert got["graph"] == (str(d["n"]), str(d["e"]))
    assert got["terms"] == (str(s), str(t))
    R, RE, ESP, bound = (float(got[k][0]) for k in ("R", "RE", "ESP", "bound"))
    assert abs(R - o["R"]) <= 1e-12 and abs(RE - (o["R"] + o["ESP"])) <= 1e-12
    assert abs(ESP - o["ESP"]) <= 1e-6 * o["ESP"] + 1e-300
    assert abs(bound - (o["ESP"] / o["R"] if o["R"] > 0 else o["ESP"])) <= 1e-6 * bound + 1e-300
    assert int(got["paths"][0]) == o["paths"] and int(got["nodes"][0]) == o["nodes"]
    assert int(got["iters"][0]) == o["iterations"]
    mb, items = got["rss"]
    assert int(items) == o["max_width"]
    W = og.mask_words
    assert abs(float(mb) - int(items) * (16 * W + 8) / 2**20) <= 0.0005 + 1e-9
    assert float(got["time"][0]) > 0.0


def test_json_line_matches_report():
    print("case cli json K8", flush=True)
    args = [nr.CLI_PATH, graph_path("COMP8-09.txt"), "0", "7", "--truncation_eps", "0",
            "--queue_capacity", "65536"]
    import json

    j = json.loads(subprocess.run(args + ["--json"], capture_output=True, text=True,
                                  check=True, timeout=120).stdout)
    rep = parse_report(subprocess.run(args, capture_output=True, text=True, check=True,
                                      timeout=120).stdout)
    assert (j["paths"], j["nodes"], j["iterations"]) == tuple(
        int(rep[k][0]) for k in ("paths", "nodes", "iters"))
    assert f"{j['R']:.12f}" == rep["R"][0]
