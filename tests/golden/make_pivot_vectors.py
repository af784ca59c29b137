#!/usr/bin/env python3
"""Writes pivot_ref.json: the reference's own edge-pivot CPU program
(workingPad/cpu_reference.cpp, compiled into oracle/_ref/ref_cpu_pivot by
`make -C oracle ref`) run on header-stripped copies of graph fixtures (its
loader reads "N E" with no type token, cpu_reference.cpp:38, and holds at
most 256 edges, :22).  It parses p as float32, so its R is the float-p
value; its node and path counts are exact and are what the oracle's pivot
mode must reproduce.  Run from the repository root after `make -C oracle ref`.
"""
import json
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_cpu_pivot")
CASES = [("test_simple.txt", 0, 3), ("test_bridge.txt", 0, 3), ("test_kite.txt", 0, 4),
         ("test_kite.txt", 0, 3), ("cycle4.txt", 0, 2), ("temp4x4.txt", 0, 3),
         ("COMP8.txt", 0, 7), ("COMP8-03.txt", 0, 7), ("COMP8-09.txt", 0, 7)]


def main():
    out = []
    with tempfile.TemporaryDirectory() as td:
        for name, s, t in CASES:
            lines = open(os.path.join(HERE, "graphs", name)).read().split("\n")
            head = lines[0].split()
            path = os.path.join(td, name)
            with open(path, "w") as f:
                f.write(" ".join(head[:2]) + "\n" + "\n".join(lines[1:]))
            txt = subprocess.run([REF, path, str(s), str(t)], capture_output=True, text=True,
                                 check=True).stdout
            out.append({"graph": name, "src": s, "dst": t,
                        "R_float_p": float(re.search(r"R = ([0-9.eE+-]+)", txt).group(1)),
                        "paths": int(re.search(r"Paths enumerated: (\d+)", txt).group(1)),
                        "nodes": int(re.search(r"Nodes processed:\s+(\d+)", txt).group(1)),
                        "source": "workingPad/cpu_reference.cpp (oracle/_ref/ref_cpu_pivot)"})
    with open(os.path.join(HERE, "pivot_ref.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
