import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE_DIR = os.path.join(ROOT, "oracle")
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)

GOLDEN = os.path.join(ROOT, "tests", "golden")
GRAPHS = os.path.join(GOLDEN, "graphs")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct GPU (gfx950)")


@pytest.fixture(scope="session")
def expected():
    with open(os.path.join(GOLDEN, "expected.json")) as f:
        return json.load(f)


def graph_path(name: str) -> str:
    return os.path.join(GRAPHS, name)


def closed_form(expected, n, p):
    for x in expected["closed_form"]:
        if x["n"] == n and x["p"] == p:
            return x["R"]
    raise KeyError((n, p))


@pytest.fixture(scope="session")
def trained_model():
    """The learned parts as the benchmark trains them, on the GPU through the
    C-ABI (relaibilitygpu_amd/learn.py): (Model with order scorer and path
    chooser, chooser summary)."""
    import relaibilitygpu_amd as nr
    from relaibilitygpu_amd import learn

    print("case training the order scorer and the path chooser", flush=True)
    order_params, _ = learn.train_order_scorer(alpha=0.5)
    ch_params, margin, info = learn.train_chooser()
    return nr.Model(order_params, alpha=0.5, chooser=ch_params, margin=margin), info
