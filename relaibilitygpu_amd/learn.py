"""Training of the learned parts on the GPU, through the C-ABI.

Both networks are trained by the engine's own kernels (nr_mlp_train: FP32
forward / backward, Adam) on data the engine logs while it solves:

* the order scorer (nr_model_create): log2 subtree sizes of children logged
  in level order (nr_log_subproblems) on K10, the reference's K10 workload;
* the path chooser (nr_model_set_chooser): for each expanded subproblem, the
  choice features of the sweep direction taken and log2 of its subtree size,
  logged by nr_log_choices while the alternating rule's direction is flipped
  for a quarter of the subproblems.  The labels then measure "this
  direction, then the alternating rule below", so sweeping where the
  prediction is smaller is a step of policy improvement over that rule.

A single chooser's choices are noisy: they rest on the difference of two
predicted log-sizes, small beside the network's error, and different seeds
disagree.  So several members are trained, the median of their predictions
is distilled into one network of the same shape (the form the expand kernel
evaluates), and its margin (the predicted advantage needed to leave the
alternating direction) is picked on validation solves of graphs that the
evaluations do not use (complete graphs truncated and capacity-bound, a grid
and a Barabasi-Albert graph of other sizes): the geometric mean of the gap
ratios over those regimes decides, since a chooser that helps in one regime
can hurt in another.  No counterpart in the reference.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import (Graph, Model, log_choices, log_subproblems, mlp_train, solve, solve_learned)
from . import workloads as wl


def complete(n: int, p: float = 0.5) -> Graph:
    """K_n with uniform p, terminals 0 -> n-1."""
    nn, edges, probs = wl.complete_graph(n, p)
    return Graph.from_edges(nn, edges, probs, 0, n - 1)


def grid(rows: int, cols: int, seed: int) -> Graph:
    """rows x cols grid, p ~ U[0.85, 0.95] seeded, corner to corner."""
    rng = wl.Xoshiro256ss(seed)
    n, edges, _ = wl.grid_graph(rows, cols, [0.0] * (2 * rows * cols))
    probs = [0.85 + 0.10 * rng.uniform() for _ in edges]
    rp, ci, ei = wl.to_csr(n, edges)
    return Graph.from_csr(rp, ci, ei, probs, 0, n - 1)


def barabasi_albert(n: int, seed: int, p: float = 0.95) -> Graph:
    """Barabasi-Albert, m = 2, uniform p, terminals 0 -> n-1."""
    nn, edges, probs = wl.barabasi_albert(n, 2, p, seed)
    rp, ci, ei = wl.to_csr(nn, edges)
    return Graph.from_csr(rp, ci, ei, probs, 0, n - 1)


def _mlp_forward(params: np.ndarray, X: np.ndarray) -> np.ndarray:
    W1 = params[:1024].reshape(16, 64)
    w2 = params[1024:1088]
    return np.maximum(X @ W1, 0.0) @ w2 + params[1088]


# ------------------------------------------------------------ order scorer
def train_order_scorer(alpha: float = 0.5, device: int = 0, records: int = 2000000):
    """The order scorer on `records` children logged from K10 solves of the
    alternating rule, the rule it orders (up to 200 000 each at delta 1e-6
    and 1e-9, the rest from the full tree): (parameters, summary)."""
    t0 = time.perf_counter()
    k10 = complete(10)
    Xs, Ys = [], []
    for eps in (1e-6, 1e-9, 0.0):
        m = 200000 if eps > 0 else records - sum(len(y) for y in Ys)
        X, Y, _ = log_subproblems(k10, m, seed=7, truncation_eps=eps, queue_capacity=1 << 22,
                                  rule="alternate", device=device)
        Xs.append(X)
        Ys.append(Y)
    X, Y = np.concatenate(Xs), np.concatenate(Ys)
    t_log = time.perf_counter() - t0
    params, _, hl = mlp_train(X, Y, None, epochs=10, batch=4096, lr=3e-3, seed=3, device=device)
    info = {"records": int(len(Y)), "label_var": float(Y.var()), "holdout_mse": float(hl[-1]),
            "log_s": t_log, "train_s": time.perf_counter() - t0 - t_log, "alpha": alpha,
            "rule": "alternate", "graph": "K10"}
    return params, info


# ------------------------------------------------------------ path chooser
#: (n of K_n, truncation, records): K10 in full (the reference's K10
#: workload) and larger complete graphs truncated at 1e-10 .. 1e-14
CHOICE_LOGS = ((10, 0.0, 1000000), (11, 1e-12, 600000), (12, 1e-12, 600000),
               (13, 1e-12, 600000), (12, 1e-10, 300000), (11, 1e-14, 300000))


def log_choice_data(logs=CHOICE_LOGS, period: int = 6, seed: int = 7, device: int = 0):
    """Choice logs of complete graphs: (X [m,16], Y [m], per-log summary)."""
    Xs, Ys, info = [], [], []
    for n, eps, m in logs:
        X, Y, r = log_choices(complete(n), m, seed=seed, truncation_eps=eps,
                              queue_capacity=1 << 26, sweep_period=period, device=device)
        Xs.append(X)
        Ys.append(Y)
        info.append({"graph": f"K{n}", "eps": eps, "records": int(len(X)),
                     "subproblems": int(r["nodes"])})
    return np.concatenate(Xs), np.concatenate(Ys), info


def validation_tasks(cap: int = 96777216):
    """Validation solves (no evaluation graph among them): K14 and K18
    truncated at 1e-12 in level order, K14 untruncated in max-mass order
    under capacities of 2^22 and 2^24 items (the capacity binds), a 7x7 grid
    and a Barabasi-Albert graph of 100 nodes."""
    k14 = complete(14)
    return {
        "K14": (k14, 6, dict(truncation_eps=1e-12, order="level", queue_capacity=cap)),
        "K18": (complete(18), 6, dict(truncation_eps=1e-12, order="level", queue_capacity=cap)),
        "K14_capacity": (k14, 6, dict(truncation_eps=0.0, order="mass", queue_capacity=1 << 22,
                                      time_budget_s=5.0)),
        "K14_capacity_2^24": (k14, 6, dict(truncation_eps=0.0, order="mass",
                                           queue_capacity=1 << 24, time_budget_s=5.0)),
        "grid7x7": (grid(7, 7, 0x7A11), 1, dict(truncation_eps=1e-10, order="level",
                                                 queue_capacity=cap)),
        "ba100": (barabasi_albert(100, 0xBA100), 1, dict(truncation_eps=1e-13, order="level",
                                                         queue_capacity=cap)),
    }


def _gap(task, model, device):
    g, period, kw = task
    if model is None:
        return solve(g, rule="alternate", sweep_period=period, device=device, **kw)["ESP"]
    return solve_learned(g, model, rule="learned", sweep_period=period, device=device,
                         **kw)["ESP"]


def train_chooser(X=None, Y=None, members: int = 8, distill_seeds=(11, 12),
                  margins=(0.0, 0.5, 1.0, 2.0), device: int = 0):
    """The path chooser: (params, margin, summary).  Members are trained on
    the choice logs with seeds 1..members; the median of their predictions
    is distilled into one network per distillation seed; the (network,
    margin) whose validation gaps have the smallest geometric mean relative
    to the alternating rule is kept."""
    t0 = time.perf_counter()
    logs = None
    if X is None:
        X, Y, logs = log_choice_data(device=device)
    t_log = time.perf_counter() - t0
    preds = []
    for seed in range(1, members + 1):
        p, _, _ = mlp_train(X, Y, None, epochs=10, batch=4096, lr=3e-3, seed=seed, device=device)
        preds.append(_mlp_forward(p, X))
    target = np.median(np.stack(preds), axis=0).astype(np.float32)
    t_train = time.perf_counter() - t0 - t_log
    tasks = validation_tasks()
    base = {k: _gap(t, None, device) for k, t in tasks.items()}
    best = None
    trials = []
    for dseed in distill_seeds:
        params, _, hl = mlp_train(X, target, None, epochs=12, batch=4096, lr=3e-3, seed=dseed,
                                  device=device)
        for margin in margins:
            m = Model(None, chooser=params, margin=margin)
            ratios = {k: base[k] / max(_gap(t, m, device), 1e-300) for k, t in tasks.items()}
            score = math.exp(sum(math.log(r) for r in ratios.values()) / len(ratios))
            trials.append({"distill_seed": dseed, "margin": margin, "validation_gain": score,
                           "ratios": ratios})
            if best is None or score > best[0]:
                best = (score, params, margin)
    summary = {"records": int(len(Y)), "label_var": float(Y.var()), "members": members,
               "log_s": t_log, "train_s": t_train,
               "select_s": time.perf_counter() - t0 - t_log - t_train,
               "margin": best[2], "validation_gain": best[0], "trials": trials,
               "validation_alternate_gap": base}
    if logs is not None:
        summary["logs"] = logs
    return best[1], best[2], summary
