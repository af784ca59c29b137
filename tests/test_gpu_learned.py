"""GPU tests of the learned work order.

* training kernels: the mean loss and gradient of nr_mlp_grad against the
  numpy oracle (oracle/mlp_oracle.py, pinned by finite differences) to FP32
  summation accuracy; Adam training of a teacher network's outputs drives
  the holdout loss down;
* inference kernel (BF16 matrix cores): nr_model_score against the oracle's
  BF16 inference form;
* tree log: every child of the K8 p=0.9 tree is logged once, and the
  multiset of labels (log2 subtree sizes) equals the one a recursive walk of
  the oracle's expansions gives;
* learned order: the same exact decomposition as level order (counts
  bit-exact, R to 1e-12 on K10 full enumeration), certified bracket under a
  deadline on K15.
"""
import math

import numpy as np
import pytest

from conftest import closed_form, graph_path

pytestmark = pytest.mark.gpu

nr = pytest.importorskip("relaibilitygpu_amd")
import mlp_oracle as mo  # noqa: E402
from oracle_ffi import OracleGraph  # noqa: E402


def rand_data(seed, n):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, mo.N_PARAMS).astype(np.float32)
    X = rng.normal(0, 1, (n, 16)).astype(np.float32)
    X[:, 15] = 1.0
    y = rng.normal(0, 2, n).astype(np.float32)
    return p, X, y


def fp32_grad_bound(p, X, y):
    """Per-entry bound of an FP32 gradient's distance to the exact one:
    each entry is a sum of at most 64 + n/64 + 8 rounded terms (within that
    many units of 2^-24 of the sum of the terms' magnitudes, the residual's
    own rounding included), plus the terms of samples whose hidden
    pre-activation is within its rounding error of 0, where FP32 and exact
    arithmetic may disagree on the ReLU mask."""
    U = 2.0 ** -24
    n = len(y)
    W1, w2, b2 = mo.unpack(p.astype(np.float64))
    Xd = X.astype(np.float64)
    Z = Xd @ W1
    Hh = np.maximum(Z, 0)
    zt = 32 * U * (np.abs(Xd) @ np.abs(W1))
    amb = np.abs(Z) <= zt
    et = 96 * U * (np.abs(b2) + np.abs(y) + (Hh + zt) @ np.abs(w2))
    rr = np.abs(b2 + Hh @ w2 - y) + et
    chain = 64 + (n + 63) // 64 + 8
    gW = np.abs(Xd).T @ (((Z > 0) | amb) * np.outer(rr, np.abs(w2))) / n
    sums = np.concatenate([gW.ravel(), (Hh + zt).T @ rr / n, [rr.mean()]])
    flip = np.concatenate([(np.abs(Xd).T @ (amb * np.outer(rr, np.abs(w2))) / n).ravel(),
                           (amb * zt).T @ rr / n, [0.0]])
    return 1e-7 + 2 * chain * U * sums + flip


@pytest.mark.parametrize("n", [1, 63, 1000, 20000])
def test_gradient_matches_oracle(n):
    print(f"case mlp grad n={n}", flush=True)
    p, X, y = rand_data(n, n)
    loss, g = nr.mlp_grad(X, y, p)
    lo, go = mo.loss_and_grad(p.astype(np.float64), X, y)
    assert abs(loss - lo) <= 1e-5 * (1 + abs(lo))
    assert np.all(np.abs(g - go) <= fp32_grad_bound(p, X, y))


def test_inference_matches_bf16_form():
    print("case mlp inference", flush=True)
    p, X, _ = rand_data(7, 1000)
    m = nr.Model(p, alpha=1.0)
    y = m.score(X)
    ref = mo.infer_bf16(p.astype(np.float64), X)
    scale = np.abs(X.astype(np.float64)) @ np.abs(mo.unpack(p)[0]) @ np.abs(mo.unpack(p)[1])
    assert np.all(np.abs(y - ref) <= 1e-5 * (1 + scale))
    # and the inference form tracks the FP32 network up to BF16 rounding
    assert np.all(np.abs(y - mo.forward(p, X)) <= 0.02 * (1 + scale))


def test_training_fits_a_teacher():
    print("case mlp training", flush=True)
    rng = np.random.default_rng(11)
    teacher = rng.normal(0, 0.4, mo.N_PARAMS)
    X = rng.uniform(-1, 1, (40000, 16)).astype(np.float32)
    X[:, 15] = 1.0
    y = mo.forward(teacher, X).astype(np.float32)
    p0 = nr.mlp_init(3)
    l0 = mo.loss_and_grad(p0.astype(np.float64), X, y)[0]
    p, tl, hl = nr.mlp_train(X, y, p0, epochs=6, batch=1024, lr=3e-3, seed=5)
    assert np.all(np.isfinite(p))
    assert hl[-1] < 0.1 * l0, (l0, list(hl))
    assert tl[-1] < tl[0]


def _walk_labels(og, s, t):
    """log2 subtree sizes of every non-root node, by recursion over the
    oracle's expansions (level-order tree = path-decomposition tree)."""
    W = og.mask_words
    labels = []

    def size(known, up, mass):
        k, succ, ck, cu, cm, _ = og.expand(s, t, known, up, mass)
        tot = 1
        for i in range(max(k, 0)):
            sz = size(ck[i], cu[i], float(cm[i]))
            labels.append(math.log2(sz))
            tot += sz
        return tot

    z = np.zeros(W, np.uint64)
    nodes = size(z, z, 1.0)
    return nodes, sorted(labels)


def test_tree_log_subtree_sizes():
    print("case tree log K8 p=0.9", flush=True)
    g = nr.Graph.load(graph_path("COMP8-09.txt"), 0, 7)
    X, Y, res = nr.log_subproblems(g, 1 << 20, seed=1, truncation_eps=0.0,
                                   queue_capacity=1 << 16)
    assert res["nodes"] == 53660 and len(Y) == res["nodes"] - 1
    assert np.all(X[:, 15] == 1.0) and np.all(np.isfinite(X))
    og = OracleGraph.load(graph_path("COMP8-09.txt"))
    nodes, labels = _walk_labels(og, 0, 7)
    assert nodes == 53660
    assert np.allclose(np.sort(Y.astype(np.float64)), np.array(labels), atol=1e-6)
    # sampling returns a subset of the same records
    Xs, Ys, _ = nr.log_subproblems(g, 1000, seed=2, truncation_eps=0.0, queue_capacity=1 << 16)
    assert len(Ys) == 1000
    assert set(np.round(Ys, 5)).issubset(set(np.round(Y, 5)))


def test_learned_order_is_exact(expected):
    print("case learned order K10 exact", flush=True)
    rng = np.random.default_rng(21)
    m = nr.Model(rng.normal(0, 0.3, mo.N_PARAMS).astype(np.float32), alpha=1.0)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    r = nr.solve_learned(g, m, truncation_eps=0.0, queue_capacity=1 << 23)
    assert (r["paths"], r["nodes"], r["dead"]) == (7433032, 14866064, 7433032)
    R = closed_form(expected, 10, "0.5")
    assert abs(r["R"] - R) <= 1e-12 * R
    assert r["ESP"] == 0.0


def test_learned_order_bracket_under_deadline(expected):
    print("case learned order K15 deadline", flush=True)
    rng = np.random.default_rng(22)
    m = nr.Model(rng.normal(0, 0.3, mo.N_PARAMS).astype(np.float32), alpha=1.0)
    g = nr.Graph.load(graph_path("COMP15-05.txt"), 0, 14)
    r = nr.solve_learned(g, m, truncation_eps=1e-15, queue_capacity=1 << 24, time_budget_s=0.01)
    R = closed_form(expected, 15, "0.5")
    assert r["deadline_hit"] == 1
    assert r["R"] <= R * (1 + 1e-12) and r["R"] + r["ESP"] >= R * (1 - 1e-12)
    assert abs(r["R"] + r["ESP"] + r["D"] - 1.0) <= 1e-12


def test_learned_order_needs_a_model():
    g = nr.Graph.load(graph_path("COMP8-09.txt"), 0, 7)
    with pytest.raises(nr.NRError) as ei:
        nr.solve(g, order="learned")
    assert ei.value.status == 4


def _feature_multiset_equal(Xg, Yg, Xo, Yo):
    """Rows of the engine's tree log against the CPU restatement: the exact
    features (counts, ratios, q, flags) and the label must match exactly,
    the four log2-based ones (f0, f8, f9, f13: float32 log2 on either side,
    then bfloat16) within one bfloat16 step."""
    from features_oracle import bf16

    exact = [1, 2, 3, 4, 5, 6, 7, 10, 11, 12, 14, 15]
    logs = [0, 8, 9, 13]
    Xo = bf16(Xo)

    def groups(X, Y):
        d = {}
        for x, y in zip(X, Y):
            key = (tuple(x[exact].view(np.uint32).tolist()), round(float(y), 4))
            d.setdefault(key, []).append(x[logs].astype(np.float64))
        return d

    gg, go = groups(Xg, Yg), groups(Xo, Yo)
    assert gg.keys() == go.keys()
    for key in gg:
        a = np.array(sorted(map(tuple, gg[key])))
        b = np.array(sorted(map(tuple, go[key])))
        assert a.shape == b.shape, key
        assert np.all(np.abs(a - b) <= 2.0 ** -7 * np.maximum(np.abs(a), np.abs(b)) + 1e-6), key


@pytest.mark.parametrize("name,s,t,eps,cap", [("COMP8-09.txt", 0, 7, 0.0, 1 << 16),
                                              ("COMP8-03.txt", 0, 7, 1e-7, 1 << 18)])
def test_tree_log_features_match_restatement(name, s, t, eps, cap):
    """Every child's 16 logged features (the learned order's inputs) against
    oracle/features_oracle.py's restatement over the same tree."""
    from features_oracle import child_feature_rows

    print(f"case tree-log features {name} eps={eps}", flush=True)
    g = nr.Graph.load(graph_path(name), s, t)
    Xg, Yg, res = nr.log_subproblems(g, 1 << 20, seed=3, truncation_eps=eps, queue_capacity=cap)
    Xo, Yo, nodes = child_feature_rows(OracleGraph.load(graph_path(name)), s, t, eps)
    assert res["nodes"] == nodes and len(Yg) == len(Yo) == nodes - 1
    _feature_multiset_equal(Xg, Yg, Xo, Yo)


@pytest.mark.parametrize("name,s,t,eps,cap", [("COMP8-09.txt", 0, 7, 0.0, 1 << 16),
                                              ("COMP8-03.txt", 0, 7, 1e-7, 1 << 18),
                                              ("test_kite.txt", 0, 4, 0.0, 1 << 16)])
def test_tree_log_conditional_reliability(name, s, t, eps, cap):
    """Second label of the tree log (nr_log_subproblems_rel): every child's
    conditional reliability against the restatement's (subtree success mass
    over the child's mass, FP64), within 1e-13; the sample, features and
    size labels equal nr_log_subproblems'."""
    from features_oracle import bf16, child_rows_with_reliability

    print(f"case tree-log conditional reliability {name} eps={eps}", flush=True)
    g = nr.Graph.load(graph_path(name), s, t)
    Xg, Yg, Cg, res = nr.log_subproblems_rel(g, 1 << 20, seed=3, truncation_eps=eps,
                                             queue_capacity=cap)
    X1, Y1, _ = nr.log_subproblems(g, 1 << 20, seed=3, truncation_eps=eps, queue_capacity=cap)
    assert np.array_equal(Xg, X1) and np.array_equal(Yg, Y1)
    Xo, Yo, Co, _, nodes = child_rows_with_reliability(OracleGraph.load(graph_path(name)), s, t,
                                                       eps)
    assert res["nodes"] == nodes and len(Cg) == len(Co) == nodes - 1
    assert np.all((Cg >= 0.0) & (Cg <= 1.0))
    exact = [1, 2, 3, 4, 5, 6, 7, 10, 11, 12, 14, 15]
    Xo = bf16(Xo)

    def groups(X, Y, C):
        d = {}
        for x, y, c in zip(X, Y, C):
            d.setdefault((tuple(x[exact].view(np.uint32).tolist()), round(float(y), 4)),
                         []).append(float(c))
        return {k: np.sort(v) for k, v in d.items()}

    gg, go = groups(Xg, Yg, Cg), groups(Xo, Yo, Co)
    assert gg.keys() == go.keys()
    worst = max(float(np.abs(gg[k] - go[k]).max()) for k in gg)
    assert worst <= 1e-13, worst


def test_tree_log_conditional_reliability_k10_repeats():
    """K10 in full (14.9 M subproblems): the fixed-point labels repeat bit
    for bit and lie in [0, 1]; leaves carry exactly 0 or 1; the sample,
    features and size labels equal the plain log's."""
    print("case tree-log conditional reliability K10 repeat", flush=True)
    g = nr.Graph.load(graph_path("COMP10-05.txt"), 0, 9)
    kw = dict(truncation_eps=0.0, queue_capacity=1 << 22)
    X1, Y1, C1, r1 = nr.log_subproblems_rel(g, 50000, seed=7, **kw)
    X2, Y2, C2, r2 = nr.log_subproblems_rel(g, 50000, seed=7, **kw)
    X0, Y0, _ = nr.log_subproblems(g, 50000, seed=7, **kw)
    assert len(C1) == 50000 and r1["nodes"] == 14866064
    assert np.array_equal(C1, C2) and np.array_equal(X1, X0) and np.array_equal(Y1, Y0)
    assert np.all((C1 >= 0.0) & (C1 <= 1.0))
    # a leaf (subtree of one node) is a success path or a dead end: its
    # conditional reliability is exactly 1 or 0
    leaf = Y1 == 0.0
    assert leaf.any() and np.all((C1[leaf] == 0.0) | (C1[leaf] == 1.0))



def test_tree_log_conditional_reliability_alternating_rule():
    """The alternating rule's tree log carries the second label too (K9 in
    full): the sample and size labels equal nr_log_subproblems', the labels
    lie in [0, 1], leaves carry exactly 0 or 1, and a second run repeats
    them bit for bit."""
    print("case tree-log conditional reliability, alternating rule, K9", flush=True)
    g = nr.Graph.load(graph_path("COMP9.txt"), 0, 8)
    kw = dict(truncation_eps=0.0, queue_capacity=1 << 22, rule="alternate", sweep_period=2)
    X1, Y1, C1, r1 = nr.log_subproblems_rel(g, 200000, seed=11, **kw)
    X2, Y2, C2, _ = nr.log_subproblems_rel(g, 200000, seed=11, **kw)
    X0, Y0, _ = nr.log_subproblems(g, 200000, seed=11, **kw)
    assert len(C1) == min(200000, r1["nodes"] - 1)
    assert np.array_equal(X1, X0) and np.array_equal(Y1, Y0) and np.array_equal(C1, C2)
    assert np.all((C1 >= 0.0) & (C1 <= 1.0))
    leaf = Y1 == 0.0
    assert leaf.any() and np.all((C1[leaf] == 0.0) | (C1[leaf] == 1.0))
    assert 0.0 < C1.mean() < 1.0
