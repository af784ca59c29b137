"""CPU restatement of the learned order's child features (TEST INFRASTRUCTURE
ONLY: tests check the tree log the engine writes on the GPU against it; the
engine never imports it).

The engine's learned order scores every child it creates on 16 features
taken from the parent's expansion (relaibilitygpu_amd/csrc/nr_learn.h,
child_features; no counterpart in the reference).  This module walks the
path-rule tree with the oracle's own expansions (oracle_ffi.OracleGraph
.expand: main.cu:214-266 restated) and computes, for each child t of an item
(t counted from the src side), in float32 as the engine does:

  f0  log2(child mass) / 64          f8  (sum of log2 p over the parent's
  f1  (u - t - 1) / E                    UNKNOWN hops after hop t) / 16
  f2  (w + t) / E                    f9  headroom: (log2 cm - log2 eps) / 64
  f3  (f + 1) / E                        clamped to [0, 1] (1 without eps)
  f4  (t + 1) / k                    f10 (allowed degree of src, less the
  f5  k / n                              hop if it touches src) / 16, <= 1
  f6  path hops / n                  f11 the same for dst
  f7  q of the hop edge              f12 1 if the hop touches src or dst
  f13 log2(parent mass) / 64         f14 (k - t - 1) / n, f15 = 1

where u, w, f are the parent's UNKNOWN, WORKING and FAILED edge counts, E
the edge count, n the node count and k the parent's UNKNOWN hops; each
value is then rounded to bfloat16 (round to nearest even) as the log stores
it.  The label of a child is log2 of its subtree's node count; its second
label (child_rows_with_reliability) is its conditional reliability: the
success mass its subtree finds (main.cu:254-262, summed over the subtree)
divided by its own mass, which for an untruncated subtree is the
probability that src reaches dst given the child's WORKING and FAILED
edges.

Pinned by identities the definition fixes, checked on the restatement
itself (tests/test_oracle.py): f1 + f2 + f3 = 1 (every edge is UNKNOWN,
WORKING or FAILED in the child), f4 = 1 and f14 = 0 for the last child,
f0 <= f13 (a child's mass is its parent's times a probability), one row per
non-root node, and labels that are log2 of whole subtree sizes.  The
conditional reliabilities of an untruncated tree are pinned against brute
force over the child's UNKNOWN edges (tests/test_oracle.py).
"""
from __future__ import annotations

import math

import numpy as np

N_FEATURES = 16


def bf16(x) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    a = np.asarray(x, np.float32).reshape(-1).copy()
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u << 16).astype(np.uint32).view(np.float32).reshape(np.shape(x))


def _log2(x: float) -> np.float32:
    """log2 of a positive double as the engine forms it: exponent plus
    log2 of the mantissa in float32."""
    if not x > 0.0:
        return np.float32(-1100.0)
    m, e = math.frexp(x)  # x = m 2^e, m in [0.5, 1)
    return np.float32(e - 1) + np.float32(np.log2(np.float32(2.0 * m)))


def _popcount(words) -> int:
    return int(sum(bin(int(w)).count("1") for w in words))


def child_feature_rows(og, src: int, dst: int, eps: float = 0.0):
    """(features [m, 16] float32 before rounding, labels [m] float64, node
    count) of every child of the path-rule tree rooted at the all-UNKNOWN
    item."""
    X, Y, _, _, nodes = child_rows_with_reliability(og, src, dst, eps)
    return X, Y, nodes


def child_rows_with_reliability(og, src: int, dst: int, eps: float = 0.0):
    """(features, log2-size labels, conditional reliabilities [m] float64,
    the children's (known, up) planes [m, 2, W] uint64, node count) of every
    child of the path-rule tree rooted at the all-UNKNOWN item."""
    rp, ci, ei, p = og.arrays()
    n = len(rp) - 1
    E = len(p)
    W = og.mask_words
    ea = np.full(E, -1, np.int64)
    eb = np.full(E, -1, np.int64)
    for u in range(n):
        for j in range(rp[u], rp[u + 1]):
            e = ei[j]
            a, b = min(u, ci[j]), max(u, ci[j])
            ea[e], eb[e] = a, b
    q = 1.0 - p
    inv_e = np.float32(1.0) / np.float32(max(E, 1))
    inv_n = np.float32(1.0) / np.float32(n)
    lp32 = [np.float32(np.log2(np.float32(x))) for x in p]

    def failed(known, up, e):
        return bool((int(known[e >> 6]) >> (e & 63)) & 1) and not (
            (int(up[e >> 6]) >> (e & 63)) & 1)

    def allowed_degree(known, up, v):
        d = 0
        for j in range(rp[v], rp[v + 1]):
            if ci[j] != v and not failed(known, up, ei[j]):
                d += 1
        return d

    rows, labels, conds, planes = [], [], [], []

    def walk(known, up, mass):
        """Expands the item, records its children, returns its subtree size
        and the success mass its subtree finds."""
        k, succ, ck, cu, cm, path = og.expand(src, dst, known, up, mass)
        if k < 0:
            return 1, 0.0
        if k == 0:
            return 1, succ
        nk = _popcount(known)
        u_p, w_p, f_p = E - nk, _popcount(up), nk - _popcount(up)
        hops = [int(e) for e in path if not ((int(known[e >> 6]) >> (e & 63)) & 1)]
        assert len(hops) == k
        lp_total = np.float32(0.0)
        for e in reversed(hops):  # the engine adds them dst side first
            lp_total = np.float32(lp_total + lp32[e])
        deg_s = allowed_degree(known, up, src)
        deg_d = allowed_degree(known, up, dst)
        log2_m = _log2(mass)
        inv_k = np.float32(1.0) / np.float32(k)
        lp_prefix = np.float32(0.0)
        first = len(rows)
        for t, e in enumerate(hops):
            lcm = _log2(float(cm[t]))
            at_s = ea[e] == src or eb[e] == src
            at_d = ea[e] == dst or eb[e] == dst
            head = np.float32(1.0)
            if eps > 0.0:
                head = np.float32(min(max((lcm - _log2(eps)) * np.float32(1.0 / 64.0),
                                          np.float32(0.0)), np.float32(1.0)))
            f = [lcm * np.float32(1.0 / 64.0),
                 np.float32(u_p - t - 1) * inv_e,
                 np.float32(w_p + t) * inv_e,
                 np.float32(f_p + 1) * inv_e,
                 np.float32(t + 1) * inv_k,
                 np.float32(k) * inv_n,
                 np.float32(len(path)) * inv_n,
                 np.float32(q[e]),
                 np.float32(np.float32(lp_total - lp_prefix) - lp32[e]) * np.float32(1.0 / 16.0),
                 head,
                 np.float32(min(deg_s - (1 if at_s else 0), 16)) * np.float32(1.0 / 16.0),
                 np.float32(min(deg_d - (1 if at_d else 0), 16)) * np.float32(1.0 / 16.0),
                 np.float32(1.0 if (at_s or at_d) else 0.0),
                 log2_m * np.float32(1.0 / 64.0),
                 np.float32(k - t - 1) * inv_n,
                 np.float32(1.0)]
            rows.append(np.array(f, np.float32))
            labels.append(0.0)
            conds.append(0.0)
            planes.append(np.stack([ck[t], cu[t]]))
            lp_prefix = np.float32(lp_prefix + lp32[e])
        total, found = 1, succ
        for t in range(k):
            sz, r = walk(ck[t], cu[t], float(cm[t])) if cm[t] >= eps else (1, 0.0)
            labels[first + t] = math.log2(sz)
            conds[first + t] = r / float(cm[t]) if cm[t] > 0.0 else 0.0
            total += sz
            found += r
        return total, found

    z = np.zeros(W, np.uint64)
    nodes, _ = walk(z, z, 1.0)
    return (np.array(rows, np.float32).reshape(-1, N_FEATURES), np.array(labels),
            np.array(conds), np.array(planes, np.uint64).reshape(-1, 2, W), nodes)
