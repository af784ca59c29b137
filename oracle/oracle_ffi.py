"""ctypes bindings of the CPU oracle (oracle/liboracle_nr.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke entry point and the
benchmark's CPU-baseline leg, as the checker.  Never used by the product.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_nr.so")
CLI_PATH = os.path.join(HERE, "bin", "cpu_reference")

MODE_PATH, MODE_PIVOT, MODE_FORKBEST, MODE_MINUNK, MODE_ALTERNATE = 0, 1, 2, 3, 4
MODE_LEARNED, MODE_EXPLORE = 5, 6
ORDER_LEVEL, ORDER_HEAP, ORDER_DFS = 0, 1, 2


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_int),
        ("order", ctypes.c_int),
        ("threads", ctypes.c_int),
        ("trunc_eps", ctypes.c_double),
        ("capacity", ctypes.c_longlong),
        ("time_budget_s", ctypes.c_double),
        ("sweep_period", ctypes.c_int),
        ("chooser", ctypes.POINTER(ctypes.c_int)),
        ("prefix_edges", ctypes.c_int),
        ("shard", ctypes.c_int),
        ("shards", ctypes.c_int),
    ]


class OrcResult(ctypes.Structure):
    _fields_ = [
        ("R", ctypes.c_double),
        ("ESP", ctypes.c_double),
        ("D", ctypes.c_double),
        ("paths", ctypes.c_ulonglong),
        ("nodes", ctypes.c_ulonglong),
        ("dead", ctypes.c_ulonglong),
        ("truncated", ctypes.c_ulonglong),
        ("overflowed", ctypes.c_ulonglong),
        ("folded", ctypes.c_ulonglong),
        ("iterations", ctypes.c_longlong),
        ("max_width", ctypes.c_longlong),
        ("elapsed_ms", ctypes.c_double),
        ("deadline_hit", ctypes.c_int),
        ("n", ctypes.c_int),
        ("e", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        L.orc_graph_load.argtypes = [ctypes.c_char_p, P(vp)]
        L.orc_graph_from_csr.argtypes = [ctypes.c_int, ctypes.c_int, P(ctypes.c_int),
                                         P(ctypes.c_int), P(ctypes.c_int),
                                         P(ctypes.c_double), P(vp)]
        L.orc_graph_free.argtypes = [vp]
        L.orc_graph_free.restype = None
        L.orc_graph_dims.argtypes = [vp] + [P(ctypes.c_int)] * 5
        L.orc_graph_dims.restype = None
        L.orc_graph_arrays.argtypes = [vp, P(ctypes.c_int), P(ctypes.c_int),
                                       P(ctypes.c_int), P(ctypes.c_double)]
        L.orc_graph_arrays.restype = None
        L.orc_solve.argtypes = [vp, ctypes.c_int, ctypes.c_int, P(OrcConfig), P(OrcResult)]
        L.orc_solve_terms.argtypes = [vp, P(ctypes.c_int), ctypes.c_int, P(OrcConfig),
                                      P(OrcResult)]
        U = P(ctypes.c_ulonglong)
        L.orc_expand.argtypes = [vp, ctypes.c_int, ctypes.c_int, U, U, ctypes.c_double,
                                 ctypes.c_int, U, U, P(ctypes.c_double), P(ctypes.c_double),
                                 P(ctypes.c_int), P(ctypes.c_int)]
        L.orc_expand_rule.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, U, U, ctypes.c_double, ctypes.c_int, U, U,
                                      P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_int)]
        L.orc_expand_choice.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, P(ctypes.c_int), ctypes.c_double, U, U,
                                        ctypes.c_double, ctypes.c_int, U, U,
                                        P(ctypes.c_double), P(ctypes.c_double),
                                        P(ctypes.c_int)]
        L.orc_subtree_stats.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, U, U,
                                        P(ctypes.c_double), ctypes.c_double, U,
                                        P(ctypes.c_double)]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class OracleGraph:
    def __init__(self, h):
        self._h = ctypes.c_void_p(h)

    @classmethod
    def load(cls, path: str) -> "OracleGraph":
        h = ctypes.c_void_p()
        rc = lib().orc_graph_load(os.fsencode(path), ctypes.byref(h))
        if rc:
            raise RuntimeError(f"orc_graph_load({path}) failed: {rc}")
        return cls(h.value)

    @classmethod
    def from_csr(cls, row_ptr, col_idx, edge_id, p) -> "OracleGraph":
        rp = np.ascontiguousarray(row_ptr, np.int32)
        ci = np.ascontiguousarray(col_idx, np.int32)
        ei = np.ascontiguousarray(edge_id, np.int32)
        pp = np.ascontiguousarray(p, np.float64)
        h = ctypes.c_void_p()
        rc = lib().orc_graph_from_csr(len(rp) - 1, len(pp), _p(rp, ctypes.c_int),
                                      _p(ci, ctypes.c_int), _p(ei, ctypes.c_int),
                                      _p(pp, ctypes.c_double), ctypes.byref(h))
        if rc:
            raise RuntimeError(f"orc_graph_from_csr failed: {rc}")
        return cls(h.value)

    def dims(self) -> dict:
        v = [ctypes.c_int() for _ in range(5)]
        lib().orc_graph_dims(self._h, *[ctypes.byref(x) for x in v])
        return dict(zip(["n", "e", "nnz", "header_e", "type"], [x.value for x in v]))

    def arrays(self):
        d = self.dims()
        rp = np.zeros(d["n"] + 1, np.int32)
        ci = np.zeros(max(d["nnz"], 1), np.int32)
        ei = np.zeros(max(d["nnz"], 1), np.int32)
        p = np.zeros(max(d["e"], 1), np.float64)
        lib().orc_graph_arrays(self._h, _p(rp, ctypes.c_int), _p(ci, ctypes.c_int),
                               _p(ei, ctypes.c_int), _p(p, ctypes.c_double))
        return rp, ci[: d["nnz"]], ei[: d["nnz"]], p[: d["e"]]

    @property
    def mask_words(self) -> int:
        return max(1, (self.dims()["e"] + 63) // 64)

    def solve(self, src: int, dst: int, mode: int = MODE_PATH, order: int = ORDER_LEVEL,
              threads: int = 1, trunc_eps: float = 0.0, capacity: int = 0,
              time_budget_s: float = 0.0, sweep_period: int = 0, chooser=None,
              prefix_edges: int = 0, shard: int = 0, shards: int = 0) -> dict:
        """chooser (MODE_LEARNED): (weights [16, 64], signs [64], margin)
        integers, the exact form the engine reports for a model's path
        chooser (margin 0 when omitted)."""
        cfg = OrcConfig(mode, order, threads, trunc_eps, capacity, time_budget_s, sweep_period)
        cfg.prefix_edges, cfg.shard, cfg.shards = prefix_edges, shard, shards
        keep = None
        if chooser is not None:
            w, sg, margin = chooser if len(chooser) == 3 else (*chooser, 0)
            keep = np.ascontiguousarray(np.concatenate([np.asarray(w, np.int32).reshape(-1),
                                                        np.asarray(sg, np.int32).reshape(-1),
                                                        np.asarray([margin], np.int32)]))
            assert keep.size == 16 * 64 + 64 + 1
            cfg.chooser = _p(keep, ctypes.c_int)
        r = OrcResult()
        rc = lib().orc_solve(self._h, src, dst, ctypes.byref(cfg), ctypes.byref(r))
        if rc:
            raise RuntimeError(f"orc_solve failed: {rc}")
        return r.as_dict()

    def solve_terms(self, terms, mode: int = MODE_PATH, order: int = ORDER_LEVEL,
                    threads: int = 1, trunc_eps: float = 0.0, capacity: int = 0,
                    time_budget_s: float = 0.0) -> dict:
        """K-terminal reliability with terms[0] as the BFS root."""
        t = np.ascontiguousarray(terms, np.int32)
        cfg = OrcConfig(mode, order, threads, trunc_eps, capacity, time_budget_s)
        r = OrcResult()
        rc = lib().orc_solve_terms(self._h, _p(t, ctypes.c_int), len(t), ctypes.byref(cfg),
                                   ctypes.byref(r))
        if rc:
            raise RuntimeError(f"orc_solve_terms failed: {rc}")
        return r.as_dict()

    def expand(self, src: int, dst: int, known, up, mass: float):
        """One path-mode expansion: (k, succ, children_known, children_up,
        children_mass, path_eids); k = -1 when dead."""
        W = self.mask_words
        n = self.dims()["n"]
        known = np.ascontiguousarray(known, np.uint64).reshape(W)
        up = np.ascontiguousarray(up, np.uint64).reshape(W)
        mc = n + 1
        ck = np.zeros((mc, W), np.uint64)
        cu = np.zeros((mc, W), np.uint64)
        cm = np.zeros(mc, np.float64)
        succ = ctypes.c_double()
        nh = ctypes.c_int()
        path = np.zeros(n + 1, np.int32)
        U = ctypes.c_ulonglong
        k = lib().orc_expand(self._h, src, dst, _p(known, U), _p(up, U), float(mass), mc,
                             _p(ck, U), _p(cu, U), _p(cm, ctypes.c_double),
                             ctypes.byref(succ), ctypes.byref(nh), _p(path, ctypes.c_int))
        kk = max(k, 0)
        return k, succ.value, ck[:kk], cu[:kk], cm[:kk], path[: max(nh.value, 0)]

    def expand_choice(self, src: int, dst: int, known, up, mass: float, mode: int,
                      period: int = 0, chooser=None, eps: float = 0.0):
        """One expansion under MODE_ALTERNATE, MODE_LEARNED (chooser =
        (weights [16, 64], signs [64])) or MODE_EXPLORE: (k, hops in sweep
        order); k = -1 when dead."""
        W = self.mask_words
        n = self.dims()["n"]
        known = np.ascontiguousarray(known, np.uint64).reshape(W)
        up = np.ascontiguousarray(up, np.uint64).reshape(W)
        path = np.zeros(n + 1, np.int32)
        keep = None
        if chooser is not None:
            w, sg, margin = chooser if len(chooser) == 3 else (*chooser, 0)
            keep = np.ascontiguousarray(np.concatenate([np.asarray(w, np.int32).reshape(-1),
                                                        np.asarray(sg, np.int32).reshape(-1),
                                                        np.asarray([margin], np.int32)]))
        U = ctypes.c_ulonglong
        k = lib().orc_expand_choice(self._h, src, dst, mode, period,
                                    None if keep is None else _p(keep, ctypes.c_int), float(eps),
                                    _p(known, U), _p(up, U), float(mass), 0, None, None, None,
                                    None, _p(path, ctypes.c_int))
        if k < -1:
            raise RuntimeError("orc_expand_choice: bad arguments")
        return k, [int(h) for h in path[: max(k, 0)]]

    def expand_rule(self, src: int, dst: int, known, up, mass: float, mode: int = MODE_PATH,
                    period: int = 0):
        """One expansion under MODE_PATH or MODE_ALTERNATE: (k, succ,
        children_known, children_up, children_mass, hops in sweep order or
        None); k = -1 when dead."""
        W = self.mask_words
        n = self.dims()["n"]
        known = np.ascontiguousarray(known, np.uint64).reshape(W)
        up = np.ascontiguousarray(up, np.uint64).reshape(W)
        mc = n + 1
        ck = np.zeros((mc, W), np.uint64)
        cu = np.zeros((mc, W), np.uint64)
        cm = np.zeros(mc, np.float64)
        succ = ctypes.c_double()
        path = np.zeros(n + 1, np.int32)
        U = ctypes.c_ulonglong
        k = lib().orc_expand_rule(self._h, src, dst, mode, period, _p(known, U), _p(up, U),
                                  float(mass), mc, _p(ck, U), _p(cu, U),
                                  _p(cm, ctypes.c_double), ctypes.byref(succ),
                                  _p(path, ctypes.c_int))
        if k < -1:
            raise RuntimeError("orc_expand_rule: bad arguments")
        kk = max(k, 0)
        hops = path[:kk] if mode == MODE_ALTERNATE else None
        return k, succ.value, ck[:kk], cu[:kk], cm[:kk], hops

    def subtree_stats(self, src, dst, known, up, mass, trunc_eps=0.0):
        W = self.mask_words
        known = np.ascontiguousarray(known, np.uint64).reshape(-1, W)
        up = np.ascontiguousarray(up, np.uint64).reshape(-1, W)
        mass = np.ascontiguousarray(mass, np.float64).reshape(-1)
        n = len(mass)
        nodes = np.zeros(n, np.uint64)
        r = np.zeros(n, np.float64)
        U = ctypes.c_ulonglong
        rc = lib().orc_subtree_stats(self._h, src, dst, n, _p(known, U), _p(up, U),
                                     _p(mass, ctypes.c_double), trunc_eps, _p(nodes, U),
                                     _p(r, ctypes.c_double))
        if rc:
            raise RuntimeError(f"orc_subtree_stats failed: {rc}")
        return nodes, r

    def close(self):
        if self._h:
            lib().orc_graph_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
