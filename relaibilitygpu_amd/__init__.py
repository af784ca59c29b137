"""relaibilitygpu_amd - MI355X (gfx950) engine for two-terminal network
reliability by path decomposition.

Python mirror of the C-ABI in ``include/nr_api.h`` (ctypes over the in-tree
``libnrgpu.so``).  The product path is the HIP library: if it has not been
built, importing the bindings raises instead of falling back to anything.

Reference surface mirrored here:
  * ``load_graph``   - workingPad/graph.h:46-145 ``graph_load``
  * ``solve``        - workingPad/main.cu:795-1056 (the whole run, report
                       fields of main.cu:1037-1056)
  * ``expand_items`` - workingPad/main.cu:114-283 on a batch of items
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnrgpu.so")
CLI_PATH = os.path.join(_HERE, "bin", "network_reliability")

NR_FLAG_TIME_LAUNCHES = 1
NR_FLAG_ROWS = 4
NR_FLAG_LEAN = 8
NR_FLAG_QUEUE_ALL = 16

NR_ORDER_LEVEL = 0
NR_ORDER_MASS = 1
NR_ORDER_LEARNED = 2
ORDERS = {"level": NR_ORDER_LEVEL, "mass": NR_ORDER_MASS, "learned": NR_ORDER_LEARNED}
NR_RULE_PATH = 0
NR_RULE_FORKBEST = 1
NR_RULE_ALTERNATE = 2
NR_RULE_PIVOT = 3
NR_RULE_LEARNED = 4
RULES = {"path": NR_RULE_PATH, "forkbest": NR_RULE_FORKBEST, "alternate": NR_RULE_ALTERNATE,
         "pivot": NR_RULE_PIVOT, "learned": NR_RULE_LEARNED}

STATUS = {
    0: "NR_OK",
    1: "NR_ERR_IO",
    2: "NR_ERR_PARSE",
    3: "NR_ERR_RANGE",
    4: "NR_ERR_ARG",
    5: "NR_ERR_UNSUPPORTED",
    6: "NR_ERR_DEVICE",
    7: "NR_ERR_NOMEM",
    8: "NR_ERR_SCRATCH",
}
STATIC_SCRATCH_BYTES = 64 << 20


class NRError(RuntimeError):
    """A non-zero nr_status returned by the library."""

    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: {STATUS.get(status, status)}")


class Config(ctypes.Structure):
    _fields_ = [
        ("queue_capacity", ctypes.c_longlong),
        ("truncation_eps", ctypes.c_double),
        ("time_budget_s", ctypes.c_double),
        ("order", ctypes.c_int),
        ("device", ctypes.c_int),
        ("block_size", ctypes.c_int),
        ("grid_blocks", ctypes.c_int),
        ("chunk_items", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("rule", ctypes.c_int),
        ("sweep_period", ctypes.c_int),
        ("prefix_edges", ctypes.c_int),
        ("shard", ctypes.c_int),
        ("shards", ctypes.c_int),
        ("ramp_items", ctypes.c_int),
    ]


class Result(ctypes.Structure):
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
        ("max_queue_items", ctypes.c_longlong),
        ("item_bytes", ctypes.c_longlong),
        ("elapsed_ms", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("kernel_launches", ctypes.c_longlong),
        ("deadline_hit", ctypes.c_int),
        ("overflowed_flag", ctypes.c_int),
        ("grid_blocks", ctypes.c_int),
        ("block_size", ctypes.c_int),
        ("lds_bytes", ctypes.c_int),
        ("blocks_per_cu", ctypes.c_int),
        ("compute_units", ctypes.c_int),
        ("rerun", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class GraphInfo(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("e", ctypes.c_int),
        ("nnz", ctypes.c_int),
        ("header_e", ctypes.c_int),
        ("type", ctypes.c_int),
        ("src", ctypes.c_int),
        ("dst", ctypes.c_int),
        ("mask_words", ctypes.c_int),
        ("symmetric", ctypes.c_int),
        ("precontracted", ctypes.c_int),
        ("predeleted", ctypes.c_int),
    ]


class TrainConfig(ctypes.Structure):
    _fields_ = [
        ("epochs", ctypes.c_int),
        ("batch", ctypes.c_int),
        ("lr", ctypes.c_float),
        ("weight_decay", ctypes.c_float),
        ("seed", ctypes.c_ulonglong),
        ("device", ctypes.c_int),
        ("holdout", ctypes.c_float),
        ("reserved", ctypes.c_int * 4),
    ]


MLP_FEATURES = 16
MLP_HIDDEN = 64
MLP_PARAMS = 16 * 64 + 64 + 1
CHOICE_FEATURES = 16

_lib = None


def lib() -> ctypes.CDLL:
    """Load libnrgpu.so (built in-tree by ``make -C relaibilitygpu_amd``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP engine first "
            "(python -c 'import __graft_entry__ as g; g.build()')"
        )
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    L.nr_config_init.argtypes = [P(Config)]
    L.nr_graph_load.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, P(vp)]
    L.nr_graph_create.argtypes = [
        ctypes.c_int, ctypes.c_int, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int),
        P(ctypes.c_double), ctypes.c_int, ctypes.c_int, P(vp),
    ]
    L.nr_graph_get_info.argtypes = [vp, P(GraphInfo)]
    L.nr_graph_get_csr.argtypes = [vp, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int),
                                   P(ctypes.c_double)]
    L.nr_graph_set_terminals.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.nr_graph_set_terminal_set.argtypes = [vp, P(ctypes.c_int), ctypes.c_int]
    L.nr_graph_get_terminal_set.argtypes = [vp, P(ctypes.c_int), ctypes.c_int, P(ctypes.c_int)]
    L.nr_graph_destroy.argtypes = [vp]
    L.nr_graph_destroy.restype = None
    L.nr_solve.argtypes = [vp, P(Config), P(Result)]
    L.nr_solve_static.argtypes = [vp, P(Config), P(Result)]
    L.nr_solve_batch.argtypes = [P(vp), ctypes.c_int, P(Config), P(Result)]
    L.nr_expand_items.argtypes = [
        vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_ulonglong), P(ctypes.c_ulonglong),
        P(ctypes.c_double), ctypes.c_int, P(ctypes.c_int), P(ctypes.c_double),
        P(ctypes.c_ulonglong), P(ctypes.c_ulonglong), P(ctypes.c_double),
    ]
    L.nr_expand_items_rule.argtypes = [
        vp, P(Config), ctypes.c_int, P(ctypes.c_ulonglong), P(ctypes.c_ulonglong),
        P(ctypes.c_double), ctypes.c_int, P(ctypes.c_int), P(ctypes.c_double),
        P(ctypes.c_ulonglong), P(ctypes.c_ulonglong), P(ctypes.c_double),
    ]
    L.nr_tune_batch.argtypes = [P(vp), ctypes.c_int, P(Config), ctypes.c_int, P(Config),
                                P(ctypes.c_double)]
    L.nr_tune_sweep_period.argtypes = [vp, P(Config), ctypes.c_double, ctypes.c_int,
                                       P(ctypes.c_int), P(ctypes.c_double)]
    L.nr_status_str.argtypes = [ctypes.c_int]
    L.nr_status_str.restype = ctypes.c_char_p
    L.nr_device_query.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int,
                                  P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_longlong)]
    fp = P(ctypes.c_float)
    L.nr_log_subproblems.argtypes = [vp, P(Config), ctypes.c_longlong, ctypes.c_ulonglong, fp,
                                     fp, P(ctypes.c_longlong), P(Result)]
    L.nr_log_subproblems_rel.argtypes = [vp, P(Config), ctypes.c_longlong, ctypes.c_ulonglong,
                                         fp, fp, P(ctypes.c_double), P(ctypes.c_longlong),
                                         P(Result)]
    L.nr_train_config_init.argtypes = [P(TrainConfig)]
    L.nr_mlp_init.argtypes = [ctypes.c_ulonglong, fp]
    L.nr_mlp_grad.argtypes = [fp, fp, ctypes.c_longlong, fp, fp, P(ctypes.c_double), ctypes.c_int]
    L.nr_mlp_train.argtypes = [fp, fp, ctypes.c_longlong, P(TrainConfig), fp, fp, fp]
    L.nr_model_create.argtypes = [fp, ctypes.c_float, P(vp)]
    L.nr_model_destroy.argtypes = [vp]
    L.nr_model_destroy.restype = None
    L.nr_model_score.argtypes = [vp, fp, ctypes.c_longlong, fp, ctypes.c_int]
    L.nr_solve_learned.argtypes = [vp, P(Config), vp, P(Result)]
    L.nr_model_set_chooser.argtypes = [vp, fp, ctypes.c_float]
    L.nr_model_get_chooser.argtypes = [vp, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int)]
    L.nr_log_choices.argtypes = [vp, P(Config), ctypes.c_longlong, ctypes.c_ulonglong, fp, fp,
                                 P(ctypes.c_longlong), P(Result)]
    L.nr_solve_checkpointed.argtypes = [vp, P(Config), vp, P(vp), P(Result)]
    L.nr_checkpoint_info.argtypes = [vp, P(ctypes.c_longlong), P(ctypes.c_double),
                                     P(ctypes.c_longlong)]
    L.nr_checkpoint_save.argtypes = [vp, ctypes.c_char_p]
    L.nr_checkpoint_load.argtypes = [ctypes.c_char_p, P(vp)]
    L.nr_checkpoint_destroy.argtypes = [vp]
    L.nr_checkpoint_destroy.restype = None
    _lib = L
    return L


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise NRError(rc, what)


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def make_config(queue_capacity: int = 1 << 20, truncation_eps: float = 1e-15,
                time_budget_s: float = 0.0, order: str | int = "level", device: int = 0,
                block_size: int = 0, grid_blocks: int = 0, chunk_items: int = 0,
                time_launches: bool = False, rule: str | int = "path",
                rows: bool = False, sweep_period: int = 0, lean: bool = False,
                prefix_edges: int = 0, shard: int = 0, shards: int = 0,
                ramp_items: int = 0, queue_all: bool = False) -> Config:
    c = Config()
    lib().nr_config_init(ctypes.byref(c))
    c.queue_capacity = int(queue_capacity)
    c.truncation_eps = float(truncation_eps)
    c.time_budget_s = float(time_budget_s)
    c.order = ORDERS[order] if isinstance(order, str) else int(order)
    c.device = int(device)
    c.block_size = int(block_size)
    c.grid_blocks = int(grid_blocks)
    c.chunk_items = int(chunk_items)
    c.flags = ((NR_FLAG_TIME_LAUNCHES if time_launches else 0) | (NR_FLAG_ROWS if rows else 0)
               | (NR_FLAG_LEAN if lean else 0) | (NR_FLAG_QUEUE_ALL if queue_all else 0))
    c.rule = RULES[rule] if isinstance(rule, str) else int(rule)
    c.sweep_period = int(sweep_period)
    c.prefix_edges = int(prefix_edges)
    c.shard = int(shard)
    c.shards = int(shards)
    c.ramp_items = int(ramp_items)
    return c


class Graph:
    """Owned handle to an nr_graph (host CSR + FP64 probabilities)."""

    def __init__(self, handle: int):
        self._h = ctypes.c_void_p(handle)

    @classmethod
    def load(cls, path: str, src: int, dst: int) -> "Graph":
        h = ctypes.c_void_p()
        _check(lib().nr_graph_load(os.fsencode(path), int(src), int(dst), ctypes.byref(h)),
               f"nr_graph_load({path})")
        return cls(h.value)

    @classmethod
    def from_csr(cls, row_ptr, col_idx, edge_id, p, src: int, dst: int) -> "Graph":
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(col_idx, dtype=np.int32)
        ei = np.ascontiguousarray(edge_id, dtype=np.int32)
        pp = np.ascontiguousarray(p, dtype=np.float64)
        h = ctypes.c_void_p()
        _check(lib().nr_graph_create(len(rp) - 1, len(pp), _ptr(rp, ctypes.c_int),
                                     _ptr(ci, ctypes.c_int), _ptr(ei, ctypes.c_int),
                                     _ptr(pp, ctypes.c_double), int(src), int(dst),
                                     ctypes.byref(h)), "nr_graph_create")
        return cls(h.value)

    @classmethod
    def from_edges(cls, n: int, edges, probs, src: int, dst: int) -> "Graph":
        """Undirected edge list -> CSR with both directions, rows sorted by
        neighbour, ids in the order given (first appearance)."""
        rows = [[] for _ in range(n)]
        for i, (a, b) in enumerate(edges):
            rows[a].append((b, i))
            if a != b:
                rows[b].append((a, i))
        rp, ci, ei = [0], [], []
        for u in range(n):
            for v, i in sorted(rows[u]):
                ci.append(v)
                ei.append(i)
            rp.append(len(ci))
        return cls.from_csr(rp, ci, ei, probs, src, dst)

    def info(self) -> GraphInfo:
        gi = GraphInfo()
        _check(lib().nr_graph_get_info(self._h, ctypes.byref(gi)), "nr_graph_get_info")
        return gi

    def csr(self):
        gi = self.info()
        rp = np.zeros(gi.n + 1, np.int32)
        ci = np.zeros(max(gi.nnz, 1), np.int32)
        ei = np.zeros(max(gi.nnz, 1), np.int32)
        p = np.zeros(max(gi.e, 1), np.float64)
        _check(lib().nr_graph_get_csr(self._h, _ptr(rp, ctypes.c_int), _ptr(ci, ctypes.c_int),
                                      _ptr(ei, ctypes.c_int), _ptr(p, ctypes.c_double)),
               "nr_graph_get_csr")
        return rp, ci[: gi.nnz], ei[: gi.nnz], p[: gi.e]

    def set_terminals(self, src: int, dst: int) -> None:
        _check(lib().nr_graph_set_terminals(self._h, int(src), int(dst)), "set_terminals")

    def set_terminal_set(self, terminals) -> None:
        """K-terminal problem rooted at terminals[0] (empty: back to src/dst)."""
        t = np.ascontiguousarray(list(terminals), dtype=np.int32)
        _check(lib().nr_graph_set_terminal_set(self._h, _ptr(t, ctypes.c_int), len(t)),
               "nr_graph_set_terminal_set")

    def terminal_set(self) -> list:
        k = ctypes.c_int()
        t = np.zeros(256, np.int32)
        _check(lib().nr_graph_get_terminal_set(self._h, _ptr(t, ctypes.c_int), 256,
                                               ctypes.byref(k)), "nr_graph_get_terminal_set")
        return [int(x) for x in t[: k.value]]

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            lib().nr_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve(graph: Graph, config: Config | None = None, **kw) -> dict:
    """Run one solve on the GPU and return the report fields as a dict."""
    cfg = config if config is not None else make_config(**kw)
    r = Result()
    _check(lib().nr_solve(graph.handle, ctypes.byref(cfg), ctypes.byref(r)), "nr_solve")
    return r.as_dict()


def solve_static(graph: Graph, config: Config | None = None, **kw) -> dict:
    """nr_solve_static: nr_solve with every device buffer in the library's
    fixed scratch array (no device allocation)."""
    cfg = config if config is not None else make_config(**kw)
    r = Result()
    _check(lib().nr_solve_static(graph.handle, ctypes.byref(cfg), ctypes.byref(r)),
           "nr_solve_static")
    return r.as_dict()


def tune_sweep_period(graph: Graph, pilot_eps: float = 1e-9, max_period: int = 6,
                      config: Config | None = None, **kw) -> tuple[int, list[float]]:
    """(period, pilot ESP per period 1..max_period): nr_tune_sweep_period."""
    cfg = config if config is not None else make_config(**kw)
    best = ctypes.c_int()
    esp = np.zeros(max_period, np.float64)
    _check(lib().nr_tune_sweep_period(graph.handle, ctypes.byref(cfg), float(pilot_eps),
                                      int(max_period), ctypes.byref(best),
                                      _ptr(esp, ctypes.c_double)), "nr_tune_sweep_period")
    return best.value, [float(x) for x in esp]


def tune_batch(graphs, sample: int = 0, config: Config | None = None, **kw):
    """(tuned Config, subproblems/s of the best trial): nr_tune_batch."""
    cfg = config if config is not None else make_config(**kw)
    hs = (ctypes.c_void_p * len(graphs))(*[g.handle.value for g in graphs])
    out = Config()
    rate = ctypes.c_double()
    _check(lib().nr_tune_batch(hs, len(graphs), ctypes.byref(cfg), int(sample), ctypes.byref(out),
                               ctypes.byref(rate)), "nr_tune_batch")
    return out, rate.value


class Batch:
    """A fixed list of instances solved repeatedly (nr_solve_batch): the
    handle array and result buffer are built once, the library keeps the
    instances' tables on the device between calls, and run() returns the
    results as a numpy structured array (fields of nr_result)."""

    def __init__(self, graphs, config: Config | None = None, **kw):
        self.graphs = list(graphs)
        self.config = config if config is not None else make_config(**kw)
        self._arr = (ctypes.c_void_p * len(self.graphs))(*[g.handle.value for g in self.graphs])
        self._res = (Result * len(self.graphs))()

    def run(self) -> np.ndarray:
        _check(lib().nr_solve_batch(self._arr, len(self.graphs), ctypes.byref(self.config),
                                    self._res), "nr_solve_batch")
        return np.ctypeslib.as_array(self._res).copy()


def solve_batch(graphs, config: Config | None = None, **kw) -> list[dict]:
    cfg = config if config is not None else make_config(**kw)
    arr = (ctypes.c_void_p * len(graphs))(*[g.handle.value for g in graphs])
    res = (Result * len(graphs))()
    _check(lib().nr_solve_batch(arr, len(graphs), ctypes.byref(cfg), res), "nr_solve_batch")
    return [r.as_dict() for r in res]


@dataclass
class Expansion:
    nchild: np.ndarray  # [count] (-1 dead)
    succ: np.ndarray  # [count]
    child_known: np.ndarray  # [count, max_children, W]
    child_up: np.ndarray
    child_mass: np.ndarray  # [count, max_children]


def expand_items(graph: Graph, known: np.ndarray, up: np.ndarray, mass: np.ndarray,
                 max_children: int, device: int = 0, rule: str | int = "path",
                 sweep_period: int = 0) -> Expansion:
    """One GPU expansion of each item (nr_expand_items; nr_expand_items_rule
    for another rule): children in sweep order."""
    W = graph.info().mask_words
    known = np.ascontiguousarray(known, dtype=np.uint64).reshape(-1, W)
    up = np.ascontiguousarray(up, dtype=np.uint64).reshape(-1, W)
    mass = np.ascontiguousarray(mass, dtype=np.float64).reshape(-1)
    n = len(mass)
    nchild = np.zeros(n, np.int32)
    succ = np.zeros(n, np.float64)
    ck = np.zeros((n, max_children, W), np.uint64)
    cu = np.zeros((n, max_children, W), np.uint64)
    cm = np.zeros((n, max_children), np.float64)
    U = ctypes.c_ulonglong
    if rule in ("path", NR_RULE_PATH) and not sweep_period:
        _check(lib().nr_expand_items(graph.handle, int(device), n, _ptr(known, U), _ptr(up, U),
                                     _ptr(mass, ctypes.c_double), int(max_children),
                                     _ptr(nchild, ctypes.c_int), _ptr(succ, ctypes.c_double),
                                     _ptr(ck, U), _ptr(cu, U), _ptr(cm, ctypes.c_double)),
               "nr_expand_items")
    else:
        cfg = make_config(device=device, rule=rule, sweep_period=sweep_period)
        _check(lib().nr_expand_items_rule(graph.handle, ctypes.byref(cfg), n, _ptr(known, U),
                                          _ptr(up, U), _ptr(mass, ctypes.c_double),
                                          int(max_children), _ptr(nchild, ctypes.c_int),
                                          _ptr(succ, ctypes.c_double), _ptr(ck, U),
                                          _ptr(cu, U), _ptr(cm, ctypes.c_double)),
               "nr_expand_items_rule")
    return Expansion(nchild, succ, ck, cu, cm)


def device_query(device: int = 0) -> dict:
    name = ctypes.create_string_buffer(256)
    cus, ws, mem = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
    _check(lib().nr_device_query(device, name, 256, ctypes.byref(cus), ctypes.byref(ws),
                                 ctypes.byref(mem)), "nr_device_query")
    return {"name": name.value.decode(), "compute_units": cus.value,
            "wave_size": ws.value, "total_mem": mem.value}


# ---------------------------------------------------------------- learned order
def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if n is not None:
        a = a.reshape(n, -1) if a.ndim > 1 else a
    return a


def log_subproblems(graph: Graph, max_records: int, seed: int = 1,
                    config: Config | None = None, **kw):
    """Tree log of a level-order solve: (features [m,16], labels [m] =
    log2 subtree nodes, result dict) for m <= max_records sampled children."""
    cfg = config if config is not None else make_config(**kw)
    X = np.zeros((max(max_records, 1), MLP_FEATURES), np.float32)
    Y = np.zeros(max(max_records, 1), np.float32)
    n = ctypes.c_longlong()
    res = Result()
    _check(lib().nr_log_subproblems(graph.handle, ctypes.byref(cfg), int(max_records), int(seed),
                                    _ptr(X, ctypes.c_float), _ptr(Y, ctypes.c_float),
                                    ctypes.byref(n), ctypes.byref(res)), "nr_log_subproblems")
    return X[: n.value], Y[: n.value], res.as_dict()


def log_subproblems_rel(graph: Graph, max_records: int, seed: int = 1,
                        config: Config | None = None, **kw):
    """Tree log with both labels (nr_log_subproblems_rel): (features [m,16],
    log2 subtree nodes [m], conditional reliabilities [m] float64, result
    dict); the sample is the one log_subproblems returns."""
    cfg = config if config is not None else make_config(**kw)
    X = np.zeros((max(max_records, 1), MLP_FEATURES), np.float32)
    Y = np.zeros(max(max_records, 1), np.float32)
    C = np.zeros(max(max_records, 1), np.float64)
    n = ctypes.c_longlong()
    res = Result()
    _check(lib().nr_log_subproblems_rel(graph.handle, ctypes.byref(cfg), int(max_records),
                                        int(seed), _ptr(X, ctypes.c_float),
                                        _ptr(Y, ctypes.c_float), _ptr(C, ctypes.c_double),
                                        ctypes.byref(n), ctypes.byref(res)),
           "nr_log_subproblems_rel")
    return X[: n.value], Y[: n.value], C[: n.value], res.as_dict()


def log_choices(graph: Graph, max_records: int, seed: int = 1,
                config: Config | None = None, **kw):
    """Choice log (nr_log_choices): (features [m,16] of the sweep direction
    taken, labels [m] = log2 subtree nodes, result dict) for m <= max_records
    subproblems of an exploring solve of the alternating rule's path."""
    cfg = config if config is not None else make_config(**kw)
    X = np.zeros((max(max_records, 1), CHOICE_FEATURES), np.float32)
    Y = np.zeros(max(max_records, 1), np.float32)
    n = ctypes.c_longlong()
    res = Result()
    _check(lib().nr_log_choices(graph.handle, ctypes.byref(cfg), int(max_records), int(seed),
                                _ptr(X, ctypes.c_float), _ptr(Y, ctypes.c_float),
                                ctypes.byref(n), ctypes.byref(res)), "nr_log_choices")
    return X[: n.value], Y[: n.value], res.as_dict()


def mlp_init(seed: int = 1) -> np.ndarray:
    p = np.zeros(MLP_PARAMS, np.float32)
    _check(lib().nr_mlp_init(int(seed), _ptr(p, ctypes.c_float)), "nr_mlp_init")
    return p


def mlp_grad(X, y, params, device: int = 0):
    X = _f32(X).reshape(-1, MLP_FEATURES)
    y = _f32(y).reshape(-1)
    p = _f32(params).reshape(MLP_PARAMS)
    g = np.zeros(MLP_PARAMS, np.float32)
    loss = ctypes.c_double()
    _check(lib().nr_mlp_grad(_ptr(X, ctypes.c_float), _ptr(y, ctypes.c_float), len(y),
                             _ptr(p, ctypes.c_float), _ptr(g, ctypes.c_float),
                             ctypes.byref(loss), int(device)), "nr_mlp_grad")
    return loss.value, g


def mlp_train(X, y, params=None, epochs: int = 8, batch: int = 4096, lr: float = 3e-3,
              weight_decay: float = 1e-5, seed: int = 1, holdout: float = 0.1, device: int = 0):
    """Train on the GPU; returns (params, train_loss[epochs], holdout_loss[epochs])."""
    X = _f32(X).reshape(-1, MLP_FEATURES)
    y = _f32(y).reshape(-1)
    p = (mlp_init(seed) if params is None else _f32(params).reshape(MLP_PARAMS)).copy()
    c = TrainConfig()
    lib().nr_train_config_init(ctypes.byref(c))
    c.epochs, c.batch, c.lr, c.weight_decay = int(epochs), int(batch), float(lr), float(weight_decay)
    c.seed, c.device, c.holdout = int(seed), int(device), float(holdout)
    tl = np.zeros(epochs, np.float32)
    hl = np.zeros(epochs, np.float32)
    _check(lib().nr_mlp_train(_ptr(X, ctypes.c_float), _ptr(y, ctypes.c_float), len(y),
                              ctypes.byref(c), _ptr(p, ctypes.c_float), _ptr(tl, ctypes.c_float),
                              _ptr(hl, ctypes.c_float)), "nr_mlp_train")
    return p, tl, hl


class Model:
    """Inference form of trained parameters (nr_model): the order scorer
    (params, may be None) and optionally a path chooser (chooser params)."""

    def __init__(self, params=None, alpha: float = 1.0, chooser=None, margin: float = 0.0):
        self.params = None if params is None else _f32(params).reshape(MLP_PARAMS).copy()
        self.alpha = float(alpha)
        h = ctypes.c_void_p()
        pp = None if self.params is None else _ptr(self.params, ctypes.c_float)
        _check(lib().nr_model_create(pp, self.alpha, ctypes.byref(h)), "nr_model_create")
        self._h = h
        self.chooser = None
        self.margin = float(margin)
        if chooser is not None:
            self.set_chooser(chooser, margin)

    def set_chooser(self, params, margin: float = 0.0) -> None:
        self.chooser = _f32(params).reshape(MLP_PARAMS).copy()
        self.margin = float(margin)
        _check(lib().nr_model_set_chooser(self._h, _ptr(self.chooser, ctypes.c_float),
                                          self.margin), "nr_model_set_chooser")

    def chooser_weights(self):
        """(weights int32 [16, 64], signs int32 [64], margin int): the
        chooser's exact form."""
        w = np.zeros(16 * 64, np.int32)
        sg = np.zeros(64, np.int32)
        mi = ctypes.c_int()
        _check(lib().nr_model_get_chooser(self._h, _ptr(w, ctypes.c_int), _ptr(sg, ctypes.c_int),
                                          ctypes.byref(mi)), "nr_model_get_chooser")
        return w.reshape(16, 64), sg, mi.value

    @property
    def handle(self):
        return self._h

    def score(self, X, device: int = 0) -> np.ndarray:
        X = _f32(X).reshape(-1, MLP_FEATURES)
        y = np.zeros(len(X), np.float32)
        _check(lib().nr_model_score(self._h, _ptr(X, ctypes.c_float), len(X),
                                    _ptr(y, ctypes.c_float), int(device)), "nr_model_score")
        return y

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.nr_model_destroy(h)
            self._h = ctypes.c_void_p()


def solve_learned(graph: Graph, model: Model, config: Config | None = None, **kw) -> dict:
    """nr_solve_learned: order "learned" (the default here) uses the model's
    order scorer, rule "learned" its path chooser."""
    if config is None:
        kw.setdefault("order", "learned")
    cfg = config if config is not None else make_config(**kw)
    r = Result()
    _check(lib().nr_solve_learned(graph.handle, ctypes.byref(cfg), model.handle, ctypes.byref(r)),
           "nr_solve_learned")
    return r.as_dict()


# ---------------------------------------------------------------- checkpoints
class Checkpoint:
    """Pending subproblems of a time-budgeted solve (nr_checkpoint)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        n, m, d = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_longlong()
        _check(lib().nr_checkpoint_info(self._h, ctypes.byref(n), ctypes.byref(m),
                                        ctypes.byref(d)), "nr_checkpoint_info")
        return {"pending": n.value, "pending_mass": m.value, "nodes_done": d.value}

    def save(self, path: str) -> None:
        _check(lib().nr_checkpoint_save(self._h, os.fsencode(path)), "nr_checkpoint_save")

    @classmethod
    def load(cls, path: str) -> "Checkpoint":
        h = ctypes.c_void_p()
        _check(lib().nr_checkpoint_load(os.fsencode(path), ctypes.byref(h)), "nr_checkpoint_load")
        return cls(h.value)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            _lib.nr_checkpoint_destroy(h)
            self._h = ctypes.c_void_p()


def solve_checkpointed(graph: Graph, resume: Checkpoint | None = None,
                       config: Config | None = None, **kw):
    """(result, Checkpoint or None): the checkpoint holds the pending
    subproblems when the time budget expired first."""
    cfg = config if config is not None else make_config(**kw)
    r = Result()
    out = ctypes.c_void_p()
    _check(lib().nr_solve_checkpointed(graph.handle, ctypes.byref(cfg),
                                       resume.handle if resume is not None else None,
                                       ctypes.byref(out), ctypes.byref(r)),
           "nr_solve_checkpointed")
    return r.as_dict(), (Checkpoint(out.value) if out.value else None)
