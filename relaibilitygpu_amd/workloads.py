"""Deterministic synthetic workloads (graphs are generated, never downloaded).

* ``er_instance(i)`` - one of the batched Erdos-Renyi instances: generator
  xoshiro256** seeded with splitmix64(0x5EED0000 + i); n ~ U{12..40}; mean
  degree d ~ U[4, 8] giving G(n, d/(n-1)) over pairs in lexicographic order;
  redrawn while E > 160 or no node pair is at BFS distance >= 3; (s, t)
  uniform among the pairs a < b at distance >= 3; p_e ~ U[0.80, 0.99] in
  edge order.  Bit-identical on every machine (integer RNG, doubles built
  from the top 53 bits).
* ``grid_graph`` - r x c grid, node = r_i * c + c_j, right and down edges.
* ``complete_graph`` - K_n with uniform p (the reference's COMP files).
* ``barabasi_albert`` - preferential attachment from a triangle.

Edge ids are the order of first appearance of (min, max) scanning nodes in
ascending order with rows sorted by neighbour (workingPad/graph.h:76-90), so
writing a generated graph to the reference file format and loading it back
gives the same ids.
"""
from __future__ import annotations

from collections import deque

MASK64 = (1 << 64) - 1


def splitmix64(x: int) -> tuple[int, int]:
    x = (x + 0x9E3779B97F4A7C15) & MASK64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return x, z ^ (z >> 31)


class Xoshiro256ss:
    """xoshiro256** (Blackman & Vigna), state seeded by splitmix64."""

    def __init__(self, seed: int):
        s = []
        x = seed & MASK64
        for _ in range(4):
            x, z = splitmix64(x)
            s.append(z)
        self.s = s

    @staticmethod
    def _rotl(x: int, k: int) -> int:
        return ((x << k) | (x >> (64 - k))) & MASK64

    def next(self) -> int:
        s = self.s
        result = (self._rotl((s[1] * 5) & MASK64, 7) * 9) & MASK64
        t = (s[1] << 17) & MASK64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / (1 << 53))


def _bfs_dist(n: int, adj: list[list[int]], s: int) -> list[int]:
    d = [-1] * n
    d[s] = 0
    q = deque([s])
    while q:
        u = q.popleft()
        for v in adj[u]:
            if d[v] < 0:
                d[v] = d[u] + 1
                q.append(v)
    return d


def er_instance(i: int, max_edges: int = 160):
    """Returns (n, edges, probs, src, dst) for batched instance i."""
    rng = Xoshiro256ss(0x5EED0000 + i)
    while True:
        n = 12 + int(rng.uniform() * 29)
        d = 4.0 + 4.0 * rng.uniform()
        pe = d / (n - 1)
        edges = [(a, b) for a in range(n) for b in range(a + 1, n) if rng.uniform() < pe]
        if not edges or len(edges) > max_edges:
            continue
        adj = [[] for _ in range(n)]
        for a, b in edges:
            adj[a].append(b)
            adj[b].append(a)
        far = []
        for a in range(n):
            da = _bfs_dist(n, adj, a)
            far += [(a, b) for b in range(a + 1, n) if da[b] >= 3]
        if not far:
            continue
        s, t = far[int(rng.uniform() * len(far))]
        probs = [0.80 + 0.19 * rng.uniform() for _ in edges]
        return n, edges, probs, s, t


def grid_graph(rows: int, cols: int, probs):
    """rows x cols grid; edges right then down per node in node order."""
    edges = []
    for r in range(rows):
        for c in range(cols):
            u = r * cols + c
            if c + 1 < cols:
                edges.append((u, u + 1))
            if r + 1 < rows:
                edges.append((u, u + cols))
    edges.sort()
    return rows * cols, edges, list(probs)[: len(edges)]


def complete_graph(n: int, p: float):
    edges = [(a, b) for a in range(n) for b in range(a + 1, n)]
    return n, edges, [p] * len(edges)


def barabasi_albert(n: int, m: int, p: float, seed: int):
    """Preferential attachment from a triangle; each new node links to m
    distinct existing nodes chosen with probability proportional to degree."""
    rng = Xoshiro256ss(seed)
    edges = {(0, 1), (0, 2), (1, 2)}
    targets = [0, 1, 1, 2, 0, 2]
    for v in range(3, n):
        chosen = set()
        while len(chosen) < m:
            chosen.add(targets[int(rng.uniform() * len(targets))])
        for u in sorted(chosen):
            edges.add((u, v))
            targets += [u, v]
    edges = sorted(edges)
    return n, edges, [p] * len(edges)


def to_csr(n: int, edges):
    """CSR with both directions, rows sorted by neighbour; edge ids in the
    order of `edges` (which is first-appearance order when sorted)."""
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
    return rp, ci, ei


def write_graph_file(path: str, n: int, edges, probs) -> None:
    """The reference's adjacency-list format with type U (graph.h:53-72)."""
    rows = [[] for _ in range(n)]
    for (a, b), p in zip(edges, probs):
        rows[a].append((b, p))
        rows[b].append((a, p))
    with open(path, "w") as f:
        f.write(f"{n} {2 * len(edges)} U\n")
        for u in range(n):
            rows[u].sort()
            f.write(f"{len(rows[u])} " + " ".join(f"{v} {p!r}" for v, p in rows[u]) + "\n")


def random_digraph(n: int, mean_out: float, p_lo: float, p_hi: float, seed: int):
    """Directed G(n, mean_out/(n-1)) over ordered pairs (u, v), u != v, in
    lexicographic order; arc i gets id i and p ~ U[p_lo, p_hi]."""
    rng = Xoshiro256ss(seed)
    pa = mean_out / (n - 1)
    arcs = [(u, v) for u in range(n) for v in range(n) if u != v and rng.uniform() < pa]
    probs = [p_lo + (p_hi - p_lo) * rng.uniform() for _ in arcs]
    return n, arcs, probs


def to_csr_directed(n: int, arcs):
    """CSR of one-way arcs (row u lists u -> v), rows sorted by head."""
    rows = [[] for _ in range(n)]
    for i, (a, b) in enumerate(arcs):
        rows[a].append((b, i))
    rp, ci, ei = [0], [], []
    for u in range(n):
        for v, i in sorted(rows[u]):
            ci.append(v)
            ei.append(i)
        rp.append(len(ci))
    return rp, ci, ei
