/*
 * nr_device.h - device-side building blocks of the expand kernels (gfx950,
 * wave64).  One lane owns one subproblem ("item"): its two edge-state bit
 * planes live in VGPRs; its allowed-adjacency rows (row u = the neighbours of
 * u over edges that are not FAILED), BFS level sets and path bookkeeping live
 * in an LDS slice laid out [slot][lane] so a wave's accesses fall in distinct
 * banks; the graph's static tables are staged in LDS per workgroup.
 *
 * Edge states: known/up bit planes.  UNKNOWN (0,0), WORKING (1,1),
 * FAILED (1,0); the reference's 2-bit codes 0b11/0b01/0b00
 * (workingPad/common.h:22-24).
 */
#ifndef NR_DEVICE_H
#define NR_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>

namespace nr {

typedef unsigned long long u64;
typedef unsigned int u32;
typedef unsigned short u16;

/* Double-double partial (hi, lo) for compensated sums. */
struct DD {
  double hi, lo;
};

__host__ __device__ inline void two_sum(double a, double b, double& s, double& e) {
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

/* Neumaier accumulator held per lane. */
struct Acc2 {
  double s, c;
  __device__ void init() { s = 0.0; c = 0.0; }
  __device__ void add(double x) {
    double t = s + x;
    if (fabs(s) >= fabs(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
  }
};

/* LDS layout of the per-workgroup graph tables followed by the per-lane
 * scratch slices.  Offsets in bytes; computed identically on host and
 * device. */
struct Layout {
  int n, e, w;
  int off_sadj, off_inc, off_ep, off_eidm, off_deg, off_p, off_q;
  int off_lvl, off_uh, off_rows;
  int off_extra; /* kernel-specific region (segment prefix table, counters) */
  int table_bytes, total_bytes;
  int block, ns_bytes;
};

inline Layout make_layout(int n, int e, int w, int block, int ns_bytes,
                          int extra_bytes = 0, bool tables = true) {
  Layout L;
  L.n = n;
  L.e = e;
  L.w = w;
  L.block = block;
  L.ns_bytes = ns_bytes;
  int o = 0;
  auto take = [&](int bytes, int align) {
    o = (o + align - 1) / align * align;
    int at = o;
    o += bytes;
    return at;
  };
  const int ee = e > 0 ? e : 1;
  const int t = tables ? 1 : 0; /* batched instances read tables from HBM */
  /* the table part is also the head of the device table blob (the host
   * packs it with these offsets), so staging it is one flat copy */
  L.off_p = take(t * 8 * ee, 16);
  L.off_q = take(t * 8 * ee, 16);
  L.off_sadj = take(t * (8 * n + 8), 16);
  L.off_inc = take(t * (8 * n * w + 8), 16);
  L.off_ep = take(t * 4 * ee, 16);
  L.off_deg = take(t * (4 * n + 4), 16);
  L.off_eidm = take(t * (2 * n * n + 4), 16);
  L.table_bytes = (o + 15) / 16 * 16;
  L.off_extra = take(extra_bytes, 16);
  L.off_rows = take(ns_bytes * n * block, 16);
  L.off_lvl = take(ns_bytes * n * block, 16);
  /* UNKNOWN hops of a path (or of a terminal tree's edges): at most n - 1 */
  L.off_uh = take(2 * (n > 1 ? n - 1 : 1) * block, 16);
  L.total_bytes = (o + 15) / 16 * 16;
  return L;
}

/* Global-memory graph tables (built on the host, see nr_engine.hip). */
struct DevGraph {
  const double* p;
  const double* q;
  const float* lk;  /* [e] forkBest's pivot key: logf of p rounded to FP32 (graph.h:138) */
  const u64* sadj;  /* [n] static neighbour set (n <= 64) */
  const u64* inc;   /* [n*w] incident edge set (n <= 64) */
  const u32* ep;    /* [e] endpoints a | b << 16 */
  const u32* deg;   /* [n] incident edge count (n <= 64) */
  const u16* eidm;  /* [n*n] edge id of (u,v), 0xffff if none (n <= 64) */
  const int* rp;    /* [n+1] out-arc rows */
  const u32* arc;   /* out-arcs v | e << 16, each row sorted by (v, e) */
  const int* irp;   /* [n+1] in-arc rows */
  const u32* iarc;  /* in-arcs u | e << 16, each row sorted by (u, e) */
  const u16* terms; /* K-terminal mode (nterms >= 2): terms[0] is the root */
  u64 tmask[4];     /* terminal set as a node set */
  int nterms;
  int n, e, w, src, dst;
  int period; /* alternating sweep: flip every `period` FAILED edges */
  int pivot;  /* two-way rule: branch on the best UNKNOWN edge of the whole graph */
  /* learned path choice: [e] round(-4 log2 p_e) and [e] round(-4 log2 q_e),
   * each clamped to 0..63 (host-computed, so device and oracle share them) */
  const unsigned char* lqp;
  const unsigned char* lqq;
};

/* What the fewest-UNKNOWN path search of one item records for the learned
 * choice of its sweep direction (both directions' features derive from it). */
struct ChoiceRaw {
  int k;                /* UNKNOWN hops of the path */
  int nS, cutS, outS;   /* src's WORKING component: nodes, allowed edges leaving
                           it (all UNKNOWN), all edges leaving it */
  int nT, cutT, outT;   /* the same for dst's component */
  int f, u;             /* FAILED and UNKNOWN edges of the item */
  int lp;               /* sum of lqp over the path's hops */
  int e_src, e_dst;     /* the hop at the src end and the hop at the dst end */
};

/* Reproducible pseudo-random draw of an item (splitmix64 finaliser over its
 * planes): 1 for a quarter of the items (top two bits zero).  A choice log
 * sweeps in the alternating rule's direction and flips it where this is 1,
 * so both directions occur in every kind of state while the subtrees below
 * follow the rule the learned choice is to improve on. */
__host__ __device__ inline unsigned long long explore_mix(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
template <int W>
__host__ __device__ inline int explore_bit(const unsigned long long (&known)[W],
                                           const unsigned long long (&up)[W]) {
  unsigned long long h = 0x9E3779B97F4A7C15ull;
  for (int i = 0; i < W; ++i) {
    h = explore_mix(h ^ known[i]);
    h = explore_mix(h ^ up[i]);
  }
  return (h >> 62) == 0 ? 1 : 0;
}

/* The workgroup size as the stride of a lane's LDS slices ([slot][lane]):
 * always a power of two, so i * stride is a shift (a 32-bit multiply is a
 * multi-pass VALU instruction on CDNA). */
struct Pow2 {
  int sh = 0;
  __host__ __device__ Pow2& operator=(int v) {
    sh = __builtin_ctz((unsigned)v);
    return *this;
  }
  __host__ __device__ operator int() const { return 1 << sh; }
};
__host__ __device__ inline int operator*(int i, Pow2 p) { return i << p.sh; }
__host__ __device__ inline int operator*(Pow2 p, int i) { return i << p.sh; }

/* One queue buffer in structure-of-arrays form: plane w of item i at
 * known[w * stride + i]; rows (when items carry them) likewise per node. */
struct QBuf {
  u64* known; /* [w][stride] */
  u64* up;    /* [w][stride] */
  void* rows; /* [n][stride] node sets (u32 or u64), or null */
  double* mass;
};

/* Cooperative copy of the tables into LDS (all threads of the block). */
/* The graph tables into LDS: the device blob starts with the table part of
 * the LDS layout, byte for byte (p at offset 0), so this is one copy of
 * table_bytes in 16-byte vectors, all loads issued before any store waits. */
__device__ inline void stage_tables(const DevGraph& g, const Layout& L,
                                    unsigned char* smem) {
  const uint4* s = (const uint4*)g.p;
  uint4* d = (uint4*)smem;
  for (int i = threadIdx.x; i < L.table_bytes / 16; i += blockDim.x) d[i] = s[i];
}

template <typename NS>
__device__ inline int ctzs(NS x);
template <>
__device__ inline int ctzs<u32>(u32 x) {
  return __builtin_ctz(x);
}
template <>
__device__ inline int ctzs<u64>(u64 x) {
  return __builtin_ctzll(x);
}

template <int W>
__device__ inline u64 pick(const u64 (&a)[W], int wi) {
  u64 r = a[0];
#pragma unroll
  for (int i = 1; i < W; ++i)
    if (wi == i) r = a[i];
  return r;
}

/* The per-lane expander: BFS + lowest-index parent path + unknown hops. */
/* forkBest's pivot among a path's UNKNOWN hops (main.cu:356-392).  The
 * reference gives hop r of the path (counted from the dst end, every hop
 * counted) to lane r mod 32; a lane keeps the first hop of highest key
 * (strict >), then a shfl_down tree over the 32 lanes keeps position i's
 * value on ties with position i + off.  So the pivot is the hop of highest
 * key, ties going to the smallest 5-bit reversal of r mod 32, then to the
 * smallest r.  Keys are logf of p in FP32, as the reference stores them. */
struct ForkPick {
  int best = -1, tie = 0;
  float key = 0.0f;
  __device__ void offer(int e, int r, float k) {
    const int t = (int)((__builtin_bitreverse32((u32)(r & 31)) >> 27) << 16) | r;
    if (best < 0 || k > key || (k == key && t < tie)) {
      best = e;
      key = k;
      tie = t;
    }
  }
};

template <int W, typename NS>
struct Expander {
  const double* p;
  const double* q;
  const u64* sadj;
  const u64* inc;
  const u32* ep;
  const u32* deg;
  const u16* eidm;
  const float* lk; /* forkBest keys (global memory) */
  NS* rows; /* this lane's allowed-adjacency rows, stride `stride` */
  NS* lvl;  /* this lane's level sets, stride `stride` */
  u16* uh;  /* this lane's unknown-hop stack, stride `stride` */
  Pow2 stride;
  int n, src, dst;
  mutable int plen; /* hops of the last path found */
  /* find<true>(): bit h set when hop h (BFS level h-1 to h) is UNKNOWN, and
   * the path's nodes (one per BFS level, dst on level plen) */
  mutable u64 hopm;
  mutable NS pathm;
  const u16* terms;
  int nterms;
  NS tmask;
  int ne, pivot; /* edges; edge-pivot rule instead of forkBest (fork_item) */

  __device__ void bind(unsigned char* smem, const Layout& L, int slot, int src_,
                       int dst_) {
    p = (const double*)(smem + L.off_p);
    q = (const double*)(smem + L.off_q);
    sadj = (const u64*)(smem + L.off_sadj);
    inc = (const u64*)(smem + L.off_inc);
    ep = (const u32*)(smem + L.off_ep);
    deg = (const u32*)(smem + L.off_deg);
    eidm = (const u16*)(smem + L.off_eidm);
    rows = (NS*)(smem + L.off_rows) + slot;
    lvl = (NS*)(smem + L.off_lvl) + slot;
    uh = (u16*)(smem + L.off_uh) + slot;
    stride = L.block;
    n = L.n;
    src = src_;
    dst = dst_;
  }

  /* ---- expander interface used by the level and items kernels ---- */
  static constexpr bool kStage = true; /* graph tables staged in LDS */
  static constexpr int kRowBytes = (int)sizeof(NS);
  static constexpr bool kChoose = false; /* sweep direction picked by the learned chooser */
  /* dead children found at creation (find_item_marked, dead_children) */
  static constexpr bool kSettle = true;
  __device__ void bind_graph(unsigned char* smem, const Layout& L, const DevGraph& g, int slot) {
    bind(smem, L, slot, g.src, g.dst);
    terms = g.terms;
    nterms = g.nterms;
    tmask = (NS)g.tmask[0];
    ne = g.e;
    pivot = g.pivot;
    lk = g.lk;
  }
  /* Queue form of the rows: with 32-bit node sets and n <= 16 nodes, rows
   * 2j and 2j+1 share the 32-bit word j (low and high half), so an item
   * carries ceil(n/2) words instead of n; otherwise one NS word per node.
   * The choice is uniform over the launch (n is the graph's). */
  __device__ bool packed() const { return sizeof(NS) == 4 && n <= 16; }
  /* the item's rows into this lane's slice: 16 loads in flight before the
   * first LDS store waits (one memory round for n <= 16) */
  __device__ void load(const QBuf& in, u64 qstride, u64 slot) const {
    if (packed()) {
      const u32* win = (const u32*)in.rows + slot;
      const int nw = (n + 1) >> 1;
      u32 v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (j < nw) v[j] = win[(u64)j * qstride];
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (j < nw) {
          rows[(2 * j) * stride] = (NS)(v[j] & 0xffffu);
          if (2 * j + 1 < n) rows[(2 * j + 1) * stride] = (NS)(v[j] >> 16);
        }
      return;
    }
    const NS* rin = (const NS*)in.rows + slot;
    for (int u0 = 0; u0 < n; u0 += 16) {
      NS v[16];
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (u0 + j < n) v[j] = rin[(u64)(u0 + j) * qstride];
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (u0 + j < n) rows[(u0 + j) * stride] = v[j];
    }
  }
  /* rows of an item given by its planes alone */
  __device__ void derive(const u64 (&known)[W], const u64 (&up)[W]) const {
    u64 Dn[W];
#pragma unroll
    for (int i = 0; i < W; ++i) Dn[i] = known[i] & ~up[i];
    build_rows(Dn);
  }
  __device__ int find_item(const u64 (&known)[W], const u64 (&)[W]) const {
    return nterms >= 2 ? find_tree(known) : find(known);
  }
  /* find() for two terminals that also records the path's nodes and its
   * UNKNOWN hops for dead_children() */
  __device__ int find_item_marked(const u64 (&known)[W], const u64 (&)[W]) const {
    return find<true>(known);
  }

  /* The children of the item find<true>() just walked whose FAILED hop is a
   * bridge between src and dst in the item's allowed graph (edges not
   * FAILED): such a child has no src-dst path left and is dead when it is
   * dequeued (main.cu:579-582); every other child has one.  Hop h joins
   * v(h-1) and v(h), v(0) = src, v(plen) = dst; v(h) is the path node of BFS
   * level h (pathm & lvl[h]).  Every node of the levels before h is reached
   * from src without hop h (along its own shortest path), and so is every
   * node of level h but v(h) (its BFS parent is on level h-1, and the edge
   * from it is not hop h).  Hop h is a bridge exactly when the set R reached
   * without it meets none of v(h..plen): it is not when v(h) has a
   * neighbour other than v(h-1) on the levels up to h (one row test);
   * otherwise R starts as those levels without v(h) and grows from level h
   * (whose neighbours may lie deeper) or, for the last hop, from level
   * plen-1 (level plen was not built), v(h-1)'s row without v(h).  A
   * WORKING hop has no child.  Bit j: the j-th UNKNOWN hop from the src
   * side (child j). */
  __device__ u64 dead_children() const {
    const NS dbit = (NS)1 << dst;
    NS below = 0; /* levels 0 .. h-1 */
    u64 dead = 0, todo = hopm;
    int j = 0;
    NS lprev = lvl[0];
    for (int h = 1; todo; ++h) {
      below |= lprev;
      const NS lh = h < plen ? lvl[h * stride] : (NS)0;
      if ((todo >> h) & 1ull) {
        todo &= todo - 1; /* bits below h are clear */
        const NS vh = h < plen ? (lh & pathm) : dbit;
        const NS vp = lprev & pathm;
        const NS rest = below | (lh & ~vh); /* reached without hop h */
        if (!(rows[ctzs<NS>(vh) * stride] & rest & ~vp)) {
          const NS suf = pathm & ~below;
          const int up = ctzs<NS>(vp);
          NS R = rest;
          NS F = h < plen ? (lh & ~vh) : lprev;
          bool hit = false;
          while (F) {
            const int u = ctzs<NS>(F);
            F &= F - 1;
            NS r = rows[u * stride];
            if (u == up) r &= ~vh;
            const NS x = r & ~R;
            if (x & suf) {
              hit = true;
              break;
            }
            R |= x;
            F |= x;
          }
          if (!hit) dead |= 1ull << j;
        }
        ++j;
      }
      lprev = lh;
    }
    return dead;
  }

  /* K-terminal form of find(): BFS from terms[0], whole levels, until every
   * terminal is reached (-1 if the frontier dies first); then each other
   * terminal, in list order, walks back to the tree built so far under the
   * same parent rule, pushing the UNKNOWN edges it crosses.  With the
   * terminals {src, dst} this is find(). */
  __device__ int find_tree(const u64 (&known)[W]) const {
    const int root = terms[0];
    NS vis = (NS)1 << root, F = vis;
    lvl[0] = F;
    NS need = tmask & ~vis;
    int L = 0;
    while (need) {
      NS nx = 0, f = F;
      while (f) {
        const int u = ctzs<NS>(f);
        f &= f - 1;
        nx |= rows[u * stride];
      }
      nx &= ~vis;
      if (!nx) return -1;
      ++L;
      lvl[L * stride] = nx;
      vis |= nx;
      F = nx;
      need &= ~nx;
    }
    NS marked = (NS)1 << root;
    int k = 0, len = 0;
    for (int i = 1; i < nterms; ++i) {
      int v = terms[i];
      while (!((marked >> v) & (NS)1)) {
        int l = 1;
        while (l <= L && !((lvl[l * stride] >> v) & (NS)1)) ++l;
        if (l > L) return -2;
        const NS cand = rows[v * stride] & lvl[(l - 1) * stride];
        if (!cand) return -2;
        const int u = ctzs<NS>(cand);
        const int e = eidm[u * n + v];
        if (!((pick<W>(known, e >> 6) >> (e & 63)) & 1ull)) uh[(k++) * stride] = (u16)e;
        ++len;
        marked |= (NS)1 << v;
        v = u;
      }
    }
    plen = len;
    return k;
  }
  __device__ void store_child(const QBuf& out, u64 qstride, u64 slot, int e) const {
    if (packed()) {
      u32* wout = (u32*)out.rows + slot;
      for (int u = 0; u < n; u += 2) {
        u32 v = (u32)child_row(u, e);
        if (u + 1 < n) v |= (u32)child_row(u + 1, e) << 16;
        wout[(u64)(u >> 1) * qstride] = v;
      }
      return;
    }
    NS* rout = (NS*)out.rows + slot;
    for (int u = 0; u < n; ++u) rout[(u64)u * qstride] = child_row(u, e);
  }
  /* neighbours of u over WORKING edges */
  __device__ NS working_row(int u, const u64 (&up)[W]) const {
    NS r = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
      u64 m = inc[u * W + i] & up[i];
      while (m) {
        const int e = i * 64 + __builtin_ctzll(m);
        m &= m - 1;
        const u32 x = ep[e];
        r |= (NS)1 << ((int)((x & 0xffffu) ^ (x >> 16)) ^ u);
      }
    }
    return r;
  }
  /* forkBest (main.cu:304-430, 550-567): 0 when a WORKING-only path joins
   * src to dst (success), -1 dead, -2 inconsistency, else 1 + the edge to
   * branch on: the path's UNKNOWN hop of highest p, the one nearest dst on
   * ties */
  __device__ int fork_item(const u64 (&known)[W], const u64 (&up)[W]) const {
    /* confirmed BFS: every goal node (dst, or all terminals) over WORKING edges */
    const NS goal = nterms >= 2 ? tmask : (NS)1 << dst;
    NS vis = (NS)1 << (nterms >= 2 ? terms[0] : src), F = vis;
    if (!(goal & ~vis)) return 0;
    while (F) {
      NS nx = 0;
      NS f = F;
      while (f) {
        const int u = ctzs<NS>(f);
        f &= f - 1;
        nx |= working_row(u, up);
      }
      nx &= ~vis;
      vis |= nx;
      if (!(goal & ~vis)) return 0;
      F = nx;
    }
    const int k = find_item(known, up);
    if (k <= 0) return k;
    if (pivot) return 1 + pivot_edge(known);
    const int b = fork_pick(known);
    return b < 0 ? -2 : 1 + b;
  }
  /* forkBest's pivot (ForkPick) over the path just found: the walk of
   * find() / find_tree() again, with each hop's index along the walk */
  __device__ int fork_pick(const u64 (&known)[W]) const {
    ForkPick fp;
    auto hop = [&](int e, int r) {
      if (!((pick<W>(known, e >> 6) >> (e & 63)) & 1ull)) fp.offer(e, r, lk[e]);
    };
    if (nterms >= 2) {
      NS marked = (NS)1 << terms[0];
      int r = 0;
      for (int i = 1; i < nterms; ++i) {
        int v = terms[i];
        while (!((marked >> v) & (NS)1)) {
          int l = 1;
          while (l < n && !((lvl[l * stride] >> v) & (NS)1)) ++l;
          const NS cand = l < n ? rows[v * stride] & lvl[(l - 1) * stride] : (NS)0;
          if (!cand) return -1;
          const int u = ctzs<NS>(cand);
          hop(eidm[u * n + v], r++);
          marked |= (NS)1 << v;
          v = u;
        }
      }
    } else {
      int v = dst;
      for (int l = plen; l >= 1; --l) {
        const int u = ctzs<NS>(rows[v * stride] & lvl[(l - 1) * stride]);
        hop(eidm[u * n + v], plen - l);
        v = u;
      }
    }
    return fp.best;
  }
  /* the edge-pivot rule (workingPad/cpu_reference.cpp:107-145): the first
   * UNKNOWN edge of highest p over the whole graph (strict >, so the lowest
   * id on ties, cpu_reference.cpp:112) */
  __device__ int pivot_edge(const u64 (&known)[W]) const {
    int best = -1;
    double bp = -1.0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
      u64 m = ~known[i];
      while (m) {
        const int e = i * 64 + __builtin_ctzll(m);
        m &= m - 1;
        if (e >= ne) break;
        if (p[e] > bp) {
          bp = p[e];
          best = e;
        }
      }
    }
    return best;
  }
  /* store_child() for the item of another lane of the same workgroup
   * (slot offset `delta`): reads that lane's rows slice */
  __device__ void store_child_of(const QBuf& out, u64 qstride, u64 slot, int e, int delta) const {
    const u32 x = ep[e];
    const int a0 = (int)(x & 0xffffu), b0 = (int)(x >> 16);
    const NS* r = rows + delta;
    if (packed()) {
      u32* wout = (u32*)out.rows + slot;
      for (int u = 0; u < n; u += 2) {
        u32 lo = (u32)r[u * stride], hi = u + 1 < n ? (u32)r[(u + 1) * stride] : 0u;
        if (u == a0) lo &= ~(1u << b0);
        if (u == b0) lo &= ~(1u << a0);
        if (u + 1 == a0) hi &= ~(1u << b0);
        if (u + 1 == b0) hi &= ~(1u << a0);
        wout[(u64)(u >> 1) * qstride] = lo | (hi << 16);
      }
      return;
    }
    NS* rout = (NS*)out.rows + slot;
    for (int u = 0; u < n; ++u) {
      NS v = r[u * stride];
      if (u == a0) v &= ~((NS)1 << b0);
      if (u == b0) v &= ~((NS)1 << a0);
      rout[(u64)u * qstride] = v;
    }
  }

  /* rows of a child that keeps the parent's allowed edges */
  __device__ void store_rows(const QBuf& out, u64 qstride, u64 slot) const {
    if (packed()) {
      u32* wout = (u32*)out.rows + slot;
      for (int u = 0; u < n; u += 2) {
        const u32 hi = u + 1 < n ? (u32)rows[(u + 1) * stride] : 0u;
        wout[(u64)(u >> 1) * qstride] = (u32)rows[u * stride] | (hi << 16);
      }
      return;
    }
    NS* rout = (NS*)out.rows + slot;
    for (int u = 0; u < n; ++u) rout[(u64)u * qstride] = rows[u * stride];
  }

  /* allowed degree of u in the loaded item (undirected: out = in) */
  __device__ int out_degree(int u, const u64 (&)[W], const u64 (&)[W]) const {
    return sizeof(NS) == 4 ? __popc((u32)rows[u * stride]) : __popcll((u64)rows[u * stride]);
  }
  __device__ int in_degree(int u, const u64 (&k)[W], const u64 (&p)[W]) const {
    return out_degree(u, k, p);
  }
  __device__ void carry(const QBuf& in, u64 slot_in, const QBuf& out, u64 slot_out,
                        u64 qstride) const {
    const NS* rin = (const NS*)in.rows + slot_in;
    NS* rout = (NS*)out.rows + slot_out;
    const int nw = packed() ? (n + 1) >> 1 : n; /* packed: u32 words, NS is u32 */
    for (int u = 0; u < nw; ++u) rout[(u64)u * qstride] = rin[(u64)u * qstride];
  }

  /* Allowed neighbours of u (edges not FAILED): the static neighbour set
   * minus the ends of u's failed edges, or the union of the ends of u's
   * allowed edges, whichever set of edges is smaller. */
  __device__ NS adj_of(int u, const u64 (&Dn)[W]) const {
    NS a = (NS)sadj[u];
    u64 dn[W];
    int cnt = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
      dn[i] = inc[u * W + i] & Dn[i];
      cnt += __popcll(dn[i]);
    }
    if (cnt == 0) return a;
    if (2 * cnt <= (int)deg[u]) {
#pragma unroll
      for (int i = 0; i < W; ++i) {
        u64 m = dn[i];
        while (m) {
          int e = i * 64 + __builtin_ctzll(m);
          m &= m - 1;
          u32 x = ep[e];
          int o = (int)((x & 0xffffu) ^ (x >> 16)) ^ u;
          a &= ~((NS)1 << o);
        }
      }
    } else {
      a = 0;
#pragma unroll
      for (int i = 0; i < W; ++i) {
        u64 m = inc[u * W + i] & ~Dn[i];
        while (m) {
          int e = i * 64 + __builtin_ctzll(m);
          m &= m - 1;
          u32 x = ep[e];
          int o = (int)((x & 0xffffu) ^ (x >> 16)) ^ u;
          a |= (NS)1 << o;
        }
      }
    }
    return a;
  }

  /* rows[u] = adj_of(u) for every node: derives an item's rows from its
   * planes (root item, caller-supplied items). */
  __device__ void build_rows(const u64 (&Dn)[W]) const {
    for (int u = 0; u < n; ++u) rows[u * stride] = adj_of(u, Dn);
  }

  /* Optimistic BFS from src (device_code.cuh:80-194) over the item's rows,
   * and the path to dst under the lowest-index parent rule: the parent of v
   * is the lowest-index node of the previous BFS level joined to v by an
   * allowed edge.  Pushes the path's UNKNOWN hops on uh[] in dst->src order.
   * Returns the number of UNKNOWN hops, -1 if dst is unreachable, -2 on an
   * internal inconsistency (never expected). */
  template <bool kMark = false>
  __device__ int find(const u64 (&known)[W]) const {
    if (kMark) {
      hopm = 0;
      pathm = (NS)1 << dst;
    }
    if (src == dst) {
      plen = 0;
      return 0;
    }
    /* the rows are symmetric (undirected graphs): dst enters the level
     * after frontier F exactly when its own row meets F, a test of one row,
     * so the levels before are expanded whole, four rows a step with their
     * loads in flight together and no test inside the step */
    const NS drow = rows[dst * stride];
    NS vis = (NS)1 << src;
    NS F = vis;
    lvl[0] = F;
    int L = 0;
    while (!(drow & F)) {
      NS nx = 0;
      NS f = F;
      while (f) {
        const int u0 = ctzs<NS>(f);
        f &= f - 1;
        const int u1 = f ? ctzs<NS>(f) : u0;
        f &= f - 1;
        const int u2 = f ? ctzs<NS>(f) : u0;
        f &= f - 1;
        const int u3 = f ? ctzs<NS>(f) : u0;
        f &= f - 1;
        nx |= (rows[u0 * stride] | rows[u1 * stride]) | (rows[u2 * stride] | rows[u3 * stride]);
      }
      nx &= ~vis;
      if (!nx) return -1;
      ++L;
      lvl[L * stride] = nx;
      vis |= nx;
      F = nx;
    }
    ++L; /* dst's level */
    plen = L;
    int v = dst, k = 0;
    for (int l = L; l >= 1; --l) {
      const NS cand = rows[v * stride] & lvl[(l - 1) * stride];
      if (!cand) return -2;
      const int u = ctzs<NS>(cand);
      const int e = eidm[u * n + v];
      if (!((pick<W>(known, e >> 6) >> (e & 63)) & 1ull)) {
        uh[(k++) * stride] = (u16)e;
        if (kMark) hopm |= 1ull << l;
      }
      if (kMark) pathm |= (NS)1 << u;
      v = u;
    }
    return k;
  }

  /* Row u derived on first use from the item's FAILED plane Dn and cached in
   * this lane's LDS slice (bit u of `have`): the form used when items carry
   * no rows. */
  __device__ NS lazy_row(int u, const u64 (&Dn)[W], NS& have) const {
    if ((have >> u) & (NS)1) return rows[u * stride];
    const NS a = adj_of(u, Dn);
    rows[u * stride] = a;
    have |= (NS)1 << u;
    return a;
  }

  /* find() for an item given by its planes alone; same BFS, same parent
   * rule, same unknown-hop stack. */
  __device__ int find_planes(const u64 (&known)[W], const u64 (&up)[W]) const {
    if (src == dst) {
      plen = 0;
      return 0;
    }
    u64 Dn[W];
#pragma unroll
    for (int i = 0; i < W; ++i) Dn[i] = known[i] & ~up[i];
    NS have = 0;
    const NS dbit = (NS)1 << dst;
    NS vis = (NS)1 << src;
    NS F = vis;
    lvl[0] = F;
    int L = 0;
    for (;;) {
      NS nx = 0;
      NS f = F;
      while (f) {
        const int u = ctzs<NS>(f);
        f &= f - 1;
        nx |= lazy_row(u, Dn, have);
        if (nx & dbit) break;
      }
      nx &= ~vis;
      if (!nx) return -1;
      ++L;
      if (nx & dbit) break;
      lvl[L * stride] = nx;
      vis |= nx;
      F = nx;
    }
    plen = L;
    int v = dst, k = 0;
    for (int l = L; l >= 1; --l) {
      const NS cand = lazy_row(v, Dn, have) & lvl[(l - 1) * stride];
      if (!cand) return -2;
      const int u = ctzs<NS>(cand);
      const int e = eidm[u * n + v];
      if (!((pick<W>(known, e >> 6) >> (e & 63)) & 1ull)) uh[(k++) * stride] = (u16)e;
      v = u;
    }
    return k;
  }

  /* find_planes() with the smallest LDS slice: adjacency derived on every
   * use (no cache) and each node's BFS parent kept as a byte (lvb: node u
   * at (u>>2)*4*stride + (u&3) of this lane's slice).  The frontier is
   * expanded in ascending node order and a node's parent is the frontier
   * node that reaches it first, so the parent of v is the lowest-index
   * neighbour on the previous level, as in find(); the walk back follows
   * the parent bytes and derives no adjacency. */
  __device__ int find_lean(const u64 (&known)[W], const u64 (&up)[W], unsigned char* lvb) const {
    if (src == dst) {
      plen = 0;
      return 0;
    }
    u64 Dn[W];
#pragma unroll
    for (int i = 0; i < W; ++i) Dn[i] = known[i] & ~up[i];
    auto setp = [&](int v, int u) { lvb[(v >> 2) * 4 * stride + (v & 3)] = (unsigned char)u; };
    auto getp = [&](int v) { return (int)lvb[(v >> 2) * 4 * stride + (v & 3)]; };
    const NS dbit = (NS)1 << dst;
    NS vis = (NS)1 << src;
    NS F = vis;
    int L = 0;
    for (;;) {
      NS nx = 0;
      NS f = F;
      while (f) {
        const int u = ctzs<NS>(f);
        f &= f - 1;
        NS nb = adj_of(u, Dn) & ~(vis | nx); /* reached first from u */
        nx |= nb;
        while (nb) {
          setp(ctzs<NS>(nb), u);
          nb &= nb - 1;
        }
        if (nx & dbit) break;
      }
      if (!nx) return -1;
      ++L;
      if (nx & dbit) break;
      vis |= nx;
      F = nx;
    }
    plen = L;
    int v = dst, k = 0;
    for (int l = L; l >= 1; --l) {
      const int u = getp(v);
      const int e = eidm[u * n + v];
      if (!((pick<W>(known, e >> 6) >> (e & 63)) & 1ull)) uh[(k++) * stride] = (u16)e;
      v = u;
    }
    return k;
  }

  /* Row u of a child whose FAILED hop is edge e: the parent's row with the
   * ends of e removed from each other's rows. */
  __device__ NS child_row(int u, int e) const {
    const u32 x = ep[e];
    const int a = (int)(x & 0xffffu), b = (int)(x >> 16);
    NS r = rows[u * stride];
    if (u == a) r &= ~((NS)1 << b);
    if (u == b) r &= ~((NS)1 << a);
    return r;
  }
};

/* NW-word node sets held in VGPRs; every index is resolved by unrolled
 * compares so the arrays never spill to scratch. */
template <int NW>
__device__ inline bool bs_test(const u64 (&s)[NW], int v) {
  u64 w = 0;
#pragma unroll
  for (int k = 0; k < NW; ++k)
    if (k == (v >> 6)) w = s[k];
  return (w >> (v & 63)) & 1ull;
}
template <int NW>
__device__ inline void bs_set(u64 (&s)[NW], int v) {
#pragma unroll
  for (int k = 0; k < NW; ++k)
    if (k == (v >> 6)) s[k] |= 1ull << (v & 63);
}

/* LDS layout of the CSR expander: per-lane node levels (one byte per node,
 * four nodes per 32-bit word, laid out [word][lane]) and the unknown-hop
 * stack ([slot][lane] u16). */
inline Layout make_layout_csr(int n, int block, int extra_bytes = 0) {
  Layout L;
  std::memset(&L, 0, sizeof L);
  L.n = n;
  L.block = block;
  L.ns_bytes = 0;
  int o = 0;
  auto take = [&](int bytes, int align) {
    o = (o + align - 1) / align * align;
    int at = o;
    o += bytes;
    return at;
  };
  L.table_bytes = 0;
  L.off_extra = take(extra_bytes, 16);
  L.off_lvl = take(4 * ((n + 3) / 4) * block, 16);
  L.off_rows = L.off_lvl;
  L.off_uh = take(2 * n * block, 16);
  L.total_bytes = (o + 15) / 16 * 16;
  return L;
}

/* Per-lane expander over the CSR arcs, for graphs wider than a 64-bit node
 * set or with one-way arcs (workingPad 'D' files).  Same search as Expander:
 * BFS from src over arcs whose edge is not FAILED, a node marked visited
 * when first discovered; the parent of v is the lowest-index node of the
 * previous level with an allowed arc into v, lowest edge id on ties (the
 * in-arc rows are sorted by (u, e), so the first match is the rule's). */
template <int W, int NW>
struct CsrExpander {
  static constexpr bool kStage = false;
  static constexpr int kRowBytes = 0;
  static constexpr bool kChoose = false;
  static constexpr bool kSettle = false;
  const double* p;
  const double* q;
  const int* rp;
  const u32* arc;
  const int* irp;
  const u32* iarc;
  const u32* ep;
  const float* lk;    /* forkBest keys */
  unsigned char* lvb; /* this lane's level bytes: node u at (u>>2)*4*stride + (u&3) */
  u16* uh;            /* this lane's unknown-hop stack, stride `stride` */
  Pow2 stride;
  int n, src, dst;
  mutable int plen; /* hops of the last path found */
  const u16* terms;
  int nterms;
  u64 tmask[NW];
  int ne, pivot; /* edges; edge-pivot rule instead of forkBest (fork_item) */

  __device__ void bind_graph(unsigned char* smem, const Layout& L, const DevGraph& g, int slot) {
    ne = g.e;
    pivot = g.pivot;
    p = g.p;
    q = g.q;
    rp = g.rp;
    arc = g.arc;
    irp = g.irp;
    iarc = g.iarc;
    ep = g.ep;
    lk = g.lk;
    terms = g.terms;
    nterms = g.nterms;
#pragma unroll
    for (int k = 0; k < NW; ++k) tmask[k] = g.tmask[k];
    lvb = smem + L.off_lvl + 4 * slot;
    uh = (u16*)(smem + L.off_uh) + slot;
    stride = L.block;
    n = g.n;
    src = g.src;
    dst = g.dst;
  }
  __device__ void load(const QBuf&, u64, u64) const {}
  __device__ void derive(const u64 (&)[W], const u64 (&)[W]) const {}
  __device__ void store_child(const QBuf&, u64, u64, int) const {}
  __device__ void store_child_of(const QBuf&, u64, u64, int, int) const {}
  __device__ void carry(const QBuf&, u64, const QBuf&, u64, u64) const {}

  __device__ void store_rows(const QBuf&, u64, u64) const {}
  /* forkBest on the CSR arcs; same contract as Expander::fork_item */
  __device__ int fork_item(const u64 (&known)[W], const u64 (&up)[W]) const {
    /* confirmed BFS: every goal node (dst, or all terminals) over WORKING arcs */
    u64 vis[NW], F[NW], goal[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) vis[k] = F[k] = goal[k] = 0;
    const int root = nterms >= 2 ? terms[0] : src;
    bs_set<NW>(vis, root);
    bs_set<NW>(F, root);
    if (nterms >= 2) {
#pragma unroll
      for (int k = 0; k < NW; ++k) goal[k] = tmask[k];
    } else {
      bs_set<NW>(goal, dst);
    }
    auto reached = [&]() {
      bool all = true;
#pragma unroll
      for (int k = 0; k < NW; ++k) all = all && !(goal[k] & ~vis[k]);
      return all;
    };
    if (reached()) return 0;
    for (bool any = true; any;) {
      u64 nx[NW];
#pragma unroll
      for (int k = 0; k < NW; ++k) nx[k] = 0;
      any = false;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        u64 fw = F[k];
        while (fw) {
          const int u = k * 64 + __builtin_ctzll(fw);
          fw &= fw - 1;
          for (int j = rp[u]; j < rp[u + 1]; ++j) {
            const u32 x = arc[j];
            const int v = (int)(x & 0xffffu), e = (int)(x >> 16);
            if (bs_test<NW>(vis, v) || !((pick<W>(up, e >> 6) >> (e & 63)) & 1ull)) continue;
            bs_set<NW>(vis, v);
            bs_set<NW>(nx, v);
            any = true;
          }
        }
      }
      if (reached()) return 0;
#pragma unroll
      for (int k = 0; k < NW; ++k) F[k] = nx[k];
    }
    const int k = find_item(known, up);
    if (k <= 0) return k;
    if (pivot) return 1 + pivot_edge(known);
    const int b = fork_pick(known, up);
    return b < 0 ? -2 : 1 + b;
  }
  /* forkBest's pivot (ForkPick) over the path just found: the walk of
   * find_item() / find_tree() again, with each hop's index along the walk */
  __device__ int fork_pick(const u64 (&known)[W], const u64 (&up)[W]) const {
    ForkPick fp;
    auto parent = [&](int v, int& pe) {
      const int l = level_of(v);
      for (int j = irp[v]; j < irp[v + 1]; ++j) {
        const u32 x = iarc[j];
        const int u = (int)(x & 0xffffu), e = (int)(x >> 16);
        if (level_of(u) == l - 1 && allowed(known, up, e)) {
          pe = e;
          return u;
        }
      }
      return -1;
    };
    auto hop = [&](int e, int r) {
      if (!((pick<W>(known, e >> 6) >> (e & 63)) & 1ull)) fp.offer(e, r, lk[e]);
    };
    if (nterms >= 2) {
      u64 marked[NW];
#pragma unroll
      for (int k = 0; k < NW; ++k) marked[k] = 0;
      bs_set<NW>(marked, terms[0]);
      int r = 0;
      for (int i = 1; i < nterms; ++i) {
        int v = terms[i];
        while (!bs_test<NW>(marked, v)) {
          int pe = -1;
          const int u = parent(v, pe);
          if (u < 0) return -1;
          hop(pe, r++);
          bs_set<NW>(marked, v);
          v = u;
        }
      }
    } else {
      /* dst's own level byte is not written by find_item (it stops at the
       * level that reaches dst): its parents lie on level plen - 1 */
      int v = dst;
      for (int l = plen; l >= 1; --l) {
        int pu = -1, pe = -1;
        for (int j = irp[v]; j < irp[v + 1]; ++j) {
          const u32 x = iarc[j];
          const int u = (int)(x & 0xffffu), e = (int)(x >> 16);
          if (level_of(u) == l - 1 && allowed(known, up, e)) {
            pu = u;
            pe = e;
            break;
          }
        }
        if (pu < 0) return -1;
        hop(pe, plen - l);
        v = pu;
      }
    }
    return fp.best;
  }
  /* the edge-pivot rule (workingPad/cpu_reference.cpp:107-145): the first
   * UNKNOWN edge of highest p over the whole graph (strict >, so the lowest
   * id on ties, cpu_reference.cpp:112) */
  __device__ int pivot_edge(const u64 (&known)[W]) const {
    int best = -1;
    double bp = -1.0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
      u64 m = ~known[i];
      while (m) {
        const int e = i * 64 + __builtin_ctzll(m);
        m &= m - 1;
        if (e >= ne) break;
        if (p[e] > bp) {
          bp = p[e];
          best = e;
        }
      }
    }
    return best;
  }
  __device__ int out_degree(int u, const u64 (&known)[W], const u64 (&up)[W]) const {
    int d = 0;
    for (int j = rp[u]; j < rp[u + 1]; ++j) d += allowed(known, up, (int)(arc[j] >> 16)) ? 1 : 0;
    return d;
  }
  __device__ int in_degree(int u, const u64 (&known)[W], const u64 (&up)[W]) const {
    int d = 0;
    for (int j = irp[u]; j < irp[u + 1]; ++j) d += allowed(known, up, (int)(iarc[j] >> 16)) ? 1 : 0;
    return d;
  }
  __device__ int level_of(int u) const { return lvb[(u >> 2) * 4 * stride + (u & 3)]; }
  __device__ void set_level(int u, int l) const {
    lvb[(u >> 2) * 4 * stride + (u & 3)] = (unsigned char)l;
  }
  __device__ static bool allowed(const u64 (&known)[W], const u64 (&up)[W], int e) {
    const int wi = e >> 6;
    const u64 b = 1ull << (e & 63);
    return !((pick<W>(known, wi) & b) && !(pick<W>(up, wi) & b));
  }

  /* Pushes the path's UNKNOWN hops on uh[] in dst->src order; returns their
   * number, -1 if dst is unreachable, -2 on an internal inconsistency. */
  /* K-terminal form (see Expander::find_tree) on the CSR arcs */
  __device__ int find_tree(const u64 (&known)[W], const u64 (&up)[W]) const {
    for (int i = 0; i < (n + 3) / 4; ++i) *(u32*)(lvb + i * 4 * stride) = 0xffffffffu;
    const int root = terms[0];
    u64 vis[NW], F[NW], need[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) vis[k] = F[k] = 0;
    bs_set<NW>(vis, root);
    bs_set<NW>(F, root);
    set_level(root, 0);
    bool left = false;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      need[k] = tmask[k] & ~vis[k];
      left |= need[k] != 0;
    }
    int L = 0;
    while (left) {
      u64 nx[NW];
#pragma unroll
      for (int k = 0; k < NW; ++k) nx[k] = 0;
      bool any = false;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        u64 fw = F[k];
        while (fw) {
          const int u = k * 64 + __builtin_ctzll(fw);
          fw &= fw - 1;
          for (int j = rp[u]; j < rp[u + 1]; ++j) {
            const u32 x = arc[j];
            const int v = (int)(x & 0xffffu), e = (int)(x >> 16);
            if (bs_test<NW>(vis, v) || !allowed(known, up, e)) continue;
            bs_set<NW>(vis, v);
            bs_set<NW>(nx, v);
            any = true;
          }
        }
      }
      if (!any) return -1;
      ++L;
      if (L >= 255) return -2;
      left = false;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        u64 m = nx[k];
        while (m) {
          set_level(k * 64 + __builtin_ctzll(m), L);
          m &= m - 1;
        }
        F[k] = nx[k];
        need[k] &= ~nx[k];
        left |= need[k] != 0;
      }
    }
    u64 marked[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) marked[k] = 0;
    bs_set<NW>(marked, root);
    int k = 0, len = 0;
    for (int i = 1; i < nterms; ++i) {
      int v = terms[i];
      while (!bs_test<NW>(marked, v)) {
        const int l = level_of(v);
        int pu = -1, pe = -1;
        for (int j = irp[v]; j < irp[v + 1]; ++j) {
          const u32 x = iarc[j];
          const int u = (int)(x & 0xffffu), e = (int)(x >> 16);
          if (level_of(u) == l - 1 && allowed(known, up, e)) {
            pu = u;
            pe = e;
            break;
          }
        }
        if (pu < 0) return -2;
        if (!((pick<W>(known, pe >> 6) >> (pe & 63)) & 1ull)) uh[(k++) * stride] = (u16)pe;
        ++len;
        bs_set<NW>(marked, v);
        v = pu;
      }
    }
    plen = len;
    return k;
  }

  __device__ int find_item(const u64 (&known)[W], const u64 (&up)[W]) const {
    if (nterms >= 2) return find_tree(known, up);
    if (src == dst) {
      plen = 0;
      return 0;
    }
    for (int i = 0; i < (n + 3) / 4; ++i) *(u32*)(lvb + i * 4 * stride) = 0xffffffffu;
    u64 vis[NW], F[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) vis[k] = F[k] = 0;
    bs_set<NW>(vis, src);
    bs_set<NW>(F, src);
    set_level(src, 0);
    int L = 0;
    for (;;) {
      u64 nx[NW];
#pragma unroll
      for (int k = 0; k < NW; ++k) nx[k] = 0;
      bool any = false, hit = false;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        u64 fw = F[k];
        while (fw && !hit) {
          const int u = k * 64 + __builtin_ctzll(fw);
          fw &= fw - 1;
          const int j1 = rp[u + 1];
          for (int j = rp[u]; j < j1; ++j) {
            const u32 x = arc[j];
            const int v = (int)(x & 0xffffu), e = (int)(x >> 16);
            if (bs_test<NW>(vis, v) || !allowed(known, up, e)) continue;
            bs_set<NW>(vis, v);
            bs_set<NW>(nx, v);
            any = true;
            if (v == dst) {
              hit = true;
              break;
            }
          }
        }
      }
      if (!any) return -1;
      ++L;
      if (hit) break;
      if (L >= 255) return -2;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        u64 m = nx[k];
        while (m) {
          set_level(k * 64 + __builtin_ctzll(m), L);
          m &= m - 1;
        }
        F[k] = nx[k];
      }
    }
    plen = L;
    int v = dst, k = 0;
    for (int l = L; l >= 1; --l) {
      int pu = -1, pe = -1;
      const int j1 = irp[v + 1];
      for (int j = irp[v]; j < j1; ++j) {
        const u32 x = iarc[j];
        const int u = (int)(x & 0xffffu), e = (int)(x >> 16);
        if (level_of(u) == l - 1 && allowed(known, up, e)) {
          pu = u;
          pe = e;
          break;
        }
      }
      if (pu < 0) return -2;
      if (!((pick<W>(known, pe >> 6) >> (pe & 63)) & 1ull)) uh[(k++) * stride] = (u16)pe;
      v = pu;
    }
    return k;
  }
};

/* The alternating-sweep rule (AltExpander's search and sweep) on the CSR
 * arcs of an undirected graph of up to 256 nodes: node sets of NW words in
 * VGPRs, levels as bytes.  Closures follow WORKING arcs from a work set; the
 * walk back takes, over the nodes v of K in ascending order, the first
 * in-arc (rows sorted by (u, e)) from the previous level, and keeps the
 * lowest u (the lowest v on ties): lowest u, then the lowest node of K it
 * reaches, then the lowest edge id. */
template <int W, int NW>
struct CsrAltExpander : CsrExpander<W, NW> {
  int period;
  const unsigned char* lqp; /* choice tables (global memory) */
  __device__ void bind_graph(unsigned char* smem, const Layout& L, const DevGraph& g, int slot) {
    CsrExpander<W, NW>::bind_graph(smem, L, g, slot);
    period = g.period > 0 ? g.period : 2;
    lqp = g.lqp;
  }
  __device__ static bool working(const u64 (&known)[W], const u64 (&up)[W], int e) {
    const int wi = e >> 6;
    const u64 b = 1ull << (e & 63);
    return (pick<W>(known, wi) & pick<W>(up, wi) & b) != 0;
  }
  /* extend x (and vis) over WORKING arcs from the nodes of x */
  __device__ __forceinline__ void close(u64 (&x)[NW], u64 (&vis)[NW], const u64 (&known)[W],
                                        const u64 (&up)[W]) const {
    u64 todo[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) todo[k] = x[k];
    for (;;) {
      int u = -1;
#pragma unroll
      for (int k = 0; k < NW; ++k)
        if (u < 0 && todo[k]) {
          u = k * 64 + __builtin_ctzll(todo[k]);
          todo[k] &= todo[k] - 1;
        }
      if (u < 0) break;
      const int j1 = this->rp[u + 1];
      for (int j = this->rp[u]; j < j1; ++j) {
        const u32 a = this->arc[j];
        const int v = (int)(a & 0xffffu), e = (int)(a >> 16);
        if (bs_test<NW>(vis, v) || !working(known, up, e)) continue;
        bs_set<NW>(vis, v);
        bs_set<NW>(x, v);
        bs_set<NW>(todo, v);
      }
    }
  }
  __device__ int find_item(const u64 (&known)[W], const u64 (&up)[W]) const {
    return find_alt<false>(known, up, nullptr);
  }
  /* CH: as AltExpander::find_alt */
  template <bool CH>
  __device__ int find_alt(const u64 (&known)[W], const u64 (&up)[W], ChoiceRaw* raw) const {
    const int n = this->n, src = this->src, dst = this->dst;
    if (src == dst) {
      this->plen = 0;
      return 0;
    }
    for (int i = 0; i < (n + 3) / 4; ++i) *(u32*)(this->lvb + i * 4 * this->stride) = 0xffffffffu;
    u64 vis[NW], F[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) vis[k] = F[k] = 0;
    bs_set<NW>(vis, src);
    bs_set<NW>(F, src);
    close(F, vis, known, up);
#pragma unroll
    for (int k = 0; k < NW; ++k)
      for (u64 w = F[k]; w; w &= w - 1) this->set_level(k * 64 + __builtin_ctzll(w), 0);
    int L = 0;
    while (!bs_test<NW>(vis, dst)) {
      u64 nx[NW];
      bool any = false;
#pragma unroll
      for (int k = 0; k < NW; ++k) nx[k] = 0;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        u64 fw = F[k];
        while (fw) {
          const int u = k * 64 + __builtin_ctzll(fw);
          fw &= fw - 1;
          const int j1 = this->rp[u + 1];
          for (int j = this->rp[u]; j < j1; ++j) {
            const u32 a = this->arc[j];
            const int v = (int)(a & 0xffffu), e = (int)(a >> 16);
            if (bs_test<NW>(vis, v) || !this->allowed(known, up, e)) continue;
            bs_set<NW>(vis, v);
            bs_set<NW>(nx, v);
            any = true;
          }
        }
      }
      if (!any) return -1;
      ++L;
      if (L >= 255) return -2;
      close(nx, vis, known, up);
#pragma unroll
      for (int k = 0; k < NW; ++k)
        for (u64 w = nx[k]; w; w &= w - 1) this->set_level(k * 64 + __builtin_ctzll(w), L);
#pragma unroll
      for (int k = 0; k < NW; ++k) F[k] = nx[k];
    }
    this->plen = L;
    u64 K[NW], none[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) K[k] = none[k] = 0;
    bs_set<NW>(K, dst);
    bs_set<NW>(none, dst);
    close(K, none, known, up);
    for (int l = L; l >= 1; --l) {
      int bu = n, be = -1;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        u64 kw = K[k];
        while (kw) {
          const int v = k * 64 + __builtin_ctzll(kw);
          kw &= kw - 1;
          const int j1 = this->irp[v + 1];
          for (int j = this->irp[v]; j < j1; ++j) {
            const u32 a = this->iarc[j];
            const int u = (int)(a & 0xffffu), e = (int)(a >> 16);
            if (u >= bu) break;
            if (this->level_of(u) == l - 1 && this->allowed(known, up, e)) {
              bu = u;
              be = e;
              break;
            }
          }
        }
      }
      if (be < 0) return -2;
      this->uh[(L - l) * this->stride] = (u16)be;
#pragma unroll
      for (int k = 0; k < NW; ++k) K[k] = none[k] = 0;
      bs_set<NW>(K, bu);
      bs_set<NW>(none, bu);
      close(K, none, known, up);
    }
    int f = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) f += __popcll(known[i] & ~up[i]);
    if (CH) {
      /* K is src's WORKING component; dst's is closed anew.  Arcs leaving a
       * component: allowed ones (UNKNOWN) and all of them. */
      auto cuts = [&](const u64 (&C)[NW], int& nodes, int& cut, int& out) {
        nodes = cut = out = 0;
#pragma unroll
        for (int k = 0; k < NW; ++k) {
          nodes += __popcll(C[k]);
          u64 w = C[k];
          while (w) {
            const int x = k * 64 + __builtin_ctzll(w);
            w &= w - 1;
            const int j1 = this->rp[x + 1];
            for (int j = this->rp[x]; j < j1; ++j) {
              const u32 a = this->arc[j];
              if (bs_test<NW>(C, (int)(a & 0xffffu))) continue;
              ++out;
              if (this->allowed(known, up, (int)(a >> 16))) ++cut;
            }
          }
        }
      };
      cuts(K, raw->nS, raw->cutS, raw->outS);
      u64 T[NW];
#pragma unroll
      for (int k = 0; k < NW; ++k) T[k] = none[k] = 0;
      bs_set<NW>(T, dst);
      bs_set<NW>(none, dst);
      close(T, none, known, up);
      cuts(T, raw->nT, raw->cutT, raw->outT);
      int kn = 0, lp = 0;
#pragma unroll
      for (int i = 0; i < W; ++i) kn += __popcll(known[i]);
      for (int i = 0; i < L; ++i) lp += lqp[this->uh[i * this->stride]];
      raw->k = L;
      raw->f = f;
      raw->u = this->ne - kn;
      raw->lp = lp;
      raw->e_dst = L > 0 ? this->uh[0] : 0;
      raw->e_src = L > 0 ? this->uh[(L - 1) * this->stride] : 0;
      return L;
    }
    if ((f / period) & 1) reverse_hops(L);
    return L;
  }
  __device__ void reverse_hops(int L) const {
    for (int i = 0, j = L - 1; i < j; ++i, --j) {
      const u16 t = this->uh[i * this->stride];
      this->uh[i * this->stride] = this->uh[j * this->stride];
      this->uh[j * this->stride] = t;
    }
  }
};

/* LDS layout of the lean row-form expander: staged graph tables, then the
 * kernel's extra region, per-lane node levels (bytes) and hop stack. */
inline Layout make_layout_lean(int n, int e, int w, int block, int extra_bytes = 0) {
  Layout L = make_layout(n, e, w, block, 0, extra_bytes, true);
  int o = L.off_extra + extra_bytes;
  o = (o + 15) / 16 * 16;
  L.off_lvl = o;
  L.off_rows = o;
  o += 4 * ((n + 3) / 4) * block;
  o = (o + 15) / 16 * 16;
  L.off_uh = o;
  o += 2 * n * block;
  L.total_bytes = (o + 15) / 16 * 16;
  return L;
}

/* Row-form graphs (undirected, N <= 64) with items that carry no rows: the
 * graph's tables are staged in LDS and each node's allowed neighbours are
 * derived from the item's FAILED plane when the search reaches it
 * (Expander::adj_of); node levels are bytes.  Used for the plain path rule;
 * the same search as Expander::find (Expander::find_lean). */
template <int W, typename NS>
struct LeanExpander : Expander<W, NS> {
  static constexpr bool kStage = true;
  static constexpr int kRowBytes = 0;
  static constexpr bool kSettle = false;
  unsigned char* lvb;
  __device__ void bind_graph(unsigned char* smem, const Layout& L, const DevGraph& g, int slot) {
    Expander<W, NS>::bind_graph(smem, L, g, slot);
    this->rows = nullptr;
    this->lvl = nullptr;
    lvb = smem + L.off_lvl + 4 * slot;
  }
  __device__ void load(const QBuf&, u64, u64) const {}
  __device__ void derive(const u64 (&)[W], const u64 (&)[W]) const {}
  __device__ void store_child(const QBuf&, u64, u64, int) const {}
  __device__ void store_child_of(const QBuf&, u64, u64, int, int) const {}
  __device__ void store_rows(const QBuf&, u64, u64) const {}
  __device__ void carry(const QBuf&, u64, const QBuf&, u64, u64) const {}
  __device__ int find_item(const u64 (&known)[W], const u64 (&up)[W]) const {
    return this->find_lean(known, up, lvb);
  }
};

/* The alternating-sweep rule on the lean row form (items carry no rows).
 * Path: fewest UNKNOWN edges.  Node sets are closed under WORKING edges
 * (close()): level 0 is src's WORKING component, level l+1 the unlevelled
 * allowed neighbours of level l with their WORKING components, until dst
 * is levelled; an allowed edge between consecutive levels is therefore
 * UNKNOWN.  The walk back starts from dst's component K: the parent is the
 * lowest-index node u on the previous level with an allowed edge into K, the
 * hop is u's edge to the lowest-index node of K it reaches, and K becomes
 * u's component.  Hops are pushed dst side first, as by find(); the stack is
 * then reversed when (FAILED edges / period) is odd, so the children are
 * formed from the dst end.  Node levels are bytes in the lane's slice, as in
 * find_lean(). */
template <int W, typename NS>
struct AltExpander : LeanExpander<W, NS> {
  int period;
  const unsigned char* lqp; /* choice tables (global memory) */
  __device__ void bind_graph(unsigned char* smem, const Layout& L, const DevGraph& g, int slot) {
    LeanExpander<W, NS>::bind_graph(smem, L, g, slot);
    period = g.period > 0 ? g.period : 2;
    lqp = g.lqp;
  }
  __device__ int find_item(const u64 (&known)[W], const u64 (&up)[W]) const {
    return find_alt(known, up);
  }
  /* allowed degree of u (items carry no rows) */
  __device__ int out_degree(int u, const u64 (&known)[W], const u64 (&up)[W]) const {
    u64 Dn[W];
#pragma unroll
    for (int i = 0; i < W; ++i) Dn[i] = known[i] & ~up[i];
    const NS a = this->adj_of(u, Dn);
    return sizeof(NS) == 4 ? __popc((u32)a) : __popcll((u64)a);
  }
  __device__ int in_degree(int u, const u64 (&k)[W], const u64 (&p)[W]) const {
    return out_degree(u, k, p);
  }

  /* CH: the learned chooser picks the sweep direction; the search records
   * its ingredients in *raw and leaves the hops in the forward sweep. */
  template <bool CH = false>
  __device__ int find_alt(const u64 (&known)[W], const u64 (&up)[W],
                          ChoiceRaw* raw = nullptr) const {
    const int n = this->n;
    const Pow2 stride = this->stride;
    if (this->src == this->dst) {
      this->plen = 0;
      return 0;
    }
    u64 Dn[W];
    int f = 0;
    NS wn = 0; /* nodes with a WORKING edge */
#pragma unroll
    for (int i = 0; i < W; ++i) {
      Dn[i] = known[i] & ~up[i];
      f += __popcll(Dn[i]);
      u64 m = up[i];
      while (m) {
        const u32 x = this->ep[i * 64 + __builtin_ctzll(m)];
        m &= m - 1;
        wn |= ((NS)1 << (x & 0xffffu)) | ((NS)1 << (x >> 16));
      }
    }
    unsigned char* lv = this->lvb;
    for (int i = 0; i < (n + 3) / 4; ++i) *(u32*)(lv + i * 4 * stride) = 0xffffffffu;
    auto setl = [&](NS m, int l) {
      while (m) {
        const int u = ctzs<NS>(m);
        m &= m - 1;
        lv[(u >> 2) * 4 * stride + (u & 3)] = (unsigned char)l;
      }
    };
    auto getl = [&](int u) { return (int)lv[(u >> 2) * 4 * stride + (u & 3)]; };
    auto close = [&](NS x) {
      NS todo = x & wn;
      while (todo) {
        const int u = ctzs<NS>(todo);
        todo &= todo - 1;
        const NS add = this->working_row(u, up) & ~x;
        x |= add;
        todo |= add;
      }
      return x;
    };
    const NS dbit = (NS)1 << this->dst;
    NS vis = close((NS)1 << this->src);
    setl(vis, 0);
    int L = 0;
    if (!(vis & dbit)) {
      NS F = vis;
      for (;;) {
        NS nx = 0, fr = F;
        while (fr) {
          const int u = ctzs<NS>(fr);
          fr &= fr - 1;
          nx |= this->adj_of(u, Dn);
        }
        nx &= ~vis;
        if (!nx) return -1;
        ++L;
        nx = close(nx);
        setl(nx, L);
        vis |= nx;
        if (nx & dbit) break;
        F = nx;
      }
    }
    this->plen = L;
    NS K = close(dbit);
    for (int l = L; l >= 1; --l) {
      int ub = 64;
      NS kk = K;
      while (kk) {
        const int v = ctzs<NS>(kk);
        kk &= kk - 1;
        NS a = this->adj_of(v, Dn) & ~K;
        while (a) {
          const int c = ctzs<NS>(a);
          if (c >= ub) break;
          a &= a - 1;
          if (getl(c) == l - 1) {
            ub = c;
            break;
          }
        }
      }
      if (ub >= n) return -2;
      const NS into = this->adj_of(ub, Dn) & K;
      if (!into) return -2;
      const int v = ctzs<NS>(into);
      this->uh[(L - l) * stride] = this->eidm[ub * n + v];
      K = close((NS)1 << ub);
    }
    if (CH) {
      /* K is src's WORKING component again; dst's is closed anew.  Edges
       * leaving a component: allowed ones (UNKNOWN, the component being
       * closed under WORKING edges) and all of them (simple graph). */
      const NS T = close(dbit);
      auto cuts = [&](NS C, int& nodes, int& cut, int& out) {
        nodes = sizeof(NS) == 4 ? __popc((u32)C) : __popcll((u64)C);
        cut = out = 0;
        NS c = C;
        while (c) {
          const int x = ctzs<NS>(c);
          c &= c - 1;
          const NS a = this->adj_of(x, Dn) & ~C, st = (NS)this->sadj[x] & ~C;
          cut += sizeof(NS) == 4 ? __popc((u32)a) : __popcll((u64)a);
          out += sizeof(NS) == 4 ? __popc((u32)st) : __popcll((u64)st);
        }
      };
      cuts(K, raw->nS, raw->cutS, raw->outS);
      cuts(T, raw->nT, raw->cutT, raw->outT);
      int kn = 0, lp = 0;
#pragma unroll
      for (int i = 0; i < W; ++i) kn += __popcll(known[i]);
      for (int i = 0; i < L; ++i) lp += lqp[this->uh[i * stride]];
      raw->k = L;
      raw->f = f;
      raw->u = this->ne - kn;
      raw->lp = lp;
      raw->e_dst = L > 0 ? this->uh[0] : 0;
      raw->e_src = L > 0 ? this->uh[(L - 1) * stride] : 0;
      return L;
    }
    if ((f / period) & 1) reverse_hops(L);
    return L;
  }
  __device__ void reverse_hops(int L) const {
    const Pow2 stride = this->stride;
    for (int i = 0, j = L - 1; i < j; ++i, --j) {
      const u16 t = this->uh[i * stride];
      this->uh[i * stride] = this->uh[j * stride];
      this->uh[j * stride] = t;
    }
  }
};

/* Wave-wide inclusive prefix sum of a small non-negative int. */
__device__ inline int wave_incl_scan(int x) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int t = __shfl_up(x, off, 64);
    if (lane >= off) x += t;
  }
  return x;
}

__device__ inline void wave_sum_dd(double& s, double& c) {
  /* combine (s, c) pairs across the wave with error-free additions */
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    double os = __shfl_xor(s, off, 64);
    double oc = __shfl_xor(c, off, 64);
    double t, e;
    two_sum(s, os, t, e);
    s = t;
    c = c + oc + e;
  }
}

__device__ inline u64 wave_sum_u64(u64 x) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

/* Merge a wave's (s, c) into a global double-double slot (slot owned by this
 * wave for the duration of the launch). */
__device__ inline void dd_merge_slot(DD* slot, double s, double c) {
  DD o = *slot;
  double t, e;
  two_sum(o.hi, s, t, e);
  double lo = o.lo + c + e;
  double hi2, e2;
  two_sum(t, lo, hi2, e2);
  slot->hi = hi2;
  slot->lo = e2;
}

}  // namespace nr

#endif
