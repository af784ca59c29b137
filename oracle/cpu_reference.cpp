/*
 * cpu_reference.cpp - CPU oracle for two-terminal network reliability by
 * path decomposition (and edge-pivot factoring).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
 * the smoke entry point and the benchmark's CPU-baseline leg use it.
 *
 * Written from scratch as a sequential restatement of the reference's rules;
 * every function names the reference lines it follows.  Differences from the
 * reference that are deliberate (FP64 instead of FP32 log space, the
 * deterministic parent rule for N > 32, src == dst giving R = 1, the success
 * term added whatever the state of the last hop) are stated where they occur.
 *
 * Edge states are two bit planes per 64 edges: known/up.
 *   UNKNOWN = (0,0)   WORKING/up = (1,1)   FAILED/down = (1,0)
 * (the reference packs the same three states as 2-bit codes 0b11/0b01/0b00,
 * workingPad/common.h:22-24).
 */
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

using u64 = uint64_t;
using Clock = std::chrono::steady_clock;

/* Neumaier compensated sum. */
struct KSum {
  double s = 0.0, c = 0.0;
  void add(double x) {
    double t = s + x;
    if (std::fabs(s) >= std::fabs(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
  }
  void merge(const KSum& o) {
    add(o.s);
    add(o.c);
  }
  double value() const { return s + c; }
};

}  // namespace

struct orc_graph {
  int n = 0, e = 0, nnz = 0, header_e = 0, type = 0;
  std::vector<int> row_ptr, col_idx, edge_id;
  std::vector<double> p, q;
  /* in-adjacency (u, eid) per v, sorted by (u, eid) ascending; used to find
   * the lowest-index parent during the backward walk */
  std::vector<int> in_ptr, in_src, in_eid;
  int w = 0;   /* edge words per plane */
  int nw = 0;  /* node words */
};

namespace {

/* ------------------------------------------------------------------ */
/* Graph loading: restates workingPad/graph.h:46-145.                  */
/*  - header "N E [T]"; T optional (graph.h:53 requires it, the README  */
/*    test graphs lack it, workingPad/test_simple.txt:1)                */
/*  - per node: degree then (nbr, p) pairs (graph.h:61-72); p parsed    */
/*    as FP64 (graph.h parses float)                                    */
/*  - edge id = order of first appearance of (min,max) scanning u      */
/*    ascending in file order (graph.h:76-90)                           */
/*  - CSR rows in file order (graph.h:104-130)                          */
/* ------------------------------------------------------------------ */
struct Cursor {
  const char* s;
  const char* end;
  void skip() {
    while (s < end && (*s == ' ' || *s == '\t' || *s == '\r' || *s == '\n' ||
                       *s == '\v' || *s == '\f'))
      ++s;
  }
  bool read_int(long* out) {
    skip();
    if (s >= end) return false;
    char* e = nullptr;
    long v = std::strtol(s, &e, 10);
    if (e == s) return false;
    s = e;
    *out = v;
    return true;
  }
  bool read_double(double* out) {
    skip();
    if (s >= end) return false;
    char* e = nullptr;
    double v = std::strtod(s, &e);
    if (e == s) return false;
    s = e;
    *out = v;
    return true;
  }
  int peek_alpha() {
    skip();
    if (s < end && ((*s >= 'A' && *s <= 'Z') || (*s >= 'a' && *s <= 'z')))
      return *s;
    return 0;
  }
  void skip_token() {
    while (s < end && !(*s == ' ' || *s == '\t' || *s == '\r' || *s == '\n'))
      ++s;
  }
};

void finish_graph(orc_graph* g) {
  g->w = (g->e + 63) / 64;
  if (g->w == 0) g->w = 1;
  g->nw = (g->n + 63) / 64;
  if (g->nw == 0) g->nw = 1;
  g->q.resize(g->e);
  for (int i = 0; i < g->e; ++i) g->q[i] = 1.0 - g->p[i];
  /* in-adjacency sorted by (u, eid) */
  std::vector<std::vector<std::pair<int, int>>> in(g->n);
  for (int u = 0; u < g->n; ++u)
    for (int j = g->row_ptr[u]; j < g->row_ptr[u + 1]; ++j)
      in[g->col_idx[j]].push_back({u, g->edge_id[j]});
  g->in_ptr.assign(g->n + 1, 0);
  g->in_src.clear();
  g->in_eid.clear();
  for (int v = 0; v < g->n; ++v) {
    std::sort(in[v].begin(), in[v].end());
    for (auto& pr : in[v]) {
      g->in_src.push_back(pr.first);
      g->in_eid.push_back(pr.second);
    }
    g->in_ptr[v + 1] = (int)g->in_src.size();
  }
}

int load_text(const char* text, size_t len, orc_graph* g) {
  Cursor c{text, text + len};
  long n = 0, ef = 0;
  if (!c.read_int(&n) || !c.read_int(&ef)) return ORC_ERR_PARSE;
  if (n <= 0) return ORC_ERR_PARSE;
  int type = c.peek_alpha();
  if (type) c.skip_token();
  std::vector<std::vector<std::pair<int, double>>> adj(n);
  for (long i = 0; i < n; ++i) {
    long deg = 0;
    if (!c.read_int(&deg) || deg < 0) return ORC_ERR_PARSE;
    adj[i].resize(deg);
    for (long j = 0; j < deg; ++j) {
      long nb = 0;
      double p = 0.0;
      if (!c.read_int(&nb) || !c.read_double(&p)) return ORC_ERR_PARSE;
      if (nb < 0 || nb >= n) return ORC_ERR_RANGE;
      adj[i][j] = {(int)nb, p};
    }
  }
  g->n = (int)n;
  g->type = type;
  /* unique ids by first appearance of (min, max) (graph.h:76-90) */
  std::unordered_map<u64, int> ids;
  g->p.clear();
  for (int u = 0; u < g->n; ++u)
    for (auto& en : adj[u]) {
      u64 a = (u64)std::min(u, en.first), b = (u64)std::max(u, en.first);
      u64 key = a * (u64)g->n + b;
      if (ids.find(key) == ids.end()) {
        int id = (int)ids.size();
        ids[key] = id;
        g->p.push_back(en.second);
      }
    }
  g->e = (int)ids.size();
  /* header E is halved for undirected files (graph.h:93) */
  g->header_e = (type == 'U') ? (int)(ef / 2) : (int)ef;
  g->row_ptr.assign(g->n + 1, 0);
  for (int u = 0; u < g->n; ++u) g->row_ptr[u + 1] = g->row_ptr[u] + (int)adj[u].size();
  g->nnz = g->row_ptr[g->n];
  g->col_idx.resize(g->nnz);
  g->edge_id.resize(g->nnz);
  for (int u = 0; u < g->n; ++u) {
    int k = g->row_ptr[u];
    for (auto& en : adj[u]) {
      u64 a = (u64)std::min(u, en.first), b = (u64)std::max(u, en.first);
      g->col_idx[k] = en.first;
      g->edge_id[k] = ids[a * (u64)g->n + b];
      ++k;
    }
  }
  finish_graph(g);
  return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* The search engine.                                                  */
/* ------------------------------------------------------------------ */
constexpr int kMaxW = 16;  /* up to 1024 edges */
constexpr int kMaxNW = 4;  /* up to 256 nodes */

struct NodeSet {
  u64 w[kMaxNW];
};

struct Engine {
  const orc_graph* g;
  int src, dst, W, NW, mode;
  double eps;
  const int* terms = nullptr; /* K-terminal mode when nterms >= 2 */
  int nterms = 0;
  int period = 2; /* ORC_MODE_ALTERNATE: sweep flips every `period` FAILED edges */
  /* ORC_MODE_LEARNED: integer chooser weights and signs; per-edge steps
   * round(-4 log2 p), round(-4 log2 q) clamped to 0..63 */
  const int* chooser = nullptr;
  const unsigned char* lqp = nullptr;
  const unsigned char* lqq = nullptr;

  bool allowed(const u64* known, const u64* up, int e) const {
    u64 bit = 1ull << (e & 63);
    int k = e >> 6;
    return !((known[k] & ~up[k]) & bit);
  }
  bool working(const u64* known, const u64* up, int e) const {
    u64 bit = 1ull << (e & 63);
    int k = e >> 6;
    return (known[k] & up[k] & bit) != 0;
  }
  static bool test(const NodeSet& s, int v) { return (s.w[v >> 6] >> (v & 63)) & 1; }
  static void set(NodeSet& s, int v) { s.w[v >> 6] |= 1ull << (v & 63); }

  /* Level-synchronous BFS from src over allowed edges (WORKING, or UNKNOWN
   * when optimistic), restating device_code.cuh:80-194.  On success returns
   * the path length L >= 0 and fills hops[0..L-1] with the edge ids of the
   * path in src->dst order.  Returns -1 when dst is unreachable.
   *
   * Parent rule ("lowest-index shortest path"): the parent of v is the
   * lowest-index node u on the previous BFS level with an allowed edge u->v.
   * This is what device_code.cuh:141-171 computes when N <= 32 (one lane
   * scans the frontier in ascending order and the first discoverer wins);
   * for N > 32 the reference is racy and this rule defines it. */
  /* K-terminal generalisation of find_path (no counterpart in the
   * reference, whose target_set is declared but unused: original/Rel_16M.cu
   * :43, original/Seq_16.cpp:33).  BFS from terms[0] over allowed edges,
   * whole levels, until every terminal is reached (-1 if the frontier dies
   * first).  Then each other terminal, in list order, is walked back to the
   * tree built so far under the same parent rule (lowest-index node of the
   * previous level, lowest edge id on ties); the walks' edges, reversed,
   * are the tree's hops, so for terms = {src, dst} this is find_path. */
  int find_tree(const u64* known, const u64* up, bool optimistic, int* hops,
                std::vector<NodeSet>& levels) const {
    auto ok = [&](int e) {
      return optimistic ? allowed(known, up, e) : working(known, up, e);
    };
    const int root = terms[0];
    std::vector<int> lv(g->n, -1);
    NodeSet vis{}, F{};
    set(vis, root);
    set(F, root);
    lv[root] = 0;
    if ((int)levels.size() < g->n + 1) levels.resize(g->n + 1);
    levels[0] = F;
    auto all_in = [&]() {
      for (int i = 0; i < nterms; ++i)
        if (!test(vis, terms[i])) return false;
      return true;
    };
    int L = 0;
    while (!all_in()) {
      NodeSet next{};
      bool any = false;
      for (int k = 0; k < NW; ++k) {
        u64 fw = F.w[k];
        while (fw) {
          int u = k * 64 + __builtin_ctzll(fw);
          fw &= fw - 1;
          for (int j = g->row_ptr[u]; j < g->row_ptr[u + 1]; ++j) {
            int v = g->col_idx[j];
            if (test(vis, v) || !ok(g->edge_id[j])) continue;
            set(vis, v);
            set(next, v);
            lv[v] = L + 1;
            any = true;
          }
        }
      }
      if (!any) return -1;
      levels[++L] = next;
      F = next;
    }
    std::vector<char> marked(g->n, 0);
    marked[root] = 1;
    std::vector<int> seq;
    for (int i = 1; i < nterms; ++i) {
      int v = terms[i];
      while (!marked[v]) {
        const int l = lv[v];
        int pu = -1, pe = -1;
        for (int k = g->in_ptr[v]; k < g->in_ptr[v + 1]; ++k) {
          int u = g->in_src[k];
          if (lv[u] == l - 1 && ok(g->in_eid[k])) {
            pu = u;
            pe = g->in_eid[k];
            break;
          }
        }
        seq.push_back(pe);
        marked[v] = 1;
        v = pu;
      }
    }
    const int m = (int)seq.size();
    for (int h = 0; h < m; ++h) hops[h] = seq[m - 1 - h];
    return m;
  }

  int find_path(const u64* known, const u64* up, bool optimistic, int* hops,
                std::vector<NodeSet>& levels) const {
    if (nterms >= 2) return find_tree(known, up, optimistic, hops, levels);
    if (src == dst) return 0;
    auto ok = [&](int e) {
      return optimistic ? allowed(known, up, e) : working(known, up, e);
    };
    NodeSet vis{}, F{};
    set(vis, src);
    set(F, src);
    if ((int)levels.size() < g->n + 1) levels.resize(g->n + 1);
    levels[0] = F;
    int L = 0;
    for (;;) {
      NodeSet next{};
      bool any = false;
      for (int k = 0; k < NW; ++k) {
        u64 fw = F.w[k];
        while (fw) {
          int u = k * 64 + __builtin_ctzll(fw);
          fw &= fw - 1;
          for (int j = g->row_ptr[u]; j < g->row_ptr[u + 1]; ++j) {
            int v = g->col_idx[j];
            if (test(vis, v) || !ok(g->edge_id[j])) continue;
            set(vis, v);
            set(next, v);
            any = true;
          }
        }
      }
      if (!any) return -1;
      levels[++L] = next;
      if (test(next, dst)) break;
      F = next;
    }
    int v = dst;
    for (int l = L; l >= 1; --l) {
      int pu = -1, pe = -1;
      for (int k = g->in_ptr[v]; k < g->in_ptr[v + 1]; ++k) {
        int u = g->in_src[k];
        if (test(levels[l - 1], u) && ok(g->in_eid[k])) {
          pu = u;
          pe = g->in_eid[k];
          break; /* in-list is sorted by (u, eid): first match is the min */
        }
      }
      hops[l - 1] = pe;
      v = pu;
    }
    return L;
  }

  /* Path-mode expansion of one item, restating main.cu:114-283:
   * for the UNKNOWN hops u_1..u_k of the path in src->dst order, child i has
   * u_1..u_{i-1} WORKING and u_i FAILED, mass m*p_1*...*p_{i-1}*q_i (FP64,
   * multiplied left to right); the success term is m*p_1*...*p_k.
   * The reference adds the success term only when the last hop is UNKNOWN
   * (main.cu:254); here it is added for every live item, which differs only
   * when p = 1 edges were pre-contracted.
   * Returns the number of children, or -1 if dead. */
  int expand_path(const u64* known, const u64* up, double mass, u64* ck, u64* cu,
                  double* cm, double* succ, int* hops, int* nh,
                  std::vector<NodeSet>& levels) const {
    int L = find_path(known, up, true, hops, levels);
    if (nh) *nh = L;
    if (L < 0) return -1;
    u64 pre[kMaxW] = {0};
    double run = mass;
    int k = 0;
    for (int h = 0; h < L; ++h) {
      int e = hops[h];
      u64 bit = 1ull << (e & 63);
      int wi = e >> 6;
      if (known[wi] & bit) continue; /* WORKING hop (FAILED cannot be on it) */
      u64* kk = ck + (size_t)k * W;
      u64* uu = cu + (size_t)k * W;
      for (int i = 0; i < W; ++i) {
        kk[i] = known[i] | pre[i];
        uu[i] = up[i] | pre[i];
      }
      kk[wi] |= bit;
      cm[k] = run * g->q[e];
      run = run * g->p[e];
      pre[wi] |= bit;
      ++k;
    }
    *succ = run;
    return k;
  }

  /* Pivot-mode expansion, restating workingPad/cpu_reference.cpp:121-145:
   * optimistic BFS prune, confirmed (WORKING-only) BFS success, else branch
   * on the UNKNOWN edge with the highest p (lowest id on ties, strict '>' as
   * cpu_reference.cpp:112).  Returns -1 dead, 0 success (*succ = mass),
   * 2 children (WORKING first). */
  int expand_pivot(const u64* known, const u64* up, double mass, u64* ck, u64* cu,
                   double* cm, double* succ, int* hops,
                   std::vector<NodeSet>& levels) const {
    if (find_path(known, up, true, hops, levels) < 0) return -1;
    if (find_path(known, up, false, hops, levels) >= 0) {
      *succ = mass;
      return 0;
    }
    int best = -1;
    double bp = -1.0;
    for (int e = 0; e < g->e; ++e) {
      u64 bit = 1ull << (e & 63);
      if (!(known[e >> 6] & bit) && g->p[e] > bp) {
        bp = g->p[e];
        best = e;
      }
    }
    *succ = 0.0;
    if (best < 0) return -1;
    u64 bit = 1ull << (best & 63);
    int wi = best >> 6;
    for (int i = 0; i < W; ++i) {
      ck[i] = known[i];
      cu[i] = up[i];
      ck[W + i] = known[i];
      cu[W + i] = up[i];
    }
    ck[wi] |= bit;
    cu[wi] |= bit;
    ck[W + wi] |= bit;
    cm[0] = mass * g->p[best];
    cm[1] = mass * g->q[best];
    return 2;
  }
};

/* forkBest expansion, restating workingPad/main.cu:550-567 and 304-430:
 * a confirmed (WORKING-only) BFS success ends the item; otherwise the
 * optimistic path (lowest-index parent rule) is walked from dst towards src
 * (the reference keeps it dst-first, main.cu:332-341) and the pivot is the
 * UNKNOWN hop that the reference's warp argmax returns (main.cu:355-388):
 * keys are logf of p in FP32 (graph.h:138), and ties are settled by the
 * lane layout and the shuffle tree, restated step by step below.  With
 * terminal sets the hops are indexed in the order the tree walk meets them.
 * Children: the pivot WORKING (mass * p) then FAILED (mass * q).  Returns
 * -1 dead, 0 success, 2 children. */
int expand_forkbest_impl(const Engine& en, const u64* known, const u64* up, double mass,
                         u64* ck, u64* cu, double* cm, double* succ, int* hops,
                         std::vector<NodeSet>& levels) {
  if (en.find_path(known, up, false, hops, levels) >= 0) {
    *succ = mass;
    return 0;
  }
  const int L = en.find_path(known, up, true, hops, levels);
  if (L < 0) return -1;
  /* per-lane argmax over hops r = lane, lane + 32, ... counted from the dst
   * end (strict >, so a lane keeps its first hop on ties), then the
   * shfl_down tree over 32 lanes, where position i keeps its own value on
   * ties with position i + off: among lanes holding the best key the winner
   * is the one whose 5-bit reversed index is smallest. */
  float lane_key[32];
  int lane_edge[32];
  for (int l = 0; l < 32; ++l) { /* -1e30 in main.cu:356; -inf keeps a lane
                                    without an UNKNOWN hop below every key */
    lane_key[l] = -INFINITY;
    lane_edge[l] = -1;
  }
  for (int r = 0; r < L; ++r) {
    const int e = hops[L - 1 - r];
    const u64 bit = 1ull << (e & 63);
    if (known[e >> 6] & bit) continue;
    const float pf = (float)en.g->p[e]; /* graph.h:138 */
    const float key = pf > 0.0f ? logf(pf) : -1e30f;
    if (key > lane_key[r & 31]) {
      lane_key[r & 31] = key;
      lane_edge[r & 31] = e;
    }
  }
  for (int off = 16; off > 0; off >>= 1)
    for (int i = 0; i < off; ++i)
      if (lane_key[i + off] > lane_key[i]) {
        lane_key[i] = lane_key[i + off];
        lane_edge[i] = lane_edge[i + off];
      }
  const int best = lane_edge[0];
  if (best < 0) { /* all-WORKING path: unreachable after the confirmed BFS */
    *succ = mass;
    return 0;
  }
  const int W = en.W;
  const u64 bit = 1ull << (best & 63);
  const int wi = best >> 6;
  for (int i = 0; i < W; ++i) {
    ck[i] = known[i];
    cu[i] = up[i];
    ck[W + i] = known[i];
    cu[W + i] = up[i];
  }
  ck[wi] |= bit;
  cu[wi] |= bit;
  ck[W + wi] |= bit;
  cm[0] = mass * en.g->p[best];
  cm[1] = mass * en.g->q[best];
  *succ = 0.0;
  return 2;
}

/* Path decomposition on a path with the fewest UNKNOWN edges.  WORKING
 * edges are contracted: components of WORKING edges are labelled by their
 * lowest node; a level-synchronous BFS over components follows UNKNOWN
 * edges; the parent of a component is the lowest-label component of the
 * previous level joined to it by an UNKNOWN edge, through the lowest such
 * edge id.  The path's UNKNOWN edges (src side first) get the children of
 * main.cu:214-266; a path with none (src and dst already joined by WORKING
 * edges) is a success. */
int expand_minunk_impl(const Engine& en, const u64* known, const u64* up, double mass,
                       u64* ck, u64* cu, double* cm, double* succ, int* hops) {
  const orc_graph* g = en.g;
  const int n = g->n, W = en.W;
  std::vector<int> lab(n);
  for (int v = 0; v < n; ++v) lab[v] = v;
  /* label propagation to the minimum over WORKING edges */
  for (bool changed = true; changed;) {
    changed = false;
    for (int e = 0; e < g->e; ++e) {
      if (!en.working(known, up, e)) continue;
      int a = -1, b = -1;
      for (int u = 0; u < n && b < 0; ++u)
        for (int j = g->row_ptr[u]; j < g->row_ptr[u + 1]; ++j)
          if (g->edge_id[j] == e) {
            if (a < 0) a = u; else if (u != a) { b = u; break; }
          }
      if (b < 0) continue;
      int m = std::min(lab[a], lab[b]);
      if (lab[a] != m || lab[b] != m) {
        for (int v = 0; v < n; ++v)
          if (lab[v] == lab[a] || lab[v] == lab[b]) lab[v] = m;
        changed = true;
      }
    }
  }
  const int cs = lab[en.src], cd = lab[en.dst];
  if (cs == cd) {
    *succ = mass;
    return 0;
  }
  /* BFS over components via UNKNOWN edges */
  std::vector<int> lvl(n, -1), par(n, -1), pedge(n, -1);
  std::vector<int> frontier{cs}, next;
  lvl[cs] = 0;
  int L = 0;
  bool found = false;
  while (!frontier.empty() && !found) {
    next.clear();
    ++L;
    std::sort(frontier.begin(), frontier.end());
    for (int c : frontier)
      for (int u = 0; u < n; ++u) {
        if (lab[u] != c) continue;
        for (int j = g->row_ptr[u]; j < g->row_ptr[u + 1]; ++j) {
          const int e = g->edge_id[j], v = g->col_idx[j];
          if (!en.allowed(known, up, e) || en.working(known, up, e)) continue;
          const int cv = lab[v];
          if (lvl[cv] >= 0 && lvl[cv] < L) continue;
          if (lvl[cv] < 0) {
            lvl[cv] = L;
            next.push_back(cv);
            par[cv] = c;
            pedge[cv] = e;
          } else if (lvl[cv] == L && (c < par[cv] || (c == par[cv] && e < pedge[cv]))) {
            par[cv] = c;
            pedge[cv] = e;
          }
        }
      }
    if (lvl[cd] >= 0) found = true;
    frontier.swap(next);
  }
  if (!found) return -1;
  int k = 0;
  for (int c = cd; c != cs; c = par[c]) hops[k++] = pedge[c];
  std::reverse(hops, hops + k);
  u64 pre[kMaxW] = {0};
  double run = mass;
  for (int h = 0; h < k; ++h) {
    const int e = hops[h];
    const u64 bit = 1ull << (e & 63);
    const int wi = e >> 6;
    u64* kk = ck + (size_t)h * W;
    u64* uu = cu + (size_t)h * W;
    for (int i = 0; i < W; ++i) {
      kk[i] = known[i] | pre[i];
      uu[i] = up[i] | pre[i];
    }
    kk[wi] |= bit;
    cm[h] = run * g->q[e];
    run = run * g->p[e];
    pre[wi] |= bit;
  }
  *succ = run;
  return k;
}

/* Whether the exploring policy of ORC_MODE_EXPLORE flips the alternating
 * direction of this item: the top two bits of a splitmix64-finaliser chain
 * over its planes are zero (a quarter of the items). */
static bool explore_flip(const u64* known, const u64* up, int W) {
  auto fin = [](u64 z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  u64 h = 0x9E3779B97F4A7C15ull;
  for (int i = 0; i < W; ++i) {
    h = fin(h ^ known[i]);
    h = fin(h ^ up[i]);
  }
  return (h >> 62) == 0;
}

/* Learned sweep choice (ORC_MODE_LEARNED): the chooser's score of sweeping
 * the path hops[0..L) (src end first) from the dst end (from_dst) or the
 * src end, in units of 2^-8, computed with integers: hidden unit h is
 * relu(sum_k x_k a[k][h]), the score the signed sum of the units.  The
 * features x restate the engine's choice features from the item itself. */
static int chooser_score(const Engine& en, const std::vector<char>& S,
                         const std::vector<char>& T, const int* hops, int L, int f, int unk,
                         double mass, const u64* known, const u64* up, bool from_dst) {
  const orc_graph* g = en.g;
  auto leaving = [&](const std::vector<char>& C, int* nodes, int* unknown, int* failed) {
    *nodes = *unknown = *failed = 0;
    for (int u = 0; u < g->n; ++u) {
      if (!C[u]) continue;
      ++*nodes;
      for (int j = g->row_ptr[u]; j < g->row_ptr[u + 1]; ++j) {
        if (C[g->col_idx[j]]) continue;
        if (en.allowed(known, up, g->edge_id[j])) ++*unknown; else ++*failed;
      }
    }
  };
  int nS, uS, fS, nT, uT, fT;
  leaving(S, &nS, &uS, &fS);
  leaving(T, &nT, &uT, &fT);
  auto lim = [](int v) { return std::max(-63, std::min(63, v)); };
  auto half_down = [](int v) { return (int)std::floor(v / 2.0); };
  auto exponent = [](double m) {
    u64 b;
    std::memcpy(&b, &m, 8);
    const int e = (int)((b >> 52) & 0x7ff);
    return e == 0 ? -1075 : e - 1023;
  };
  const int lm = exponent(mass);
  /* headroom against eps, or 2^-64 below that, saturating at 32 half-binades */
  const int le = en.eps > std::ldexp(1.0, -64) ? exponent(en.eps) : -64;
  const int head = std::max(0, std::min(32, half_down(lm - le)));
  int lp = 0;
  for (int i = 0; i < L; ++i) lp += en.lqp[hops[i]];
  const int first = from_dst ? hops[L - 1] : hops[0];
  const int last = from_dst ? hops[0] : hops[L - 1];
  const int nA = from_dst ? nT : nS, uA = from_dst ? uT : uS, fA = from_dst ? fT : fS;
  const int nB = from_dst ? nS : nT, uB = from_dst ? uS : uT, fB = from_dst ? fS : fT;
  const int E = std::max(g->e, 1);
  const int x[16] = {lim(L), lim(nA), lim(uA), lim(fA), lim(nB), lim(uB), lim(fB),
                     (32 * f) / E, (32 * unk) / E, 0, head, -lim(lp),
                     -(int)en.lqq[first], -(int)en.lqq[last], lim(uA - uB), 1};
  long long score = 0;
  for (int h = 0; h < 64; ++h) {
    long long z = 0;
    for (int k = 0; k < 16; ++k) z += (long long)x[k] * en.chooser[k * 64 + h];
    if (z > 0) score += en.chooser[16 * 64 + h] ? -z : z;
  }
  return (int)score;
}

/* Alternating-sweep rule (ORC_MODE_ALTERNATE; no counterpart in the
 * reference).  Any src->dst path, and any order of its UNKNOWN hops, splits
 * an item into disjoint children plus the all-WORKING success term, so this
 * rule keeps the decomposition exact; it only changes the tree's shape.
 *
 * Path: fewest UNKNOWN hops, WORKING edges contracted.  Level 0 is the set
 * of nodes joined to src by WORKING edges; level l+1 is every node without a
 * level that has an allowed edge from level l, together with the nodes
 * joined to those by WORKING edges.  Every level is closed under WORKING
 * edges, so an allowed edge between consecutive levels is UNKNOWN.  The
 * walk back starts from K = dst's WORKING component; at level l the parent
 * is the lowest-index node u of level l-1 with an allowed edge to K, the hop
 * is u's edge to the lowest-index node of K it reaches, and the walk goes on
 * from u's WORKING component.
 *
 * Sweep: with f the item's FAILED edge count, the children are formed from
 * the src end (child i: the first i-1 hops WORKING, hop i FAILED) when
 * (f / period) is even and from the dst end otherwise; the masses multiply
 * along the sweep.  Each child has one more FAILED edge than its parent, so
 * the direction flips every `period` tree levels. */
int expand_alt_impl(const Engine& en, const u64* known, const u64* up, double mass,
                    u64* ck, u64* cu, double* cm, double* succ, int* hops) {
  const orc_graph* g = en.g;
  const int n = g->n, W = en.W;
  if (en.src == en.dst) {
    *succ = mass;
    return 0;
  }
  std::vector<int> lev(n, -1);
  /* extend `set` (nodes just given level l) over WORKING edges */
  auto close = [&](std::vector<int>& set, int l) {
    for (size_t i = 0; i < set.size(); ++i) {
      const int u = set[i];
      for (int j = g->row_ptr[u]; j < g->row_ptr[u + 1]; ++j) {
        const int v = g->col_idx[j];
        if (lev[v] < 0 && en.working(known, up, g->edge_id[j])) {
          lev[v] = l;
          set.push_back(v);
        }
      }
    }
  };
  std::vector<int> cur{en.src};
  lev[en.src] = 0;
  close(cur, 0);
  int L = 0;
  while (lev[en.dst] < 0) {
    std::vector<int> nxt;
    for (int u : cur)
      for (int j = g->row_ptr[u]; j < g->row_ptr[u + 1]; ++j) {
        const int v = g->col_idx[j];
        if (lev[v] < 0 && en.allowed(known, up, g->edge_id[j])) {
          lev[v] = L + 1;
          nxt.push_back(v);
        }
      }
    if (nxt.empty()) return -1;
    ++L;
    close(nxt, L);
    cur.swap(nxt);
  }
  /* WORKING component of x as flags */
  auto comp = [&](int x) {
    std::vector<char> in(n, 0);
    std::vector<int> st{x};
    in[x] = 1;
    while (!st.empty()) {
      const int u = st.back();
      st.pop_back();
      for (int j = g->row_ptr[u]; j < g->row_ptr[u + 1]; ++j) {
        const int v = g->col_idx[j];
        if (!in[v] && en.working(known, up, g->edge_id[j])) {
          in[v] = 1;
          st.push_back(v);
        }
      }
    }
    return in;
  };
  std::vector<char> K = comp(en.dst);
  for (int l = L; l >= 1; --l) {
    int bu = -1, be = -1;
    for (int u = 0; u < n && bu < 0; ++u) {
      if (lev[u] != l - 1) continue;
      int bv = n;
      for (int j = g->row_ptr[u]; j < g->row_ptr[u + 1]; ++j) {
        const int v = g->col_idx[j], e = g->edge_id[j];
        if (K[v] && en.allowed(known, up, e) && (v < bv || (v == bv && e < be))) {
          bv = v;
          be = e;
        }
      }
      if (bv < n) bu = u;
    }
    if (bu < 0) return -1; /* not reached: every level joins the previous one */
    hops[l - 1] = be;
    K = comp(bu);
  }
  int f = 0;
  for (int i = 0; i < W; ++i) f += __builtin_popcountll(known[i] & ~up[i]);
  const int per = en.period > 0 ? en.period : 2;
  bool rev = ((f / per) & 1) != 0;
  if (en.mode == ORC_MODE_EXPLORE && L >= 2) {
    rev = rev != explore_flip(known, up, W);
  } else if (en.mode == ORC_MODE_LEARNED && L >= 2) {
    /* src's component is the level-0 node set; dst's is closed anew */
    std::vector<char> S(n, 0);
    for (int v = 0; v < n; ++v) S[v] = lev[v] == 0;
    const std::vector<char> T = comp(en.dst);
    int unk = 0;
    for (int e = 0; e < g->e; ++e)
      if (!((known[e >> 6] >> (e & 63)) & 1)) ++unk;
    const int score_fwd = chooser_score(en, S, T, hops, L, f, unk, mass, known, up, false);
    const int score_rev = chooser_score(en, S, T, hops, L, f, unk, mass, known, up, true);
    /* leave the alternating direction only for an advantage above the margin */
    const int s_alt = rev ? score_rev : score_fwd, s_other = rev ? score_fwd : score_rev;
    if (s_alt - s_other > en.chooser[16 * 64 + 64]) rev = !rev;
  }
  if (rev) std::reverse(hops, hops + L);
  u64 pre[kMaxW] = {0};
  double run = mass;
  for (int h = 0; h < L; ++h) {
    const int e = hops[h];
    const u64 bit = 1ull << (e & 63);
    const int wi = e >> 6;
    u64* kk = ck + (size_t)h * W;
    u64* uu = cu + (size_t)h * W;
    for (int i = 0; i < W; ++i) {
      kk[i] = known[i] | pre[i];
      uu[i] = up[i] | pre[i];
    }
    kk[wi] |= bit;
    cm[h] = run * g->q[e];
    run = run * g->p[e];
    pre[wi] |= bit;
  }
  *succ = run;
  return L;
}

/* Flat item store: planes (known, up) of W words + FP64 mass. */
struct Items {
  int W = 1;
  std::vector<u64> planes;
  std::vector<double> mass;
  size_t size() const { return mass.size(); }
  void clear() {
    planes.clear();
    mass.clear();
  }
  void push(const u64* k, const u64* u, double m) {
    planes.insert(planes.end(), k, k + W);
    planes.insert(planes.end(), u, u + W);
    mass.push_back(m);
  }
  const u64* known(size_t i) const { return planes.data() + i * 2 * W; }
  const u64* up(size_t i) const { return planes.data() + i * 2 * W + W; }
  void pop_back() {
    planes.resize(planes.size() - 2 * W);
    mass.pop_back();
  }
};

struct Acc {
  KSum R, ESP, D;
  unsigned long long paths = 0, nodes = 0, dead = 0, truncated = 0,
                     overflowed = 0, folded = 0;
  void merge(const Acc& o) {
    R.merge(o.R);
    ESP.merge(o.ESP);
    D.merge(o.D);
    paths += o.paths;
    nodes += o.nodes;
    dead += o.dead;
    truncated += o.truncated;
    overflowed += o.overflowed;
    folded += o.folded;
  }
};

/* Processes one dequeued item: counting, truncation and dead-end handling of
 * main.cu:533-582, then expansion.  Children are appended through `emit`.
 * Returns false when the item was truncated or dead. */
struct Worker {
  const Engine& en;
  Acc acc;
  std::vector<int> hops;
  std::vector<NodeSet> levels;
  std::vector<u64> ck, cu;
  std::vector<double> cm;
  explicit Worker(const Engine& e) : en(e) {
    hops.resize(e.g->n + 1);
    levels.resize(e.g->n + 1);
    ck.resize((size_t)(e.g->n + 2) * e.W);
    cu.resize((size_t)(e.g->n + 2) * e.W);
    cm.resize(e.g->n + 2);
  }
  /* returns number of children in ck/cu/cm (or 0) */
  int process(const u64* known, const u64* up, double mass) {
    acc.nodes++;
    if (mass < en.eps) { /* main.cu:537-543 */
      acc.ESP.add(mass);
      acc.truncated++;
      return 0;
    }
    double succ = 0.0;
    int k;
    if (en.mode == ORC_MODE_PIVOT)
      k = en.expand_pivot(known, up, mass, ck.data(), cu.data(), cm.data(), &succ,
                          hops.data(), levels);
    else if (en.mode == ORC_MODE_FORKBEST)
      k = expand_forkbest_impl(en, known, up, mass, ck.data(), cu.data(), cm.data(), &succ,
                               hops.data(), levels);
    else if (en.mode == ORC_MODE_MINUNK)
      k = expand_minunk_impl(en, known, up, mass, ck.data(), cu.data(), cm.data(), &succ,
                             hops.data());
    else if (en.mode == ORC_MODE_ALTERNATE || en.mode == ORC_MODE_LEARNED ||
             en.mode == ORC_MODE_EXPLORE)
      k = expand_alt_impl(en, known, up, mass, ck.data(), cu.data(), cm.data(), &succ,
                          hops.data());
    else
      k = en.expand_path(known, up, mass, ck.data(), cu.data(), cm.data(), &succ,
                         hops.data(), nullptr, levels);
    if (k < 0) { /* main.cu:579-582 */
      acc.D.add(mass);
      acc.dead++;
      return 0;
    }
    if (en.mode == ORC_MODE_PATH || en.mode == ORC_MODE_MINUNK ||
        en.mode == ORC_MODE_ALTERNATE || en.mode == ORC_MODE_LEARNED ||
        en.mode == ORC_MODE_EXPLORE || k == 0) {
      acc.R.add(succ);
      acc.paths++;
    }
    return k;
  }
};

void initial_item(const orc_graph* g, std::vector<u64>& k, std::vector<u64>& u) {
  /* main.cu:717-748: p <= 0 FAILED, p >= 1 WORKING, else UNKNOWN */
  int W = g->w;
  k.assign(W, 0);
  u.assign(W, 0);
  for (int e = 0; e < g->e; ++e) {
    u64 bit = 1ull << (e & 63);
    if (g->p[e] <= 0.0) {
      k[e >> 6] |= bit;
    } else if (g->p[e] >= 1.0) {
      k[e >> 6] |= bit;
      u[e >> 6] |= bit;
    }
  }
}

struct Deadline {
  bool on;
  Clock::time_point t;
  explicit Deadline(double s) : on(s > 0) {
    t = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                           std::chrono::duration<double>(s));
  }
  bool hit() const { return on && Clock::now() >= t; }
};

/* The items a solve starts from: the root, or the seeds of the static
 * prefix partition (original/Rel_16M.cu:97-101, thread-id bits fixing the
 * first edges) in seed order, masses multiplied in edge-id order. */
template <class Push>
void for_each_start(const Engine& en, const orc_config& cfg, Push push) {
  std::vector<u64> k0, u0;
  initial_item(en.g, k0, u0);
  if (cfg.prefix_edges <= 0) {
    push(k0.data(), u0.data(), 1.0);
    return;
  }
  std::vector<int> pick;
  for (int e = 0; e < en.g->e && (int)pick.size() < cfg.prefix_edges; ++e)
    if (!((k0[e >> 6] >> (e & 63)) & 1)) pick.push_back(e);
  const unsigned long long n_seeds = 1ull << pick.size();
  const unsigned long long step = cfg.shards > 1 ? (unsigned long long)cfg.shards : 1;
  std::vector<u64> k(en.W), u(en.W);
  for (unsigned long long i = (unsigned long long)cfg.shard; i < n_seeds; i += step) {
    k = k0;
    u = u0;
    double m = 1.0;
    for (size_t j = 0; j < pick.size(); ++j) {
      const int e = pick[j];
      k[e >> 6] |= 1ull << (e & 63);
      const bool working = (i >> j) & 1;
      if (working) u[e >> 6] |= 1ull << (e & 63);
      m *= working ? en.g->p[e] : en.g->q[e];
    }
    push(k.data(), u.data(), m);
  }
}

/* Level order (main.cu:963-1008): the whole current level is processed
 * before the next; children beyond `capacity` go to ESP and count as paths
 * (main.cu:245-252). iterations = number of levels; max_width = widest
 * level (main.cu:972-973). */
void run_level(const Engine& en, const orc_config& cfg, orc_result* r, Acc& acc) {
  Worker wk(en);
  Items cur, nxt;
  cur.W = nxt.W = en.W;
  for_each_start(en, cfg, [&](const u64* k, const u64* u, double m) { cur.push(k, u, m); });
  Deadline dl(cfg.time_budget_s);
  long long iters = 0, maxw = 0;
  size_t cap = cfg.capacity > 0 ? (size_t)cfg.capacity : (size_t)-1;
  bool stop = false;
  while (cur.size() && !stop) {
    ++iters;
    maxw = std::max(maxw, (long long)cur.size());
    nxt.clear();
    size_t i = 0;
    for (; i < cur.size(); ++i) {
      if ((i & 1023) == 0 && dl.hit()) {
        stop = true;
        break;
      }
      int k = wk.process(cur.known(i), cur.up(i), cur.mass[i]);
      for (int c = 0; c < k; ++c) {
        if (nxt.size() < cap) {
          nxt.push(&wk.ck[(size_t)c * en.W], &wk.cu[(size_t)c * en.W], wk.cm[c]);
        } else {
          wk.acc.ESP.add(wk.cm[c]);
          wk.acc.paths++;
          wk.acc.overflowed++;
        }
      }
    }
    if (stop) {
      for (; i < cur.size(); ++i) {
        wk.acc.ESP.add(cur.mass[i]);
        wk.acc.folded++;
      }
      for (size_t j = 0; j < nxt.size(); ++j) {
        wk.acc.ESP.add(nxt.mass[j]);
        wk.acc.folded++;
      }
      r->deadline_hit = 1;
      break;
    }
    std::swap(cur, nxt);
  }
  r->iterations = iters;
  r->max_width = maxw;
  acc.merge(wk.acc);
}

/* Max-mass binary heap order (seq/seq.cpp:49-197, 481-592): pop the largest
 * mass; a child that finds the heap full goes to ESP (seq.cpp:552-561) and is
 * counted as a path like the GPU's overflow (main.cu:251). */
void run_heap(const Engine& en, const orc_config& cfg, orc_result* r, Acc& acc) {
  Worker wk(en);
  const int W = en.W;
  std::vector<u64> planes;
  std::vector<double> key;
  auto sw = [&](size_t a, size_t b) {
    std::swap(key[a], key[b]);
    for (int i = 0; i < 2 * W; ++i) std::swap(planes[a * 2 * W + i], planes[b * 2 * W + i]);
  };
  auto push = [&](const u64* k, const u64* u, double m) {
    planes.insert(planes.end(), k, k + W);
    planes.insert(planes.end(), u, u + W);
    key.push_back(m);
    size_t i = key.size() - 1;
    while (i > 0) {
      size_t par = (i - 1) / 2;
      if (key[par] >= key[i]) break;
      sw(par, i);
      i = par;
    }
  };
  std::vector<u64> tk(W), tu(W);
  for_each_start(en, cfg, [&](const u64* k, const u64* u, double m) { push(k, u, m); });
  Deadline dl(cfg.time_budget_s);
  size_t cap = cfg.capacity > 0 ? (size_t)cfg.capacity : (size_t)-1;
  long long maxw = 1;
  unsigned long long pops = 0;
  while (!key.empty()) {
    if ((pops & 1023) == 0 && dl.hit()) {
      for (double m : key) {
        wk.acc.ESP.add(m);
        wk.acc.folded++;
      }
      r->deadline_hit = 1;
      break;
    }
    ++pops;
    double m = key[0];
    std::copy(planes.begin(), planes.begin() + W, tk.begin());
    std::copy(planes.begin() + W, planes.begin() + 2 * W, tu.begin());
    size_t last = key.size() - 1;
    if (last > 0) sw(0, last);
    key.pop_back();
    planes.resize(planes.size() - 2 * W);
    size_t i = 0, n = key.size();
    while (2 * i + 1 < n) {
      size_t l = 2 * i + 1, rr = l + 1, big = i;
      if (key[l] > key[big]) big = l;
      if (rr < n && key[rr] > key[big]) big = rr;
      if (big == i) break;
      sw(i, big);
      i = big;
    }
    int k = wk.process(tk.data(), tu.data(), m);
    for (int c = 0; c < k; ++c) {
      if (key.size() < cap) {
        push(&wk.ck[(size_t)c * W], &wk.cu[(size_t)c * W], wk.cm[c]);
      } else {
        wk.acc.ESP.add(wk.cm[c]);
        wk.acc.paths++;
        wk.acc.overflowed++;
      }
    }
    maxw = std::max(maxw, (long long)key.size());
  }
  r->iterations = 1;
  r->max_width = maxw;
  acc.merge(wk.acc);
}

/* Depth-first order, optionally multithreaded.  The tree is first expanded
 * level by level until the frontier holds >= 64*T items; T threads then pull
 * frontier items from a shared counter and finish each subtree depth-first.
 * Without capacity limits or a deadline the tree does not depend on the
 * order, so counts equal the level order's exactly. */
void run_dfs(const Engine& en, const orc_config& cfg, orc_result* r, Acc& acc) {
  int T = cfg.threads > 1 ? cfg.threads : 1;
  Worker seed(en);
  Items cur, nxt;
  cur.W = nxt.W = en.W;
  for_each_start(en, cfg, [&](const u64* k, const u64* u, double m) { cur.push(k, u, m); });
  size_t want = (size_t)64 * T;
  long long iters = 0;
  Deadline dl(cfg.time_budget_s);
  while (cur.size() && cur.size() < want && !dl.hit()) {
    ++iters;
    nxt.clear();
    for (size_t i = 0; i < cur.size(); ++i) {
      int k = seed.process(cur.known(i), cur.up(i), cur.mass[i]);
      for (int c = 0; c < k; ++c)
        nxt.push(&seed.ck[(size_t)c * en.W], &seed.cu[(size_t)c * en.W], seed.cm[c]);
    }
    std::swap(cur, nxt);
  }
  std::atomic<size_t> next_idx{0};
  std::atomic<bool> expired{false};
  std::vector<Acc> accs(T);
  auto body = [&](int t) {
    Worker wk(en);
    Items st;
    st.W = en.W;
    unsigned long long since = 0;
    for (;;) {
      size_t i = next_idx.fetch_add(1);
      if (i >= cur.size()) break;
      if (expired.load(std::memory_order_relaxed)) {
        wk.acc.ESP.add(cur.mass[i]);
        wk.acc.folded++;
        continue;
      }
      st.clear();
      st.push(cur.known(i), cur.up(i), cur.mass[i]);
      while (st.size()) {
        if ((++since & 1023) == 0 && dl.hit()) expired.store(true);
        if (expired.load(std::memory_order_relaxed)) {
          for (size_t j = 0; j < st.size(); ++j) {
            wk.acc.ESP.add(st.mass[j]);
            wk.acc.folded++;
          }
          st.clear();
          break;
        }
        size_t top = st.size() - 1;
        u64 tk[kMaxW], tu[kMaxW];
        std::memcpy(tk, st.known(top), sizeof(u64) * en.W);
        std::memcpy(tu, st.up(top), sizeof(u64) * en.W);
        double m = st.mass[top];
        st.pop_back();
        int k = wk.process(tk, tu, m);
        for (int c = k - 1; c >= 0; --c)
          st.push(&wk.ck[(size_t)c * en.W], &wk.cu[(size_t)c * en.W], wk.cm[c]);
      }
    }
    accs[t] = wk.acc;
  };
  if (T == 1) {
    body(0);
  } else {
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) th.emplace_back(body, t);
    for (auto& x : th) x.join();
  }
  acc.merge(seed.acc);
  for (auto& a : accs) acc.merge(a);
  if (expired.load() || dl.hit()) r->deadline_hit = expired.load() ? 1 : r->deadline_hit;
  r->iterations = iters;
  r->max_width = (long long)cur.size();
}

}  // namespace

extern "C" {

int orc_graph_load(const char* path, orc_graph** out) {
  if (!path || !out) return ORC_ERR_ARG;
  FILE* f = std::fopen(path, "rb");
  if (!f) return ORC_ERR_IO;
  std::string text;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
  std::fclose(f);
  orc_graph* g = new orc_graph();
  int rc = load_text(text.data(), text.size(), g);
  if (rc != ORC_OK) {
    delete g;
    return rc;
  }
  *out = g;
  return ORC_OK;
}

int orc_graph_from_csr(int n, int e, const int* row_ptr, const int* col_idx,
                       const int* edge_id, const double* p, orc_graph** out) {
  if (n <= 0 || e < 0 || !row_ptr || !out) return ORC_ERR_ARG;
  orc_graph* g = new orc_graph();
  g->n = n;
  g->e = e;
  g->header_e = e;
  g->row_ptr.assign(row_ptr, row_ptr + n + 1);
  g->nnz = row_ptr[n];
  g->col_idx.assign(col_idx, col_idx + g->nnz);
  g->edge_id.assign(edge_id, edge_id + g->nnz);
  g->p.assign(p, p + e);
  for (int j = 0; j < g->nnz; ++j)
    if (g->col_idx[j] < 0 || g->col_idx[j] >= n || g->edge_id[j] < 0 ||
        g->edge_id[j] >= e) {
      delete g;
      return ORC_ERR_RANGE;
    }
  finish_graph(g);
  *out = g;
  return ORC_OK;
}

void orc_graph_free(orc_graph* g) { delete g; }

void orc_graph_dims(const orc_graph* g, int* n, int* e, int* nnz, int* header_e,
                    int* type) {
  if (n) *n = g->n;
  if (e) *e = g->e;
  if (nnz) *nnz = g->nnz;
  if (header_e) *header_e = g->header_e;
  if (type) *type = g->type;
}

void orc_graph_arrays(const orc_graph* g, int* row_ptr, int* col_idx, int* edge_id,
                      double* p) {
  if (row_ptr) std::copy(g->row_ptr.begin(), g->row_ptr.end(), row_ptr);
  if (col_idx) std::copy(g->col_idx.begin(), g->col_idx.end(), col_idx);
  if (edge_id) std::copy(g->edge_id.begin(), g->edge_id.end(), edge_id);
  if (p) std::copy(g->p.begin(), g->p.end(), p);
}

/* ------------------------------------------------------------------ */
/* Graphs beyond the fixed node-set and plane sizes (more than 256 nodes */
/* or 1024 edges): the same path rule and level order with per-node      */
/* level arrays and planes of any width.  Path mode, level order only.   */
/* ------------------------------------------------------------------ */
namespace {

struct BigEngine {
  const orc_graph* g;
  int src, dst, W;
  std::vector<int> lev, cur, nxt;
  bool allowed(const u64* known, const u64* up, int e) const {
    const u64 b = 1ull << (e & 63);
    return !((known[e >> 6] & b) && !(up[e >> 6] & b));
  }
  /* BFS from src over allowed arcs, whole levels, until the level holding
   * dst (device_code.cuh:80-194); the walk back takes the lowest-index
   * node of the previous level with an allowed arc, lowest edge id on ties
   * (in-arc rows are sorted by (u, eid)).  Returns the path length (hops in
   * src->dst order) or -1 when dst is unreachable. */
  int find(const u64* known, const u64* up, std::vector<int>& hops) {
    const int n = g->n;
    if (src == dst) return 0;
    lev.assign(n, -1);
    cur.assign(1, src);
    lev[src] = 0;
    int L = 0;
    while (lev[dst] < 0) {
      nxt.clear();
      for (int u : cur)
        for (int j = g->row_ptr[u]; j < g->row_ptr[u + 1]; ++j) {
          const int v = g->col_idx[j];
          if (lev[v] < 0 && allowed(known, up, g->edge_id[j])) {
            lev[v] = L + 1;
            nxt.push_back(v);
          }
        }
      if (nxt.empty()) return -1;
      ++L;
      cur.swap(nxt);
    }
    hops.assign(L, -1);
    int v = dst;
    for (int l = L; l >= 1; --l) {
      int pu = -1, pe = -1;
      for (int k = g->in_ptr[v]; k < g->in_ptr[v + 1]; ++k)
        if (lev[g->in_src[k]] == l - 1 && allowed(known, up, g->in_eid[k])) {
          pu = g->in_src[k];
          pe = g->in_eid[k];
          break;
        }
      if (pu < 0) return -1; /* not reached: every levelled node has a parent */
      hops[l - 1] = pe;
      v = pu;
    }
    return L;
  }
};

void run_level_big(const orc_graph* g, int src, int dst, const orc_config& cfg, orc_result* r,
                   Acc& acc) {
  const int W = g->w;
  BigEngine en{g, src, dst, W, {}, {}, {}};
  std::vector<u64> cur, nxt; /* items: known[W], up[W] */
  std::vector<double> cm, nm;
  cur.resize(2 * (size_t)W);
  {
    std::vector<u64> k0, u0;
    initial_item(g, k0, u0);
    std::copy(k0.begin(), k0.end(), cur.begin());
    std::copy(u0.begin(), u0.end(), cur.begin() + W);
  }
  cm.assign(1, 1.0);
  const size_t cap = cfg.capacity > 0 ? (size_t)cfg.capacity : (size_t)-1;
  const double eps = cfg.trunc_eps;
  Deadline dl(cfg.time_budget_s);
  std::vector<int> hops;
  std::vector<u64> pre(W);
  long long iters = 0, maxw = 0;
  bool stop = false;
  while (!cm.empty() && !stop) {
    ++iters;
    maxw = std::max(maxw, (long long)cm.size());
    nxt.clear();
    nm.clear();
    size_t i = 0;
    for (; i < cm.size(); ++i) {
      if ((i & 255) == 0 && dl.hit()) {
        stop = true;
        break;
      }
      const u64* known = &cur[i * 2 * W];
      const u64* up = known + W;
      const double mass = cm[i];
      acc.nodes++;
      if (mass < eps) {
        acc.ESP.add(mass);
        acc.truncated++;
        continue;
      }
      const int L = en.find(known, up, hops);
      if (L < 0) {
        acc.D.add(mass);
        acc.dead++;
        continue;
      }
      std::fill(pre.begin(), pre.end(), 0ull);
      double run = mass;
      for (int h = 0; h < L; ++h) {
        const int e = hops[h];
        const u64 bit = 1ull << (e & 63);
        if (known[e >> 6] & bit) continue; /* a WORKING hop */
        const double c = run * g->q[e];
        if (nm.size() < cap) {
          const size_t o = nxt.size();
          nxt.resize(o + 2 * (size_t)W);
          for (int w = 0; w < W; ++w) {
            nxt[o + w] = known[w] | pre[w];
            nxt[o + W + w] = up[w] | pre[w];
          }
          nxt[o + (e >> 6)] |= bit;
          nm.push_back(c);
        } else { /* main.cu:245-252 */
          acc.ESP.add(c);
          acc.paths++;
          acc.overflowed++;
        }
        run = run * g->p[e];
        pre[e >> 6] |= bit;
      }
      acc.R.add(run);
      acc.paths++;
    }
    if (stop) { /* fold what is left of this level and the next */
      for (size_t j = i; j < cm.size(); ++j) {
        acc.ESP.add(cm[j]);
        acc.folded++;
      }
      for (double m : nm) {
        acc.ESP.add(m);
        acc.folded++;
      }
      r->deadline_hit = 1;
      break;
    }
    cur.swap(nxt);
    cm.swap(nm);
  }
  r->iterations = iters;
  r->max_width = maxw;
}

}  // namespace

static int check_common(const orc_graph* g, int src, int dst) {
  if (!g) return ORC_ERR_ARG;
  if (src < 0 || src >= g->n || dst < 0 || dst >= g->n) return ORC_ERR_RANGE;
  if (g->w > kMaxW || g->nw > kMaxNW) return ORC_ERR_SIZE;
  return ORC_OK;
}

/* Per-edge probability steps of the learned choice: round(-4 log2 p) and
 * round(-4 log2 q), clamped to 0..63 (63 for a zero probability). */
static void choice_steps(const orc_graph* g, std::vector<unsigned char>& lqp,
                         std::vector<unsigned char>& lqq) {
  auto step = [](double x) -> unsigned char {
    if (!(x > 0.0)) return 63;
    return (unsigned char)std::max(0.0, std::min(63.0, std::nearbyint(-4.0 * std::log2(x))));
  };
  lqp.resize(g->e);
  lqq.resize(g->e);
  for (int e = 0; e < g->e; ++e) {
    lqp[e] = step(g->p[e]);
    lqq[e] = step(g->q[e]);
  }
}

int orc_solve(const orc_graph* g, int src, int dst, const orc_config* cfg,
              orc_result* out) {
  int rc = check_common(g, src, dst);
  if (rc == ORC_ERR_SIZE && cfg && out && cfg->mode == ORC_MODE_PATH &&
      cfg->order == ORC_ORDER_LEVEL) { /* beyond the fixed sizes */
    std::memset(out, 0, sizeof(*out));
    Acc acc;
    auto t0 = Clock::now();
    run_level_big(g, src, dst, *cfg, out, acc);
    auto t1 = Clock::now();
    out->elapsed_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    out->R = acc.R.value();
    out->ESP = acc.ESP.value();
    out->D = acc.D.value();
    out->paths = acc.paths;
    out->nodes = acc.nodes;
    out->dead = acc.dead;
    out->truncated = acc.truncated;
    out->overflowed = acc.overflowed;
    out->folded = acc.folded;
    out->n = g->n;
    out->e = g->e;
    return ORC_OK;
  }
  if (rc) return rc;
  if (!cfg || !out) return ORC_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  Engine en{g, src, dst, g->w, g->nw, cfg->mode, cfg->trunc_eps};
  en.period = cfg->sweep_period;
  if (cfg->mode == ORC_MODE_LEARNED && !cfg->chooser) return ORC_ERR_ARG;
  en.chooser = cfg->chooser;
  std::vector<unsigned char> lqp, lqq;
  choice_steps(g, lqp, lqq);
  en.lqp = lqp.data();
  en.lqq = lqq.data();
  Acc acc;
  auto t0 = Clock::now();
  switch (cfg->order) {
    case ORC_ORDER_LEVEL: run_level(en, *cfg, out, acc); break;
    case ORC_ORDER_HEAP: run_heap(en, *cfg, out, acc); break;
    case ORC_ORDER_DFS: run_dfs(en, *cfg, out, acc); break;
    default: return ORC_ERR_ARG;
  }
  auto t1 = Clock::now();
  out->elapsed_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  out->R = acc.R.value();
  out->ESP = acc.ESP.value();
  out->D = acc.D.value();
  out->paths = acc.paths;
  out->nodes = acc.nodes;
  out->dead = acc.dead;
  out->truncated = acc.truncated;
  out->overflowed = acc.overflowed;
  out->folded = acc.folded;
  out->n = g->n;
  out->e = g->e;
  return ORC_OK;
}

int orc_solve_terms(const orc_graph* g, const int* terms, int k, const orc_config* cfg,
                    orc_result* out) {
  if (!g || !terms || k < 2 || !cfg || !out) return ORC_ERR_ARG;
  for (int i = 0; i < k; ++i)
    if (terms[i] < 0 || terms[i] >= g->n) return ORC_ERR_RANGE;
  int rc = check_common(g, terms[0], terms[1]);
  if (rc) return rc;
  std::memset(out, 0, sizeof(*out));
  if (cfg->mode == ORC_MODE_ALTERNATE) return ORC_ERR_ARG; /* two terminals only */
  Engine en{g, terms[0], terms[1], g->w, g->nw, cfg->mode, cfg->trunc_eps, terms, k};
  Acc acc;
  auto t0 = Clock::now();
  switch (cfg->order) {
    case ORC_ORDER_LEVEL: run_level(en, *cfg, out, acc); break;
    case ORC_ORDER_HEAP: run_heap(en, *cfg, out, acc); break;
    case ORC_ORDER_DFS: run_dfs(en, *cfg, out, acc); break;
    default: return ORC_ERR_ARG;
  }
  auto t1 = Clock::now();
  out->elapsed_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  out->R = acc.R.value();
  out->ESP = acc.ESP.value();
  out->D = acc.D.value();
  out->paths = acc.paths;
  out->nodes = acc.nodes;
  out->dead = acc.dead;
  out->truncated = acc.truncated;
  out->overflowed = acc.overflowed;
  out->folded = acc.folded;
  out->n = g->n;
  out->e = g->e;
  return ORC_OK;
}

int orc_expand(const orc_graph* g, int src, int dst, const uint64_t* known,
               const uint64_t* up, double mass, int max_children,
               uint64_t* child_known, uint64_t* child_up, double* child_mass,
               double* succ, int* nhops, int* path_eids) {
  if (check_common(g, src, dst)) return -2;
  Engine en{g, src, dst, g->w, g->nw, ORC_MODE_PATH, 0.0};
  std::vector<NodeSet> levels(g->n + 1);
  std::vector<int> hops(g->n + 1);
  std::vector<u64> ck((size_t)(g->n + 1) * g->w), cu((size_t)(g->n + 1) * g->w);
  std::vector<double> cm(g->n + 1);
  double s = 0.0;
  int nh = 0;
  int k = en.expand_path(known, up, mass, ck.data(), cu.data(), cm.data(), &s,
                         hops.data(), &nh, levels);
  if (nhops) *nhops = nh;
  if (succ) *succ = (k < 0) ? 0.0 : s;
  if (k < 0) return -1;
  if (path_eids)
    for (int i = 0; i < nh; ++i) path_eids[i] = hops[i];
  int m = std::min(k, max_children);
  if (child_known) std::copy(ck.begin(), ck.begin() + (size_t)m * g->w, child_known);
  if (child_up) std::copy(cu.begin(), cu.begin() + (size_t)m * g->w, child_up);
  if (child_mass) std::copy(cm.begin(), cm.begin() + m, child_mass);
  return k;
}

int orc_expand_rule(const orc_graph* g, int src, int dst, int mode, int period,
                    const uint64_t* known, const uint64_t* up, double mass, int max_children,
                    uint64_t* child_known, uint64_t* child_up, double* child_mass, double* succ,
                    int* path_eids) {
  if (check_common(g, src, dst)) return -2;
  if (mode != ORC_MODE_PATH && mode != ORC_MODE_ALTERNATE) return -2;
  Engine en{g, src, dst, g->w, g->nw, mode, 0.0};
  en.period = period;
  std::vector<NodeSet> levels(g->n + 1);
  std::vector<int> hops(g->n + 1);
  std::vector<u64> ck((size_t)(g->n + 1) * g->w), cu((size_t)(g->n + 1) * g->w);
  std::vector<double> cm(g->n + 1);
  double s = 0.0;
  int k = mode == ORC_MODE_ALTERNATE
              ? expand_alt_impl(en, known, up, mass, ck.data(), cu.data(), cm.data(), &s,
                                hops.data())
              : en.expand_path(known, up, mass, ck.data(), cu.data(), cm.data(), &s,
                               hops.data(), nullptr, levels);
  if (succ) *succ = (k < 0) ? 0.0 : s;
  if (k < 0) return -1;
  if (path_eids && mode == ORC_MODE_ALTERNATE)
    for (int i = 0; i < k; ++i) path_eids[i] = hops[i];
  const int m = std::min(k, max_children);
  if (child_known) std::copy(ck.begin(), ck.begin() + (size_t)m * g->w, child_known);
  if (child_up) std::copy(cu.begin(), cu.begin() + (size_t)m * g->w, child_up);
  if (child_mass) std::copy(cm.begin(), cm.begin() + m, child_mass);
  return k;
}

int orc_expand_choice(const orc_graph* g, int src, int dst, int mode, int period,
                      const int* chooser, double eps, const uint64_t* known,
                      const uint64_t* up, double mass, int max_children, uint64_t* child_known,
                      uint64_t* child_up, double* child_mass, double* succ, int* path_eids) {
  if (check_common(g, src, dst)) return -2;
  if (mode != ORC_MODE_ALTERNATE && mode != ORC_MODE_LEARNED && mode != ORC_MODE_EXPLORE)
    return -2;
  if (mode == ORC_MODE_LEARNED && !chooser) return -2;
  Engine en{g, src, dst, g->w, g->nw, mode, eps};
  en.period = period;
  en.chooser = chooser;
  std::vector<unsigned char> lqp, lqq;
  choice_steps(g, lqp, lqq);
  en.lqp = lqp.data();
  en.lqq = lqq.data();
  std::vector<int> hops(g->n + 1);
  std::vector<u64> ck((size_t)(g->n + 1) * g->w), cu((size_t)(g->n + 1) * g->w);
  std::vector<double> cm(g->n + 1);
  double s = 0.0;
  const int k =
      expand_alt_impl(en, known, up, mass, ck.data(), cu.data(), cm.data(), &s, hops.data());
  if (succ) *succ = (k < 0) ? 0.0 : s;
  if (k < 0) return -1;
  if (path_eids)
    for (int i = 0; i < k; ++i) path_eids[i] = hops[i];
  const int m = std::min(k, max_children);
  if (child_known) std::copy(ck.begin(), ck.begin() + (size_t)m * g->w, child_known);
  if (child_up) std::copy(cu.begin(), cu.begin() + (size_t)m * g->w, child_up);
  if (child_mass) std::copy(cm.begin(), cm.begin() + m, child_mass);
  return k;
}

int orc_subtree_stats(const orc_graph* g, int src, int dst, int count,
                      const uint64_t* known, const uint64_t* up, const double* mass,
                      double trunc_eps, unsigned long long* nodes_out,
                      double* r_out) {
  int rc = check_common(g, src, dst);
  if (rc) return rc;
  Engine en{g, src, dst, g->w, g->nw, ORC_MODE_PATH, trunc_eps};
  Worker wk(en);
  Items st;
  st.W = en.W;
  for (int i = 0; i < count; ++i) {
    wk.acc = Acc();
    st.clear();
    st.push(known + (size_t)i * en.W, up + (size_t)i * en.W, mass[i]);
    while (st.size()) {
      size_t top = st.size() - 1;
      u64 tk[kMaxW], tu[kMaxW];
      std::memcpy(tk, st.known(top), sizeof(u64) * en.W);
      std::memcpy(tu, st.up(top), sizeof(u64) * en.W);
      double m = st.mass[top];
      st.pop_back();
      int k = wk.process(tk, tu, m);
      for (int c = k - 1; c >= 0; --c)
        st.push(&wk.ck[(size_t)c * en.W], &wk.cu[(size_t)c * en.W], wk.cm[c]);
    }
    if (nodes_out) nodes_out[i] = wk.acc.nodes;
    if (r_out) r_out[i] = wk.acc.R.value();
  }
  return ORC_OK;
}

}  // extern "C"
