/*
 * nr_graph.cpp - graph file loader and CSR construction (host C++).
 *
 * Keeps the reference's file format and edge numbering so graphs drop in
 * unchanged (workingPad/graph.h:46-145):
 *   header "N E [T]"            graph.h:53 (T optional here; the README
 *                               test graphs omit it, workingPad/README.md:126)
 *   per node "deg (nbr p)*"     graph.h:61-72
 *   edge id = first appearance of the pair (min,max) scanning nodes in
 *   ascending order, entries in file order          graph.h:76-90
 *   header E halved for type U, mismatch only warns graph.h:93-96
 *   CSR rows in file order                          graph.h:104-130
 * Probabilities are parsed and kept in FP64; q = 1 - p is formed once here
 * (graph.h:132-140 keeps float logs instead).
 */
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>

#include "nr_internal.h"

namespace nr {

unsigned long long next_stamp() {
  static std::atomic<unsigned long long> counter{0};
  return ++counter;
}

namespace {

inline bool is_space(char c) {
  return c == ' ' || c == '\t' || c == '\r' || c == '\n' || c == '\v' || c == '\f';
}

struct Scanner {
  const char* s;
  const char* end;
  void skip() {
    while (s < end && is_space(*s)) ++s;
  }
  bool integer(long long* v) {
    skip();
    if (s >= end) return false;
    char buf[32];
    size_t n = 0;
    while (s + n < end && !is_space(s[n]) && n < sizeof(buf) - 1) {
      buf[n] = s[n];
      ++n;
    }
    buf[n] = 0;
    char* stop = nullptr;
    errno = 0;
    long long x = std::strtoll(buf, &stop, 10);
    if (stop == buf || *stop != 0 || errno) return false;
    s += n;
    *v = x;
    return true;
  }
  bool real(double* v) {
    skip();
    if (s >= end) return false;
    char buf[64];
    size_t n = 0;
    while (s + n < end && !is_space(s[n]) && n < sizeof(buf) - 1) {
      buf[n] = s[n];
      ++n;
    }
    buf[n] = 0;
    char* stop = nullptr;
    double x = std::strtod(buf, &stop);
    if (stop == buf || *stop != 0) return false;
    s += n;
    *v = x;
    return true;
  }
  /* a header type token: one alphabetic word, e.g. "U" or "D" */
  int type_token() {
    skip();
    if (s < end && ((*s >= 'A' && *s <= 'Z') || (*s >= 'a' && *s <= 'z'))) {
      int t = *s;
      while (s < end && !is_space(*s)) ++s;
      return t;
    }
    return 0;
  }
};

}  // namespace

int load_graph_text(const char* text, size_t len, nr_graph* g) {
  Scanner sc{text, text + len};
  long long n = 0, ef = 0;
  if (!sc.integer(&n) || !sc.integer(&ef) || n <= 0 || n > (1 << 24) || ef < 0)
    return NR_ERR_PARSE;
  int type = sc.type_token();
  struct Entry {
    int nbr;
    double p;
  };
  std::vector<std::vector<Entry>> adj((size_t)n);
  for (long long u = 0; u < n; ++u) {
    long long deg = 0;
    /* each (nbr, p) pair takes at least two tokens of one byte and a
     * separator each: a degree the rest of the text cannot hold is
     * malformed (and must not size an allocation) */
    if (!sc.integer(&deg) || deg < 0 || deg > (long long)(sc.end - sc.s) / 2)
      return NR_ERR_PARSE;
    adj[u].reserve((size_t)deg);
    for (long long j = 0; j < deg; ++j) {
      long long v = 0;
      double p = 0;
      if (!sc.integer(&v) || !sc.real(&p)) return NR_ERR_PARSE;
      if (v < 0 || v >= n) return NR_ERR_RANGE;
      adj[u].push_back({(int)v, p});
    }
  }
  g->n = (int)n;
  g->type = type;
  g->header_e = type == 'U' ? (int)(ef / 2) : (int)ef;
  std::unordered_map<uint64_t, int> id_of;
  id_of.reserve(1024);
  g->p.clear();
  auto key = [n](int a, int b) {
    return (uint64_t)std::min(a, b) * (uint64_t)n + (uint64_t)std::max(a, b);
  };
  for (int u = 0; u < g->n; ++u)
    for (const Entry& en : adj[u]) {
      auto it = id_of.find(key(u, en.nbr));
      if (it == id_of.end()) {
        id_of.emplace(key(u, en.nbr), (int)g->p.size());
        g->p.push_back(en.p);
      }
    }
  g->e = (int)g->p.size();
  g->row_ptr.assign((size_t)g->n + 1, 0);
  for (int u = 0; u < g->n; ++u) g->row_ptr[u + 1] = g->row_ptr[u] + (int)adj[u].size();
  g->nnz = g->row_ptr[g->n];
  g->col_idx.resize(g->nnz);
  g->edge_id.resize(g->nnz);
  for (int u = 0, k = 0; u < g->n; ++u)
    for (const Entry& en : adj[u]) {
      g->col_idx[k] = en.nbr;
      g->edge_id[k] = id_of[key(u, en.nbr)];
      ++k;
    }
  return finish_graph(g);
}

/* Derived per-edge data: endpoints, CSR direction coverage, q = 1 - p and
 * the counts of pre-resolved edges (main.cu:729-745). */
int finish_graph(nr_graph* g) {
  g->ea.assign(g->e, -1);
  g->eb.assign(g->e, -1);
  g->dirmask.assign(g->e, 0);
  for (int u = 0; u < g->n; ++u)
    for (int j = g->row_ptr[u]; j < g->row_ptr[u + 1]; ++j) {
      int v = g->col_idx[j], id = g->edge_id[j];
      if (v < 0 || v >= g->n || id < 0 || id >= g->e) return NR_ERR_RANGE;
      int a = std::min(u, v), b = std::max(u, v);
      if (g->ea[id] < 0) {
        g->ea[id] = a;
        g->eb[id] = b;
      } else if (g->ea[id] != a || g->eb[id] != b) {
        return NR_ERR_RANGE; /* one id must name one node pair */
      }
      g->dirmask[id] |= (u == a) ? 1 : 2;
      if (a == b) g->dirmask[id] = 3;
    }
  g->symmetric = true;
  for (int id = 0; id < g->e; ++id) {
    if (g->ea[id] < 0) return NR_ERR_RANGE; /* edge id never used */
    if (g->dirmask[id] != 3) g->symmetric = false;
  }
  {
    std::vector<uint64_t> pairs((size_t)g->e);
    for (int id = 0; id < g->e; ++id)
      pairs[id] = (uint64_t)(uint32_t)g->ea[id] << 32 | (uint32_t)g->eb[id];
    std::sort(pairs.begin(), pairs.end());
    g->simple = std::adjacent_find(pairs.begin(), pairs.end()) == pairs.end();
  }
  g->q.resize(g->e);
  g->precontracted = g->predeleted = 0;
  for (int id = 0; id < g->e; ++id) {
    g->q[id] = 1.0 - g->p[id];
    if (g->p[id] <= 0.0)
      g->predeleted++;
    else if (g->p[id] >= 1.0)
      g->precontracted++;
  }
  return NR_OK;
}

}  // namespace nr

extern "C" {

int nr_graph_load(const char* path, int src, int dst, nr_graph** out) {
  return nr::guarded([&]() -> int {
    if (!path || !out) return NR_ERR_ARG;
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return NR_ERR_IO;
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    std::fclose(f);
    nr_graph* g = new nr_graph();
    int rc = nr::load_graph_text(text.data(), text.size(), g);
    if (rc == NR_OK) rc = nr_graph_set_terminals(g, src, dst);
    if (rc != NR_OK) {
      delete g;
      return rc;
    }
    *out = g;
    return NR_OK;
  });
}

int nr_graph_create(int n, int e, const int* row_ptr, const int* col_idx,
                    const int* edge_id, const double* p, int src, int dst,
                    nr_graph** out) {
  return nr::guarded([&]() -> int {
    if (!out || n <= 0 || e < 0 || !row_ptr || (e > 0 && !p)) return NR_ERR_ARG;
    *out = nullptr;
    int nnz = row_ptr[n];
    if (nnz < 0 || row_ptr[0] != 0) return NR_ERR_ARG;
    for (int u = 0; u < n; ++u)
      if (row_ptr[u + 1] < row_ptr[u]) return NR_ERR_ARG;
    if (nnz > 0 && (!col_idx || !edge_id)) return NR_ERR_ARG;
    nr_graph* g = new nr_graph();
    g->n = n;
    g->e = e;
    g->nnz = nnz;
    g->header_e = e;
    g->row_ptr.assign(row_ptr, row_ptr + n + 1);
    g->col_idx.assign(col_idx, col_idx + nnz);
    g->edge_id.assign(edge_id, edge_id + nnz);
    g->p.assign(p, p + e);
    int rc = nr::finish_graph(g);
    if (rc == NR_OK) rc = nr_graph_set_terminals(g, src, dst);
    if (rc != NR_OK) {
      delete g;
      return rc;
    }
    *out = g;
    return NR_OK;
  });
}

int nr_graph_set_terminals(nr_graph* g, int src, int dst) {
  return nr::guarded([&]() -> int {
    if (!g) return NR_ERR_ARG;
    if (src < 0 || src >= g->n || dst < 0 || dst >= g->n) return NR_ERR_RANGE;
    g->src = src;
    g->dst = dst;
    g->terms.clear(); /* back to the two-terminal problem */
    g->stamp = nr::next_stamp();
    return NR_OK;
  });
}

int nr_graph_get_info(const nr_graph* g, nr_graph_info* info) {
  return nr::guarded([&]() -> int {
    if (!g || !info) return NR_ERR_ARG;
    info->n = g->n;
    info->e = g->e;
    info->nnz = g->nnz;
    info->header_e = g->header_e;
    info->type = g->type;
    info->src = g->src;
    info->dst = g->dst;
    info->mask_words = nr::mask_words(g);
    info->symmetric = g->symmetric ? 1 : 0;
    info->precontracted = g->precontracted;
    info->predeleted = g->predeleted;
    return NR_OK;
  });
}

int nr_graph_get_csr(const nr_graph* g, int* row_ptr, int* col_idx, int* edge_id,
                     double* p) {
  return nr::guarded([&]() -> int {
    if (!g) return NR_ERR_ARG;
    if (row_ptr) std::copy(g->row_ptr.begin(), g->row_ptr.end(), row_ptr);
    if (col_idx) std::copy(g->col_idx.begin(), g->col_idx.end(), col_idx);
    if (edge_id) std::copy(g->edge_id.begin(), g->edge_id.end(), edge_id);
    if (p) std::copy(g->p.begin(), g->p.end(), p);
    return NR_OK;
  });
}

void nr_graph_destroy(nr_graph* g) { delete g; }

int nr_graph_set_terminal_set(nr_graph* g, const int* terminals, int k) {
  return nr::guarded([&]() -> int {
    if (!g || k < 0 || k == 1 || (k > 0 && !terminals)) return NR_ERR_ARG;
    if (k > 256) return NR_ERR_UNSUPPORTED;
    for (int i = 0; i < k; ++i)
      if (terminals[i] < 0 || terminals[i] >= g->n) return NR_ERR_RANGE;
    g->terms.assign(terminals, terminals + k);
    if (k >= 2) {
      g->src = terminals[0];
      g->dst = terminals[1];
    }
    g->stamp = nr::next_stamp();
    return NR_OK;
  });
}

int nr_graph_get_terminal_set(const nr_graph* g, int* terminals, int max_k, int* k) {
  return nr::guarded([&]() -> int {
    if (!g || !k) return NR_ERR_ARG;
    *k = (int)g->terms.size();
    for (int i = 0; i < *k && i < max_k && terminals; ++i) terminals[i] = g->terms[i];
    return NR_OK;
  });
}

int nr_config_init(nr_config* cfg) {
  return nr::guarded([&]() -> int {
    if (!cfg) return NR_ERR_ARG;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->queue_capacity = 1 << 20; /* DEFAULT_QUEUE_CAPACITY, common.h:33 */
    cfg->truncation_eps = 1e-15;   /* DEFAULT_TRUNCATION_EPS, common.h:34 */
    cfg->order = NR_ORDER_LEVEL;
    return NR_OK;
  });
}

const char* nr_status_str(int s) {
  switch (s) {
    case NR_OK: return "ok";
    case NR_ERR_IO: return "cannot open graph file";
    case NR_ERR_PARSE: return "malformed graph file";
    case NR_ERR_RANGE: return "index out of range";
    case NR_ERR_ARG: return "invalid argument";
    case NR_ERR_UNSUPPORTED: return "graph beyond supported width classes";
    case NR_ERR_DEVICE: return "HIP runtime error";
    case NR_ERR_NOMEM: return "device allocation failed";
    case NR_ERR_SCRATCH: return "solve needs more than the fixed scratch area";
    default: return "unknown status";
  }
}

}  // extern "C"
