/*
 * oracle.h - C-ABI of the CPU oracle for two-terminal reliability by path
 * decomposition.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests, the smoke
 * entry point and the CPU-baseline leg of the benchmark call it.  The product
 * (the HIP engine and its CLI) never links or loads it.
 *
 * It restates, in plain sequential C++:
 *   - workingPad/graph.h:46-145        graph file parse, edge ids, CSR
 *   - workingPad/main.cu:717-748       initial edge states (p<=0 down, p>=1 up)
 *   - workingPad/device_code.cuh:80-194 optimistic BFS; parent rule made
 *                                       deterministic (lowest-index parent)
 *   - workingPad/main.cu:114-283       k children + success term per item
 *   - workingPad/main.cu:465-629       per-item counting, truncation, dead ends
 *   - workingPad/main.cu:635-711,963-1008 level (FIFO by depth) order,
 *                                       per-level capacity, iterations, RSS
 *   - seq/seq.cpp:49-197,481-592       max-mass heap order, ESP on a full queue
 *   - original/Seq_16.cpp:44,227       time-budget fold of the queue into ESP
 *   - workingPad/cpu_reference.cpp:85-145 edge-pivot factoring ("pivot" mode)
 *   - workingPad/main.cu:304-430,550-567 forkBest: confirmed BFS success, else
 *                                       two-way branch on the path edge of
 *                                       highest p ("forkbest" mode)
 * with FP64 masses and compensated (Neumaier) sums.
 */
#ifndef NR_ORACLE_H
#define NR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* orders */
#define ORC_ORDER_LEVEL 0 /* FIFO by tree depth (the reference GPU's order) */
#define ORC_ORDER_HEAP 1  /* max-mass binary heap (seq/seq.cpp) */
#define ORC_ORDER_DFS 2   /* depth-first; used by the multithreaded baseline */

/* modes */
#define ORC_MODE_PATH 0  /* path decomposition, k children per item */
#define ORC_MODE_PIVOT 1 /* edge-pivot factoring (workingPad/cpu_reference.cpp) */
#define ORC_MODE_FORKBEST 2 /* two-way branch on the best edge of the path
                               (workingPad/main.cu:304-430, 550-567) */
#define ORC_MODE_MINUNK 3   /* path decomposition on a path with the fewest
                               UNKNOWN edges (BFS over WORKING components) */
#define ORC_MODE_ALTERNATE 4 /* fewest-UNKNOWN path (lowest-index walk back over
                                WORKING components), children swept from the src
                                end or the dst end by the FAILED edge count */
#define ORC_MODE_LEARNED 5   /* the same path, swept as ORC_MODE_ALTERNATE unless
                                the integer chooser (orc_config.chooser) scores
                                the other direction lower by more than its
                                margin */
#define ORC_MODE_EXPLORE 6   /* the same path, swept as ORC_MODE_ALTERNATE but
                                in the other direction where the top two bits
                                of a splitmix64 hash of the planes are zero
                                (the exploring policy of the engine's choice
                                log) */

/* status codes */
#define ORC_OK 0
#define ORC_ERR_IO -1
#define ORC_ERR_PARSE -2
#define ORC_ERR_RANGE -3
#define ORC_ERR_ARG -4
#define ORC_ERR_SIZE -5

typedef struct orc_config {
  int mode;              /* ORC_MODE_* */
  int order;             /* ORC_ORDER_* */
  int threads;           /* >1 only with ORC_ORDER_DFS */
  double trunc_eps;      /* items with mass < trunc_eps go to ESP; 0 = off */
  long long capacity;    /* queue capacity in items; 0 = unbounded */
  double time_budget_s;  /* 0 = none; on expiry the queue folds into ESP */
  int sweep_period;      /* ORC_MODE_ALTERNATE: flip every this many FAILED
                            edges (<= 0: 2) */
  const int* chooser;    /* ORC_MODE_LEARNED: a[16*64] integer weights (row =
                            feature), s[64] signs (1 = negative unit), then
                            the margin (score units) needed to leave the
                            alternating direction */
  int prefix_edges;      /* > 0: start from the 2^t seeds of the static
                            prefix partition (original/Rel_16M.cu:97-101)
                            over the root's first t UNKNOWN edges: seed i
                            has edge j WORKING iff bit j of i is set */
  int shard, shards;     /* with prefix_edges: only seeds i % shards == shard */
} orc_config;

typedef struct orc_result {
  double R, ESP, D;            /* reliability, certified gap, dead mass */
  unsigned long long paths;    /* successes + overflowed children */
  unsigned long long nodes;    /* dequeued items incl. dead and truncated */
  unsigned long long dead, truncated, overflowed, folded;
  long long iterations;        /* levels (level order) */
  long long max_width;         /* widest level / largest queue, items */
  double elapsed_ms;           /* solve only, excludes load */
  int deadline_hit;
  int n, e;
} orc_result;

typedef struct orc_graph orc_graph;

int orc_graph_load(const char* path, orc_graph** out);
int orc_graph_from_csr(int n, int e, const int* row_ptr, const int* col_idx,
                       const int* edge_id, const double* p, orc_graph** out);
void orc_graph_free(orc_graph* g);
/* n, e, nnz, header E, type char (0 if absent) */
void orc_graph_dims(const orc_graph* g, int* n, int* e, int* nnz,
                    int* header_e, int* type);
void orc_graph_arrays(const orc_graph* g, int* row_ptr, int* col_idx,
                      int* edge_id, double* p);

int orc_solve(const orc_graph* g, int src, int dst, const orc_config* cfg,
              orc_result* out);

/* K-terminal reliability: probability that every terminal is reached from
 * terms[0] (undirected: all mutually connected).  The path becomes the
 * terminal tree of Engine::find_tree; k >= 2, terms = {s, t} is orc_solve. */
int orc_solve_terms(const orc_graph* g, const int* terms, int k, const orc_config* cfg,
                    orc_result* out);

/* Expand one item (path mode).  known/up are bit planes of W=ceil(E/64)
 * words.  Children are written in src->dst order of their failed hop:
 * child_known/child_up [max_children*W], child_mass[max_children].
 * Returns number of children, or -1 if the item is dead (dst unreachable).
 * *succ receives the success mass; *nhops the path length in edges.
 * path_eids (optional, size >= n) receives the path's edge ids src->dst. */
int orc_expand(const orc_graph* g, int src, int dst, const uint64_t* known,
               const uint64_t* up, double mass, int max_children,
               uint64_t* child_known, uint64_t* child_up, double* child_mass,
               double* succ, int* nhops, int* path_eids);

/* Expand one item under ORC_MODE_PATH or ORC_MODE_ALTERNATE (with its
 * sweep period): children in sweep order, *succ the success mass, path_eids
 * (alternate rule only) the hops in sweep order.  Returns the number of
 * children, -1 if dead, -2 on bad arguments. */
int orc_expand_rule(const orc_graph* g, int src, int dst, int mode, int period,
                    const uint64_t* known, const uint64_t* up, double mass, int max_children,
                    uint64_t* child_known, uint64_t* child_up, double* child_mass, double* succ,
                    int* path_eids);

/* orc_expand_rule for ORC_MODE_ALTERNATE, ORC_MODE_LEARNED (chooser as in
 * orc_config, eps its truncation for the headroom feature) and
 * ORC_MODE_EXPLORE; path_eids receives the hops in sweep order. */
int orc_expand_choice(const orc_graph* g, int src, int dst, int mode, int period,
                      const int* chooser, double eps, const uint64_t* known,
                      const uint64_t* up, double mass, int max_children, uint64_t* child_known,
                      uint64_t* child_up, double* child_mass, double* succ, int* path_eids);

/* Per-instance exact counts for a list of items processed to completion in
 * DFS order (full subtree under each item, with truncation eps). */
int orc_subtree_stats(const orc_graph* g, int src, int dst, int count,
                      const uint64_t* known, const uint64_t* up,
                      const double* mass, double trunc_eps,
                      unsigned long long* nodes_out, double* r_out);

#ifdef __cplusplus
}
#endif
#endif
