/*
 * nr_api.h - C-ABI of the MI355X two-terminal reliability engine
 * (path decomposition / factoring, hand-written HIP kernels for gfx950).
 *
 * Plain C types only: no HIP, no torch, no C++ in the signatures.  Every
 * entry point returns an nr_status (0 = NR_OK); none calls exit() or prints.
 * Calls are blocking; the library owns every device allocation it makes and
 * the caller owns its input arrays for the duration of a call.  Calls on one
 * nr_graph are not reentrant.
 *
 * Reference surface each entry point replaces (paths relative to the
 * reference repository):
 *   nr_graph_load      workingPad/graph.h:46-145   graph_load()
 *   nr_graph_create    workingPad/graph.h:104-145  CSR build + probabilities
 *   nr_graph_destroy   workingPad/graph.h:181-188  graph_free_device()
 *   nr_solve           workingPad/main.cu:795-1056 main(): initial mask
 *                      (717-748), queue setup (635-711, 847-880), level loop
 *                      (963-1008), factoring_kernel (465-629), stats readback
 *                      (1017-1029)
 *   nr_solve_batch     parallel/CUDA_Reliability.h:95-122 (the reference's
 *                      only extern "C" host API: batched GPU work over many
 *                      masks) generalised to many graphs in one launch
 *   nr_expand_items    workingPad/main.cu:114-283 path_edge_sweep_warp() on a
 *                      caller-supplied batch of items (one expansion each)
 *   nr_config_init     workingPad/common.h:33-36 defaults, main.cu:753-790
 */
#ifndef NR_API_H
#define NR_API_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum nr_status {
  NR_OK = 0,
  NR_ERR_IO = 1,          /* graph file cannot be opened */
  NR_ERR_PARSE = 2,       /* malformed graph file */
  NR_ERR_RANGE = 3,       /* node/edge index or terminal out of range */
  NR_ERR_ARG = 4,         /* null pointer or invalid config value */
  NR_ERR_UNSUPPORTED = 5, /* graph shape beyond the engine's width classes */
  NR_ERR_DEVICE = 6,      /* HIP runtime error */
  NR_ERR_NOMEM = 7,       /* device allocation failed */
  NR_ERR_SCRATCH = 8      /* nr_solve_static: the solve needs more device
                             memory than the fixed scratch area holds */
} nr_status;

/* Work order.  Any order gives an exact decomposition; order only changes
 * which subproblems are resolved before a capacity/time limit. */
typedef enum nr_order {
  NR_ORDER_LEVEL = 0, /* FIFO by tree depth: the reference's order
                         (main.cu:963-1008); reports levels as iterations */
  NR_ORDER_MASS = 1,  /* max-mass first (bucketed device priority queue);
                         the reference's CPU heap order (seq/seq.cpp:49-197) */
  NR_ORDER_LEARNED = 2 /* learned scorer orders the queue: nr_solve_learned
                          (nr_solve returns NR_ERR_ARG: it has no model) */
} nr_order;

typedef struct nr_config {
  long long queue_capacity; /* items per queue buffer (--queue_capacity,
                               main.cu:781-782); default 1<<20 */
  double truncation_eps;    /* items with mass < eps go to ESP (main.cu:537);
                               FP64; 0 disables; default 1e-15 */
  double time_budget_s;     /* 0 = none; when it expires every queued item is
                               folded into ESP, keeping R <= R_true <= R+ESP */
  int order;                /* nr_order */
  int device;               /* HIP device ordinal */
  int block_size;           /* threads per block: 64, 128 or 256; other
                               values (0) = auto, the largest that fits LDS */
  int grid_blocks;          /* blocks per launch; 0 = auto from occupancy */
  int chunk_items;          /* nr_solve_batch: the queue chunk, the number of
                               items the refill between two levels tops the
                               next level up to (spilled children first, then
                               new instance roots); 0 = a quarter of the
                               capacity; nr_tune_batch picks it */
  int flags;                /* NR_FLAG_* */
  int rule;                 /* nr_rule: which decomposition each item gets */
  int sweep_period;         /* NR_RULE_ALTERNATE: the sweep direction flips
                               every this many FAILED edges (tree levels);
                               0 = 2 */
  int prefix_edges;         /* static prefix partition (the reference's
                               TID_TO_MASK, original/Rel_16M.cu:97-101): the
                               root is split over its first prefix_edges
                               UNKNOWN edges (edge id order) into 2^t seed
                               subproblems, seed i with edge j WORKING when
                               bit j of i is set and FAILED otherwise, mass
                               the product of their p and q; the solve starts
                               from the seeds (0 = from the root; at most 24).
                               Exact: the seeds partition the states */
  int shard, shards;        /* with prefix_edges: only the seeds i with
                               i % shards == shard (0 <= shard < shards;
                               shards 0 or 1 = all); the shards' R, D, ESP
                               and counts add up to the whole solve's */
  int ramp_items;           /* fresh level-order solves of the path and
                               alternating rules: after the levels one
                               workgroup runs alone, every workgroup expands
                               its own slice's subtrees level after level
                               (no launch per level) until the levels hold
                               about this many items; 0 = 1024 per compute
                               unit (four workgroups per unit; at most a
                               quarter of the workgroups' region slots),
                               < 0 = one launch per level from there.  Same
                               results either way: a ramp that overflows a
                               workgroup's region is discarded and the solve
                               runs again without it (nr_result.rerun) */
} nr_config;

/* Decomposition rule (nr_config.rule).  Both are exact; they build
 * different trees. */
typedef enum nr_rule {
  NR_RULE_PATH = 0,    /* k children per path (main.cu:114-283), default */
  NR_RULE_FORKBEST = 1, /* success on a confirmed WORKING path, else a two-way
                          branch on the path's most reliable UNKNOWN edge
                          (FP32 logf keys), ties settled as the reference's
                          warp argmax settles them (main.cu:304-430, 550-567) */
  NR_RULE_ALTERNATE = 2, /* no counterpart in the reference: k children per
                          path as NR_RULE_PATH, on a path with the fewest
                          UNKNOWN edges (WORKING edges contracted; walk back
                          to the lowest-index node of the previous level,
                          then the lowest-index node of the component it
                          joins), children swept from the src end when
                          (FAILED edges / sweep_period) is even and from the
                          dst end otherwise.  Undirected graphs of up to 256
                          nodes, two terminals; NR_ERR_UNSUPPORTED
                          otherwise */
  NR_RULE_LEARNED = 4,  /* no counterpart in the reference: the path of
                          NR_RULE_ALTERNATE, swept from the end the model's
                          path chooser predicts the smaller subtree for, by
                          more than its margin (nr_model_set_chooser; else
                          the alternating direction).  nr_solve_learned only;
                          same graphs as NR_RULE_ALTERNATE */
  NR_RULE_PIVOT = 3     /* edge-pivot factoring (workingPad/cpu_reference.cpp
                          :107-145): success on a confirmed WORKING path,
                          dead when dst is unreachable, else a two-way branch
                          on the UNKNOWN edge of highest p over the whole
                          graph, lowest id on ties (WORKING child first);
                          graphs of up to 256 nodes, orders as forkBest */
} nr_rule;

/* nr_config.flags */
#define NR_FLAG_TIME_LAUNCHES 1 /* time every expand launch with HIP events
                                   (fills kernel_ms / kernel_launches; adds a
                                   few microseconds between launches) */
#define NR_FLAG_ROWS 4          /* undirected graphs of <= 64 nodes: items always
                                   carry their adjacency rows (by default they
                                   do only while the rows take <= 64 bytes, and
                                   are derived from the planes otherwise; same
                                   results) */
#define NR_FLAG_LEAN 8          /* items never carry rows (NR_FLAG_ROWS wins when
                                   both are set); same results */
#define NR_FLAG_QUEUE_ALL 16    /* level order of the path rule on a row-form
                                   graph: queue every child, as the reference
                                   does, instead of counting the dead ones and
                                   those below the truncation when they are
                                   created; same results */

typedef struct nr_result {
  double R;                     /* reliability lower bound (exact when ESP=0) */
  double ESP;                   /* certified gap: truncated + overflowed +
                                   folded mass */
  double D;                     /* mass of dead subproblems; R+D+ESP = 1 */
  unsigned long long paths;     /* successes + overflowed children
                                   (main.cu:251,262) */
  unsigned long long nodes;     /* dequeued items incl. dead and truncated
                                   (main.cu:533-534) */
  unsigned long long dead, truncated, overflowed, folded;
  long long iterations;         /* levels (level order) or refill rounds */
  long long max_queue_items;    /* widest level / largest queue occupancy */
  long long item_bytes;         /* bytes per queued item in HBM */
  double elapsed_ms;            /* first launch to final sync, excludes load
                                   and allocation (main.cu:951-1014) */
  double kernel_ms;             /* summed expand-kernel time (HIP events) */
  long long kernel_launches;    /* expand-kernel launches */
  int deadline_hit;
  int overflowed_flag;
  int grid_blocks;              /* launch geometry actually used */
  int block_size;
  int lds_bytes;                /* dynamic LDS per workgroup */
  int blocks_per_cu;            /* occupancy reported by the runtime */
  int compute_units;
  int rerun;                    /* nr_solve_batch: 1 if the instance overflowed
                                   the shared queue and was solved again alone
                                   (same capacity: overflowed > 0 if it
                                   overflows even alone);
                                   nr_solve: 1 if the ramp (ramp_items)
                                   overflowed and the solve ran again without
                                   it */
} nr_result;

typedef struct nr_graph_info {
  int n, e, nnz;
  int header_e;   /* E from the header (halved for type U, graph.h:93) */
  int type;       /* 'U', 'D', or 0 when the header has no type token */
  int src, dst;
  int mask_words; /* 64-bit words per edge-state plane */
  int symmetric;  /* every CSR entry u->v has a matching v->u entry */
  int precontracted, predeleted; /* p>=1 / p<=0 edges (main.cu:729-745) */
} nr_graph_info;

typedef struct nr_graph nr_graph;

/* Defaults of workingPad/common.h:33-34: capacity 2^20, eps 1e-15. */
int nr_config_init(nr_config* cfg);

/* Parse the reference's adjacency-list file: header "N E [T]" (the type
 * token is optional), then per node "deg (nbr p)*"; CRLF accepted; p parsed
 * as FP64.  Edge ids follow first appearance of (min,max), CSR rows follow
 * the file. */
int nr_graph_load(const char* path, int src, int dst, nr_graph** out);

/* Build from CSR arrays (row_ptr[n+1], col_idx[nnz], edge_id[nnz], p[e]). */
int nr_graph_create(int n, int e, const int* row_ptr, const int* col_idx,
                    const int* edge_id, const double* p, int src, int dst,
                    nr_graph** out);

int nr_graph_get_info(const nr_graph* g, nr_graph_info* info);

/* Copy out the host CSR (any pointer may be null). */
int nr_graph_get_csr(const nr_graph* g, int* row_ptr, int* col_idx,
                     int* edge_id, double* p);

/* Two-terminal problem src -> dst (clears a terminal set). */
int nr_graph_set_terminals(nr_graph* g, int src, int dst);

/* K-terminal reliability (no counterpart in the reference, whose target_set
 * is declared but unused: original/Rel_16M.cu:43, original/Seq_16.cpp:33):
 * the probability that every terminal is reached from terminals[0] (on an
 * undirected graph: that all are connected).  Each subproblem branches on
 * the UNKNOWN edges of a terminal tree (BFS from terminals[0]; the others,
 * in list order, walk back to the tree under the lowest-index parent rule)
 * instead of a path; {s, t} gives the two-terminal tree exactly.  k = 0
 * returns to the src/dst pair; k = 1 is an error; at most 256 terminals.
 * Not available to nr_solve_batch. */
int nr_graph_set_terminal_set(nr_graph* g, const int* terminals, int k);
int nr_graph_get_terminal_set(const nr_graph* g, int* terminals, int max_k, int* k);

void nr_graph_destroy(nr_graph* g);

/* Solve one instance on the GPU. */
int nr_solve(const nr_graph* g, const nr_config* cfg, nr_result* out);

/* nr_solve that allocates nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Its queue
 * buffers, graph tables, per-workgroup partial sums and control block are
 * carved from a fixed 64 MiB device array (NR_STATIC_SCRATCH_BYTES, one per
 * GPU, kept until the process ends) on the null stream, with no events and
 * no pinned host memory; results equal nr_solve's.  Sizes beyond the array
 * (queue_capacity * item bytes * 2 plus tables and slots) return
 * NR_ERR_SCRATCH.  Graphs past the per-lane forms (more than 256 nodes or
 * 512 edges) return NR_ERR_UNSUPPORTED; elapsed_ms is host time.  Calls on
 * one GPU are serialised by the library within a process; the array keeps
 * the last call's contents (each call clears what it needs).  kernel_ms and
 * kernel_launches stay 0. */
#define NR_STATIC_SCRATCH_BYTES (64ll << 20)
int nr_solve_static(const nr_graph* g, const nr_config* cfg, nr_result* out);

/* Solve `count` independent instances (graphs and/or terminal pairs) in one
 * persistent launch sequence; results[i] belongs to graphs[i].  Counts are
 * per instance; elapsed_ms/kernel_ms are the whole batch's. */
int nr_solve_batch(const nr_graph* const* graphs, int count,
                   const nr_config* cfg, nr_result* results);

/* Launch schedule of nr_solve_batch for these instances, picked by timing
 * trial batches of the first `sample` instances (all of them when 0):
 * workgroups of 256, 128 or 64 threads and 1, 2 or 4 workgroups per CU or
 * the occupancy limit; then, for the fastest of those, a queue chunk
 * (chunk_items) of an eighth, a quarter or half of the capacity; the rest of
 * cfg (or the defaults) unchanged.  *tuned receives cfg with block_size,
 * grid_blocks and chunk_items of the fastest trial, and *best_rate
 * (optional) its subproblems per second.  No counterpart in the reference
 * (its launch shape is fixed, main.cu:882-948). */
int nr_tune_batch(const nr_graph* const* graphs, int count, const nr_config* cfg,
                  int sample, nr_config* tuned, double* best_rate);

/* Expand `count` caller-supplied items once (one subproblem each) on the
 * GPU.  Item i has planes known[i*W..], up[i*W..] (W = mask_words) and mass
 * mass[i].  Outputs: nchild[i] (-1 dead), succ[i], and children in slots
 * i*max_children.. of child_known/child_up (W words each) and child_mass,
 * ordered by their failed hop from src to dst. */
int nr_expand_items(const nr_graph* g, int device, int count,
                    const unsigned long long* known,
                    const unsigned long long* up, const double* mass,
                    int max_children, int* nchild, double* succ,
                    unsigned long long* child_known,
                    unsigned long long* child_up, double* child_mass);

/* Sweep period of NR_RULE_ALTERNATE for graph g, picked by pilot solves:
 * level order at truncation pilot_eps (> 0) for periods 1..max_period (at
 * most 64), the other settings from cfg (or the defaults).  *best = the
 * period with the smallest certified gap, the fewest subproblems on ties;
 * pilot_esp[0..max_period-1] (optional) receives each pilot's ESP.  No
 * counterpart in the reference. */
int nr_tune_sweep_period(const nr_graph* g, const nr_config* cfg, double pilot_eps,
                         int max_period, int* best, double* pilot_esp);

/* nr_expand_items() under cfg->rule (NR_RULE_PATH or NR_RULE_ALTERNATE with
 * cfg->sweep_period) on device cfg->device; children in sweep order.  No
 * counterpart in the reference beyond nr_expand_items'. */
int nr_expand_items_rule(const nr_graph* g, const nr_config* cfg, int count,
                         const unsigned long long* known,
                         const unsigned long long* up, const double* mass,
                         int max_children, int* nchild, double* succ,
                         unsigned long long* child_known,
                         unsigned long long* child_up, double* child_mass);

const char* nr_status_str(int status);

/* Free the per-device workspace (queue buffers, graph tables, stream,
 * events) that solves on `device` keep between calls; the next call
 * reallocates it.  Replaces the end-of-run frees of workingPad/main.cu:
 * 1058-1069 and graph_free_device (graph.h:181-188). */
int nr_device_release(int device);

/* Device name / compute units / wavefront size of a HIP device. */
int nr_device_query(int device, char* name, int name_len, int* compute_units,
                    int* wave_size, long long* total_mem);

/* ------------------------------------------------------------------------
 * Checkpoint / resume of a time-budgeted solve (original/Seq_16.cpp:44,227
 * folds the pending queue into ESP when the time is up; seq/seq.cpp:189-196
 * reports the pending mass).  Here the pending subproblems can instead be
 * kept and the solve continued later, on this or another process or device.
 * ------------------------------------------------------------------------ */
typedef struct nr_checkpoint nr_checkpoint;

/* nr_solve that, when cfg->time_budget_s expires first, keeps the pending
 * subproblems in *kept (else *kept = NULL): the result is then the bracket
 * so far, ESP including the pending mass (folded stays 0).  With `resume`
 * (from an earlier call on the same graph, rule and order) the solve starts
 * from its pending subproblems, sums and counters instead of the root; a
 * chain of such calls ends with the counts, levels and sums of one
 * uninterrupted solve.  The queue capacity must hold the pending items. */
int nr_solve_checkpointed(const nr_graph* g, const nr_config* cfg,
                          const nr_checkpoint* resume, nr_checkpoint** kept,
                          nr_result* out);
int nr_checkpoint_info(const nr_checkpoint* ckpt, long long* pending,
                       double* pending_mass, long long* nodes_done);
/* Binary file: magic, shape, graph fingerprint, sums, counters, items. */
int nr_checkpoint_save(const nr_checkpoint* ckpt, const char* path);
int nr_checkpoint_load(const char* path, nr_checkpoint** out);
void nr_checkpoint_destroy(nr_checkpoint* ckpt);

/* ------------------------------------------------------------------------
 * Learned work order.  No counterpart in the reference: a small MLP scores
 * each child subproblem (predicted log2 of its subtree size under the plain
 * method) and the class-round queue expands children in order of
 * log2(mass) - alpha * score.  Any order expands the same exact
 * decomposition; the score only changes what is resolved before a time or
 * capacity limit, and R <= R_true <= R + ESP holds whatever it predicts.
 *
 * Features (16, see relaibilitygpu_amd/csrc/nr_learn.h child_features): the
 * child's log2 mass, its UNKNOWN/WORKING/FAILED edge fractions, the failed
 * hop's position on the parent's path, the path's length and unknown hops,
 * q of the failed edge, the path suffix's log2 success probability, mass
 * headroom over eps, allowed degrees of src and dst, terminal incidence,
 * the parent's log2 mass, remaining unknown hops, and a constant 1.
 * Model: W1[16][64] (row 15 = hidden bias), ReLU, w2[64], b2: NR_MLP_PARAMS
 * floats in that order.
 * ------------------------------------------------------------------------ */
#define NR_MLP_FEATURES 16
#define NR_MLP_HIDDEN 64
#define NR_MLP_PARAMS (16 * 64 + 64 + 1)

typedef struct nr_model nr_model;

/* ------------------------------------------------------------------------
 * Learned path choice (NR_RULE_LEARNED).  No counterpart in the reference:
 * its one path rule (device_code.cuh:80-194, main.cu:114-283) is joined by a
 * per-subproblem choice between two exact decompositions of the same
 * fewest-UNKNOWN path: children swept from the src end or from the dst end.
 * A second network of the same shape (16 -> 64 ReLU -> 1) predicts log2 of
 * the subproblem's subtree size for each direction (trained on choice logs:
 * a direction, then the alternating rule below); the sweep leaves the
 * alternating rule's direction when the other is predicted smaller by more
 * than a margin.
 *
 * Choice features (NR_CHOICE_FEATURES integers, |x| <= 63), for a sweep that
 * starts at end A (src or dst) and ends at B; "component" is the end's node
 * set joined by WORKING edges:
 *   0 UNKNOWN hops of the path        1 nodes of A's component
 *   2 UNKNOWN edges leaving A's comp. 3 FAILED edges leaving A's component
 *   4-6 the same for B                7 32*FAILED/E  8 32*UNKNOWN/E
 *   9 0 (unused)                      10 floor((floor(log2 mass) - le) / 2)
 *                                        in 0..32, le = floor(log2 eps), or
 *                                        -64 when eps < 2^-64 or 0
 *  11 -sum over the hops of lq(p)    12 -lq(q) of A's hop  13 -lq(q) of B's
 *  14 (2) - (5)                       15 constant 1
 * where lq(x) = round(-4 log2 x) clamped to 0..63, and features 0-6, 11, 14
 * are clamped to +-63.  Only mass relative to the truncation enters (10), so
 * a solve truncated deeper than the training logs, or not at all, shows the
 * chooser the same kind of states.  The inference form is exact: weights a/256 with
 * integer |a| <= 255 (W1 times |w2|, one scale), so BF16 operands and FP32
 * sums are exact and the choice is an integer comparison.
 * ------------------------------------------------------------------------ */
#define NR_CHOICE_FEATURES 16

/* Attach a path chooser (NR_MLP_PARAMS floats in the layout of the order
 * scorer, trained on choice features) to a model.  A subproblem leaves the
 * alternating direction only when the other direction's predicted log2
 * subtree size is smaller by more than `margin` (>= 0; a huge margin keeps
 * the alternating rule). */
int nr_model_set_chooser(nr_model* model, const float* params, float margin);

/* The chooser's exact integer form: weights[16*64] (row k = feature, column
 * h = hidden unit, whose value is relu(sum_k x_k a[k][h])), signs[64] (1:
 * the unit counts negatively; a direction's score is the signed sum of the
 * units) and the margin in score units: the sweep leaves the alternating
 * direction d when score(d) - score(other) > margin.  NR_ERR_ARG without a
 * chooser. */
int nr_model_get_chooser(const nr_model* model, int* weights, int* signs, int* margin);

/* Solve in level order with the alternating rule (cfg->sweep_period) while
 * the sweep direction explores: it is flipped for a reproducible
 * pseudo-random quarter of the subproblems (the top two bits of a
 * splitmix64-finaliser chain over the planes are zero), so the labels
 * measure each direction followed by the rule the chooser improves on.
 * Logs for every expanded
 * subproblem with at least two UNKNOWN hops its choice features of the
 * direction taken; labels = log2 of its subtree nodes under this policy.
 * Sampling and the other arguments as nr_log_subproblems. */
int nr_log_choices(const nr_graph* g, const nr_config* cfg, long long max_records,
                   unsigned long long seed, float* features, float* labels,
                   long long* n_out, nr_result* result);

/* Solve in level order while logging every child (features at creation,
 * parent, depth), then derive each child's subtree size bottom-up on the
 * device and return up to max_records of them: the records with the
 * smallest seeded hashes of their content (features and size), in hash
 * order — a uniform sample without replacement that is the same for the
 * same tree and seed whatever order the device logged it in.
 * features[max_records * 16], labels = log2(subtree nodes).  The solve's
 * own result goes to *result.  Fails with NR_ERR_ARG if the queue
 * overflowed (the tree would be incomplete). */
int nr_log_subproblems(const nr_graph* g, const nr_config* cfg, long long max_records,
                       unsigned long long seed, float* features, float* labels,
                       long long* n_out, nr_result* result);

/* nr_log_subproblems with a second label per record: cond_rel[max_records]
 * receives the child's conditional reliability, the success mass its
 * subtree finds (main.cu:254-262 summed bottom-up over the subtree) divided
 * by its own mass.  For an untruncated subtree this is P(src reaches dst |
 * the child's WORKING and FAILED edges); under truncation it is the lower
 * end of that probability's bracket.  Accumulated on the device in 62-bit
 * fixed point relative to each record's mass (exactly repeatable; absolute
 * error below depth * children * 2^-62).  The sample, features and size
 * labels are those nr_log_subproblems returns for the same arguments.
 * No counterpart in the reference (its tree is never logged). */
int nr_log_subproblems_rel(const nr_graph* g, const nr_config* cfg, long long max_records,
                           unsigned long long seed, float* features, float* labels,
                           double* cond_rel, long long* n_out, nr_result* result);

typedef struct nr_train_config {
  int epochs;         /* default 8 */
  int batch;          /* samples per Adam step, default 4096 */
  float lr;           /* default 3e-3 */
  float weight_decay; /* decoupled, default 1e-5 */
  unsigned long long seed;
  int device;
  float holdout;      /* fraction held out for the holdout loss, default 0.1 */
  int reserved[4];
} nr_train_config;

int nr_train_config_init(nr_train_config* cfg);

/* He-style random initial parameters (host). */
int nr_mlp_init(unsigned long long seed, float* params);

/* Mean loss (y_hat - y)^2 / 2 over n samples and its gradient with respect
 * to the NR_MLP_PARAMS parameters, computed by the training kernels. */
int nr_mlp_grad(const float* X, const float* y, long long n, const float* params,
                float* grad, double* loss, int device);

/* Train on the GPU (FP32 forward / backward kernels, Adam); params is the
 * initial point on entry and the result on exit; per-epoch mean training
 * and holdout losses (arrays of cfg->epochs, may be null). */
int nr_mlp_train(const float* X, const float* y, long long n, const nr_train_config* cfg,
                 float* params, float* train_loss, float* holdout_loss);

/* Inference form of trained parameters (BF16 weights, FP32 accumulation);
 * params may be NULL for a model that only carries a path chooser. */
int nr_model_create(const float* params, float alpha, nr_model** out);
void nr_model_destroy(nr_model* model);

/* Scores of n feature rows by the inference kernel (BF16 matrix cores). */
int nr_model_score(const nr_model* model, const float* features, long long n,
                   float* scores, int device);

/* nr_solve with a model: cfg->order NR_ORDER_LEARNED orders the queue by
 * the model's order scorer (NR_ORDER_LEVEL and NR_ORDER_MASS as nr_solve);
 * cfg->rule NR_RULE_LEARNED picks sweeps by its path chooser (NR_RULE_PATH
 * and NR_RULE_ALTERNATE as nr_solve).  NR_ERR_ARG when the model lacks the
 * part asked for. */
int nr_solve_learned(const nr_graph* g, const nr_config* cfg, const nr_model* model,
                     nr_result* out);

#ifdef __cplusplus
}
#endif
#endif /* NR_API_H */
