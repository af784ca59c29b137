/*
 * nr_cli.cpp - the `network_reliability` command line (host C++ over the
 * C-ABI of include/nr_api.h).
 *
 *   network_reliability <graph_file> <src> <dst> [--truncation_eps f]
 *       [--queue_capacity n] [--thresh_high f] [--thresh_low f]
 *       [--order level|mass|learned] [--rule path|forkbest|alternate|pivot|learned]
 *       [--sweep_period n] [--model file] [--chooser file] [--margin m]
 *       [--alpha a] [--terminals a,b,c,...] [--time_budget s] [--device d] [--json]
 *
 * Argument handling follows workingPad/main.cu:753-790: fewer than three
 * positionals print the usage on stderr and exit 1; unknown flags and flags
 * missing their value are ignored (an unknown --order or --rule value prints
 * the usage and exits 1 instead of running another method); --thresh_high/
 * --thresh_low are accepted
 * and, as in the reference (main.cu:783-786), unused.  The report reproduces
 * the printf layout of main.cu:1037-1056 and the pre-run lines of
 * graph.h:142-143 and main.cu:743-745, 838-848, 944-948.
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/nr_api.h"

static void usage(const char* prog, long long cap) {
  std::fprintf(stderr,
               "Usage: %s <graph_file> <src> <dst> [options]\n"
               "Options:\n"
               "  --truncation_eps <float>   Truncation epsilon (default 1e-15)\n"
               "  --queue_capacity <int>     Per-queue capacity  (default %lld)\n"
               "  --thresh_high <float>      High-priority log-prob threshold\n"
               "  --thresh_low  <float>      Low-priority  log-prob threshold\n"
               "  --order <level|mass|learned> Work order (default level)\n"
               "  --rule <path|forkbest|alternate|pivot|learned> Decomposition rule (default path)\n"
               "  --sweep_period <int>       Alternate rule: flip every n levels (default 2)\n"
               "  --model <file>             Learned order: NR_MLP_PARAMS raw float32\n"
               "  --alpha <float>            Learned order: score weight (default 0.5)\n"
               "  --chooser <file>           Learned rule: NR_MLP_PARAMS raw float32\n"
               "  --margin <float>           Learned rule: log2 advantage to leave the\n"
               "                             alternating direction (default 0)\n"
               "  --terminals <a,b,...>      K-terminal reliability (first is the root)\n"
               "  --time_budget <seconds>    Fold the queue into ESP after this\n"
               "  --device <int>             HIP device ordinal\n"
               "  --json                     Print one JSON line instead\n",
               prog, cap);
}

int main(int argc, char** argv) {
  nr_config cfg;
  nr_config_init(&cfg);
  if (argc < 4) {
    usage(argv[0], cfg.queue_capacity);
    return 1;
  }
  const char* file = argv[1];
  int src = std::atoi(argv[2]);
  int dst = std::atoi(argv[3]);
  bool json = false;
  const char* model_file = nullptr;
  const char* chooser_file = nullptr;
  const char* terminals = nullptr;
  float alpha = 0.5f;
  float margin = 0.0f; /* learned rule: prediction advantage to leave the alternation */
  double thresh_high = -4.60517, thresh_low = -9.21034; /* common.h:35-36 */
  for (int i = 4; i < argc; ++i) {
    const bool has = i + 1 < argc;
    if (!std::strcmp(argv[i], "--truncation_eps") && has) {
      cfg.truncation_eps = std::atof(argv[++i]);
    } else if (!std::strcmp(argv[i], "--queue_capacity") && has) {
      cfg.queue_capacity = std::atoll(argv[++i]);
    } else if (!std::strcmp(argv[i], "--thresh_high") && has) {
      thresh_high = std::atof(argv[++i]);
    } else if (!std::strcmp(argv[i], "--thresh_low") && has) {
      thresh_low = std::atof(argv[++i]);
    } else if (!std::strcmp(argv[i], "--order") && has) {
      const char* o = argv[++i];
      if (!std::strcmp(o, "level")) {
        cfg.order = NR_ORDER_LEVEL;
      } else if (!std::strcmp(o, "mass")) {
        cfg.order = NR_ORDER_MASS;
      } else if (!std::strcmp(o, "learned")) {
        cfg.order = NR_ORDER_LEARNED;
      } else { /* never run another method than the one asked for */
        std::fprintf(stderr, "unknown --order %s\n", o);
        usage(argv[0], cfg.queue_capacity);
        return 1;
      }
    } else if (!std::strcmp(argv[i], "--rule") && has) {
      const char* r = argv[++i];
      if (!std::strcmp(r, "path")) {
        cfg.rule = NR_RULE_PATH;
      } else if (!std::strcmp(r, "forkbest")) {
        cfg.rule = NR_RULE_FORKBEST;
      } else if (!std::strcmp(r, "alternate")) {
        cfg.rule = NR_RULE_ALTERNATE;
      } else if (!std::strcmp(r, "pivot")) {
        cfg.rule = NR_RULE_PIVOT;
      } else if (!std::strcmp(r, "learned")) {
        cfg.rule = NR_RULE_LEARNED;
      } else {
        std::fprintf(stderr, "unknown --rule %s\n", r);
        usage(argv[0], cfg.queue_capacity);
        return 1;
      }
    } else if (!std::strcmp(argv[i], "--sweep_period") && has) {
      cfg.sweep_period = std::atoi(argv[++i]);
    } else if (!std::strcmp(argv[i], "--model") && has) {
      model_file = argv[++i];
    } else if (!std::strcmp(argv[i], "--chooser") && has) {
      chooser_file = argv[++i];
    } else if (!std::strcmp(argv[i], "--margin") && has) {
      margin = (float)std::atof(argv[++i]);
    } else if (!std::strcmp(argv[i], "--terminals") && has) {
      terminals = argv[++i];
    } else if (!std::strcmp(argv[i], "--alpha") && has) {
      alpha = (float)std::atof(argv[++i]);
    } else if (!std::strcmp(argv[i], "--time_budget") && has) {
      cfg.time_budget_s = std::atof(argv[++i]);
    } else if (!std::strcmp(argv[i], "--device") && has) {
      cfg.device = std::atoi(argv[++i]);
    } else if (!std::strcmp(argv[i], "--json")) {
      json = true;
    }
  }
  (void)thresh_high;
  (void)thresh_low;

  nr_graph* g = nullptr;
  int rc = nr_graph_load(file, src, dst, &g);
  if (rc != NR_OK) {
    std::fprintf(stderr, "Cannot load %s: %s\n", file, nr_status_str(rc));
    return 1;
  }
  if (terminals) {
    int t[256], k = 0;
    for (const char* c = terminals; *c && k < 256;) {
      t[k++] = std::atoi(c);
      while (*c && *c != ',') ++c;
      if (*c == ',') ++c;
    }
    rc = nr_graph_set_terminal_set(g, t, k);
    if (rc != NR_OK) {
      std::fprintf(stderr, "bad --terminals: %s\n", nr_status_str(rc));
      nr_graph_destroy(g);
      return 1;
    }
  }
  nr_graph_info info;
  nr_graph_get_info(g, &info);
  if (info.header_e != info.e)
    std::fprintf(stderr, "Warning: header E=%d, found %d unique edges\n", info.header_e,
                 info.e);
  if (!json) {
    std::printf("Graph loaded: N=%d  E=%d  mask_words=%d  src=%d  dst=%d\n", info.n, info.e,
                info.mask_words, src, dst);
    if (info.precontracted || info.predeleted)
      std::printf("Preprocessing: %d edges contracted (p=1), %d deleted (p=0)\n",
                  info.precontracted, info.predeleted);
    char name[256] = {0};
    int cus = 0, ws = 0;
    long long mem = 0;
    if (nr_device_query(cfg.device, name, sizeof name, &cus, &ws, &mem) == NR_OK)
      std::printf("GPU: %s, %d CUs, wave%d, %.2f GB\n", name, cus, ws,
                  mem / (1024.0 * 1024.0 * 1024.0));
    const double item_bytes = 16.0 * info.mask_words + 8.0;
    std::printf("Queue capacity: %lld items (%.2f MB per buffer x 2)\n", cfg.queue_capacity,
                cfg.queue_capacity * item_bytes / (1024.0 * 1024.0));
    std::fflush(stdout);
  }
  nr_result r;
  if (cfg.order == NR_ORDER_LEARNED || cfg.rule == NR_RULE_LEARNED) {
    /* raw float32 parameter files: the order scorer, the path chooser */
    auto read_params = [](const char* path, float* params) {
      FILE* f = path ? std::fopen(path, "rb") : nullptr;
      const size_t got = f ? std::fread(params, sizeof(float), NR_MLP_PARAMS, f) : 0;
      if (f) std::fclose(f);
      return got == NR_MLP_PARAMS;
    };
    static float params[NR_MLP_PARAMS], cparams[NR_MLP_PARAMS];
    const bool need_order = cfg.order == NR_ORDER_LEARNED;
    const bool need_chooser = cfg.rule == NR_RULE_LEARNED;
    if ((need_order && !read_params(model_file, params)) ||
        (need_chooser && !read_params(chooser_file, cparams))) {
      std::fprintf(stderr,
                   "--order learned needs --model and --rule learned needs --chooser, each "
                   "a file of %d float32 parameters\n",
                   NR_MLP_PARAMS);
      nr_graph_destroy(g);
      return 1;
    }
    nr_model* m = nullptr;
    rc = nr_model_create(need_order ? params : nullptr, alpha, &m);
    if (rc == NR_OK && need_chooser) rc = nr_model_set_chooser(m, cparams, margin);
    if (rc == NR_OK) rc = nr_solve_learned(g, &cfg, m, &r);
    nr_model_destroy(m);
  } else {
    rc = nr_solve(g, &cfg, &r);
  }
  nr_graph_destroy(g);
  if (rc != NR_OK) {
    std::fprintf(stderr, "solve failed: %s\n", nr_status_str(rc));
    return 1;
  }
  if (json) {
    std::printf(
        "{\"R\": %.17g, \"ESP\": %.17g, \"D\": %.17g, \"paths\": %llu, \"nodes\": %llu, "
        "\"dead\": %llu, \"truncated\": %llu, \"overflowed\": %llu, \"folded\": %llu, "
        "\"iterations\": %lld, \"max_queue_items\": %lld, \"item_bytes\": %lld, "
        "\"elapsed_ms\": %.6f, \"kernel_ms\": %.6f, \"kernel_launches\": %lld, "
        "\"deadline_hit\": %d, \"N\": %d, \"E\": %d, \"src\": %d, \"dst\": %d}\n",
        r.R, r.ESP, r.D, r.paths, r.nodes, r.dead, r.truncated, r.overflowed, r.folded,
        r.iterations, r.max_queue_items, r.item_bytes, r.elapsed_ms, r.kernel_ms,
        r.kernel_launches, r.deadline_hit, info.n, info.e, src, dst);
    return 0;
  }
  std::printf("Occupancy: %d blocks/CU x %d CUs = %d blocks total\n", r.blocks_per_cu,
              r.compute_units, r.grid_blocks);
  std::printf("Launch config: %d blocks x %d threads, lds=%d B, CUs=%d\n", r.grid_blocks,
              r.block_size, r.lds_bytes, r.compute_units);
  const double rss = (double)r.max_queue_items * (double)r.item_bytes;
  std::printf("\n");
  std::printf("============================================\n");
  std::printf("  Two-Terminal Network Reliability Results\n");
  std::printf("============================================\n");
  std::printf("Graph:            N = %d, E = %d\n", info.n, info.e);
  std::printf("Terminals:        src = %d, dst = %d\n", src, dst);
  std::printf("--------------------------------------------\n");
  std::printf("Reliability:      R(G) = %.12f\n", r.R);
  std::printf("R + ESP:                 %.12f\n", r.R + r.ESP);
  std::printf("ESP (truncated):         %.6e\n", r.ESP);
  std::printf("Truncation bound:        %.6e\n", (r.R > 0) ? r.ESP / r.R : r.ESP);
  std::printf("--------------------------------------------\n");
  std::printf("Paths enumerated: %llu\n", r.paths);
  std::printf("Nodes processed:  %llu\n", r.nodes);
  std::printf("Kernel iterations:%lld\n", r.iterations);
  std::printf("Max queue RSS:    %.3f MB  (%lld items)\n", rss / (1024.0 * 1024.0),
              r.max_queue_items);
  std::printf("Total time:       %.3f ms\n", r.elapsed_ms);
  std::printf("============================================\n");
  return 0;
}
