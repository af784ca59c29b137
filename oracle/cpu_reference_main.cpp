/*
 * cpu_reference_main.cpp - command line for the CPU oracle (test
 * infrastructure; see oracle.h).
 *
 *   cpu_reference <graph_file> <src> <dst> [--mode path|pivot|forkbest|minunk|alternate]
 *       [--order level|heap|dfs] [--threads T] [--truncation_eps f]
 *       [--queue_capacity n] [--time_budget s] [--sweep_period n] [--json]
 *
 * The text report follows the layout of workingPad/main.cu:1037-1056 so the
 * oracle and the engine can be diffed line by line (minus the timing line).
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "oracle.h"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr,
                 "Usage: %s <graph_file> <src> <dst> [--mode path|pivot|forkbest|minunk|alternate] "
                 "[--order level|heap|dfs] [--threads T] [--truncation_eps f] "
                 "[--queue_capacity n] [--time_budget s] [--sweep_period n] [--json]\n",
                 argv[0]);
    return 1;
  }
  orc_config cfg;
  std::memset(&cfg, 0, sizeof cfg);
  cfg.mode = ORC_MODE_PATH;
  cfg.order = ORC_ORDER_LEVEL;
  cfg.threads = 1;
  bool json = false;
  for (int i = 4; i < argc; ++i) {
    const char* a = argv[i];
    bool has = i + 1 < argc;
    if (!std::strcmp(a, "--mode") && has) {
      ++i;
      cfg.mode = !std::strcmp(argv[i], "pivot")      ? ORC_MODE_PIVOT
                 : !std::strcmp(argv[i], "forkbest") ? ORC_MODE_FORKBEST
                 : !std::strcmp(argv[i], "minunk")   ? ORC_MODE_MINUNK
                 : !std::strcmp(argv[i], "alternate") ? ORC_MODE_ALTERNATE
                                                     : ORC_MODE_PATH;
    } else if (!std::strcmp(a, "--order") && has) {
      ++i;
      cfg.order = !std::strcmp(argv[i], "heap")  ? ORC_ORDER_HEAP
                  : !std::strcmp(argv[i], "dfs") ? ORC_ORDER_DFS
                                                 : ORC_ORDER_LEVEL;
    } else if (!std::strcmp(a, "--threads") && has) {
      cfg.threads = std::atoi(argv[++i]);
    } else if (!std::strcmp(a, "--truncation_eps") && has) {
      cfg.trunc_eps = std::atof(argv[++i]);
    } else if (!std::strcmp(a, "--queue_capacity") && has) {
      cfg.capacity = std::atoll(argv[++i]);
    } else if (!std::strcmp(a, "--time_budget") && has) {
      cfg.time_budget_s = std::atof(argv[++i]);
    } else if (!std::strcmp(a, "--sweep_period") && has) {
      cfg.sweep_period = std::atoi(argv[++i]);
    } else if (!std::strcmp(a, "--json")) {
      json = true;
    }
  }
  int src = std::atoi(argv[2]), dst = std::atoi(argv[3]);
  orc_graph* g = nullptr;
  int rc = orc_graph_load(argv[1], &g);
  if (rc) {
    std::fprintf(stderr, "cannot load %s (status %d)\n", argv[1], rc);
    return 1;
  }
  int n, e, nnz, he, type;
  orc_graph_dims(g, &n, &e, &nnz, &he, &type);
  if (he != e)
    std::fprintf(stderr, "Warning: header E=%d, found %d unique edges\n", he, e);
  orc_result r;
  rc = orc_solve(g, src, dst, &cfg, &r);
  orc_graph_free(g);
  if (rc) {
    std::fprintf(stderr, "solve failed (status %d)\n", rc);
    return 1;
  }
  if (json) {
    std::printf(
        "{\"R\": %.17g, \"ESP\": %.17g, \"D\": %.17g, \"paths\": %llu, "
        "\"nodes\": %llu, \"dead\": %llu, \"truncated\": %llu, "
        "\"overflowed\": %llu, \"folded\": %llu, \"iterations\": %lld, "
        "\"max_width\": %lld, \"elapsed_ms\": %.3f, \"deadline_hit\": %d, "
        "\"N\": %d, \"E\": %d}\n",
        r.R, r.ESP, r.D, r.paths, r.nodes, r.dead, r.truncated, r.overflowed,
        r.folded, r.iterations, r.max_width, r.elapsed_ms, r.deadline_hit, r.n,
        r.e);
    return 0;
  }
  std::printf("============================================\n");
  std::printf("  CPU Oracle - Two-Terminal Network Reliability\n");
  std::printf("============================================\n");
  std::printf("Graph:            N = %d, E = %d\n", r.n, r.e);
  std::printf("Terminals:        src = %d, dst = %d\n", src, dst);
  std::printf("--------------------------------------------\n");
  std::printf("Reliability:      R(G) = %.12f\n", r.R);
  std::printf("R + ESP:                 %.12f\n", r.R + r.ESP);
  std::printf("ESP (truncated):         %.6e\n", r.ESP);
  std::printf("Truncation bound:        %.6e\n", r.R > 0 ? r.ESP / r.R : r.ESP);
  std::printf("--------------------------------------------\n");
  std::printf("Paths enumerated: %llu\n", r.paths);
  std::printf("Nodes processed:  %llu\n", r.nodes);
  std::printf("Kernel iterations:%lld\n", r.iterations);
  std::printf("Max queue width:  %lld items\n", r.max_width);
  std::printf("Total time:       %.3f ms\n", r.elapsed_ms);
  std::printf("============================================\n");
  return 0;
}
