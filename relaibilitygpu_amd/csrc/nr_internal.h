/*
 * nr_internal.h - host-side types shared by the loader, the HIP runtime and
 * the command line.  Not part of the C-ABI.
 */
#ifndef NR_INTERNAL_H
#define NR_INTERNAL_H

#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nr_api.h"

/* Host graph: CSR in file order plus per-edge FP64 probabilities.
 * Mirrors the fields of GraphHost (workingPad/graph.h:13-27) with FP64
 * p/q instead of float logs. */
struct nr_graph {
  int n = 0, e = 0, nnz = 0;
  int header_e = 0, type = 0;
  int src = 0, dst = 0;
  std::vector<int> row_ptr, col_idx, edge_id;
  std::vector<double> p, q;
  /* derived */
  std::vector<int> ea, eb;      /* endpoints of each edge id, ea <= eb */
  std::vector<uint8_t> dirmask; /* bit0: ea->eb in CSR, bit1: eb->ea */
  bool symmetric = true;
  bool simple = true; /* no two edge ids join the same pair of nodes */
  int precontracted = 0, predeleted = 0;
  std::vector<int> terms; /* K-terminal set (empty: the src/dst pair) */
  /* unique per state: new at creation and at every change of terminals
   * (keys the device copy of a batch's instance tables) */
  unsigned long long stamp = 0;
};

namespace nr {

int load_graph_text(const char* text, size_t len, nr_graph* g);
int finish_graph(nr_graph* g);
inline int mask_words(const nr_graph* g) { return g->e > 0 ? (g->e + 63) / 64 : 1; }
unsigned long long next_stamp();

/* Runs the body of a C-ABI entry point so that no C++ exception crosses the
 * boundary: a failed host allocation (std::bad_alloc, or std::length_error
 * for an impossible size) returns NR_ERR_NOMEM, anything else NR_ERR_ARG. */
template <class F>
int guarded(F&& body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return NR_ERR_NOMEM;
  } catch (const std::length_error&) {
    return NR_ERR_NOMEM;
  } catch (...) {
    return NR_ERR_ARG;
  }
}

}  // namespace nr

#endif
