/*
 * nr_lean.hip - level-kernel instantiations with the lean row-form expander
 * (undirected graphs of up to 64 nodes, plain path rule, two terminals):
 * items are edge-state planes and a mass only (24 B for one plane word).
 */
#include "nr_level.h"

namespace nr {

template <int W, typename NS>
struct KernLean {
  static void* level() { return (void*)&k_expand_level<W, LeanExpander<W, NS>, 0>; }
  static void* small() { return (void*)&k_small_levels<W, LeanExpander<W, NS>>; }
  static void* ramp() { return (void*)&k_ramp_levels<W, LeanExpander<W, NS>>; }
};

template <typename NS>
static Variant lean_for_w(int W) {
  Variant v{nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr, 1, nullptr};
  switch (W) {
#define NR_W(X)                            \
  case X:                                  \
    v.level = &KernLean<X, NS>::level;     \
    v.small = &KernLean<X, NS>::small;     \
    v.ramp = &KernLean<X, NS>::ramp;       \
    break;
    NR_W(1) NR_W(2) NR_W(3) NR_W(4) NR_W(5) NR_W(6) NR_W(7) NR_W(8)
#undef NR_W
    default: break;
  }
  v.ns_bytes = (int)sizeof(NS); /* node-set width (items carry no rows) */
  return v;
}

Variant variant_lean(int W, int ns_bytes) {
  return ns_bytes == 4 ? lean_for_w<u32>(W) : lean_for_w<u64>(W);
}

}  // namespace nr
