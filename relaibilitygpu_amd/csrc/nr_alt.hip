/*
 * nr_alt.hip - kernel instantiations for the alternating-sweep rule
 * (AltExpander: fewest-UNKNOWN path, sweep direction flipping with the
 * FAILED edge count) on undirected graphs of up to 64 nodes; items are
 * edge-state planes and a mass only.
 */
#include "nr_level.h"

namespace nr {

template <int W, typename NS>
struct KernAlt {
  using EX = AltExpander<W, NS>;
  static void* level() { return (void*)&k_expand_level<W, EX, 0>; }
  static void* log() { return (void*)&k_expand_level<W, EX, 1>; }
  static void* learn() { return (void*)&k_expand_level<W, EX, 2>; }
  static void* mass() { return (void*)&k_expand_level<W, EX, 4>; }
  static void* small() { return (void*)&k_small_levels<W, EX>; }
  static void* ramp() { return (void*)&k_ramp_levels<W, EX>; }
  static void* items() { return (void*)&k_expand_items<W, EX>; }
};

template <typename NS>
static Variant alt_for_w(int W) {
  Variant v{nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr, 1, nullptr};
  switch (W) {
#define NR_W(X)                              \
  case X:                                    \
    v.level = &KernAlt<X, NS>::level;        \
    v.items = &KernAlt<X, NS>::items;        \
    v.level_log = &KernAlt<X, NS>::log;      \
    v.level_learn = &KernAlt<X, NS>::learn;  \
    v.level_mass = &KernAlt<X, NS>::mass;    \
    v.small = &KernAlt<X, NS>::small;        \
    v.ramp = &KernAlt<X, NS>::ramp;          \
    break;
    NR_W(1) NR_W(2) NR_W(3) NR_W(4) NR_W(5) NR_W(6) NR_W(7) NR_W(8)
#undef NR_W
    default: break;
  }
  v.ns_bytes = (int)sizeof(NS); /* node-set width (items carry no rows) */
  return v;
}

Variant variant_alt(int W, int ns_bytes) {
  return ns_bytes == 4 ? alt_for_w<u32>(W) : alt_for_w<u64>(W);
}

}  // namespace nr
