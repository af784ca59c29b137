/*
 * nr_choose.hip - kernel instantiations of the learned path choice
 * (ChooseAltExpander: the alternating rule's fewest-UNKNOWN path, its sweep
 * direction picked per item by the chooser network on the matrix cores) on
 * undirected graphs of up to 64 nodes: level order, tree log (exploring
 * directions), learned order and max-mass order.
 */
#include "nr_level.h"

namespace nr {

template <int W, typename NS>
struct KernChoose {
  using EX = ChooseAltExpander<W, NS>;
  static void* level() { return (void*)&k_expand_level<W, EX, 0>; }
  static void* log() { return (void*)&k_expand_level<W, EX, 1>; }
  static void* learn() { return (void*)&k_expand_level<W, EX, 2>; }
  static void* mass() { return (void*)&k_expand_level<W, EX, 4>; }
};

template <typename NS>
static Variant choose_for_w(int W) {
  Variant v{nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr, 1, nullptr};
  switch (W) {
#define NR_W(X)                                  \
  case X:                                        \
    v.level = &KernChoose<X, NS>::level;         \
    v.level_log = &KernChoose<X, NS>::log;       \
    v.level_learn = &KernChoose<X, NS>::learn;   \
    v.level_mass = &KernChoose<X, NS>::mass;     \
    break;
    NR_W(1) NR_W(2) NR_W(3) NR_W(4) NR_W(5) NR_W(6) NR_W(7) NR_W(8)
#undef NR_W
    default: break;
  }
  v.ns_bytes = (int)sizeof(NS); /* node-set width (items carry no rows) */
  v.choose = 1;
  return v;
}

Variant variant_choose(int W, int ns_bytes) {
  return ns_bytes == 4 ? choose_for_w<u32>(W) : choose_for_w<u64>(W);
}

}  // namespace nr
