/*
 * nr_choose_csr.hip - kernel instantiations of the learned path choice on
 * the CSR arcs (CsrChooseAltExpander): undirected graphs of 65 to 256 nodes.
 */
#include "nr_level.h"

namespace nr {

template <int W, int NW>
struct KernChooseCsr {
  using EX = CsrChooseAltExpander<W, NW>;
  static void* level() { return (void*)&k_expand_level<W, EX, 0>; }
  static void* log() { return (void*)&k_expand_level<W, EX, 1>; }
  static void* learn() { return (void*)&k_expand_level<W, EX, 2>; }
  static void* mass() { return (void*)&k_expand_level<W, EX, 4>; }
};

template <int NW>
static Variant choose_csr_nw(int W) {
  Variant v{nullptr, nullptr, nullptr, nullptr, 0, NW, nullptr, nullptr, nullptr, nullptr, 0, nullptr};
  switch (W) {
#define NR_W(X)                                     \
  case X:                                           \
    v.level = &KernChooseCsr<X, NW>::level;         \
    v.level_log = &KernChooseCsr<X, NW>::log;       \
    v.level_learn = &KernChooseCsr<X, NW>::learn;   \
    v.level_mass = &KernChooseCsr<X, NW>::mass;     \
    break;
    NR_W(1) NR_W(2) NR_W(3) NR_W(4) NR_W(5) NR_W(6) NR_W(7) NR_W(8)
#undef NR_W
    default: break;
  }
  v.choose = 1;
  return v;
}

Variant variant_choose_csr(int W, int NW) {
  switch (NW) {
    case 1:
    case 2: return choose_csr_nw<2>(W);
    case 3:
    case 4: return choose_csr_nw<4>(W);
    default: return Variant{nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr, 0, nullptr};
  }
}

}  // namespace nr
