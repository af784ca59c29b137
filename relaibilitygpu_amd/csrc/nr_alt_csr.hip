/*
 * nr_alt_csr.hip - kernel instantiations for the alternating-sweep rule on
 * the CSR arcs (CsrAltExpander): undirected graphs of 65 to 256 nodes.
 */
#include "nr_level.h"

namespace nr {

template <int W, int NW>
struct KernAltCsr {
  using EX = CsrAltExpander<W, NW>;
  static void* level() { return (void*)&k_expand_level<W, EX, 0>; }
  static void* log() { return (void*)&k_expand_level<W, EX, 1>; }
  static void* learn() { return (void*)&k_expand_level<W, EX, 2>; }
  static void* mass() { return (void*)&k_expand_level<W, EX, 4>; }
  static void* small() { return (void*)&k_small_levels<W, EX>; }
  static void* ramp() { return (void*)&k_ramp_levels<W, EX>; }
  static void* items() { return (void*)&k_expand_items<W, EX>; }
};

template <int NW>
static Variant alt_csr_nw(int W) {
  Variant v{nullptr, nullptr, nullptr, nullptr, 0, NW, nullptr, nullptr, nullptr, nullptr, 0, nullptr};
  switch (W) {
#define NR_W(X)                                 \
  case X:                                       \
    v.level = &KernAltCsr<X, NW>::level;        \
    v.items = &KernAltCsr<X, NW>::items;        \
    v.level_log = &KernAltCsr<X, NW>::log;      \
    v.level_learn = &KernAltCsr<X, NW>::learn;  \
    v.level_mass = &KernAltCsr<X, NW>::mass;    \
    v.small = &KernAltCsr<X, NW>::small;        \
    v.ramp = &KernAltCsr<X, NW>::ramp;          \
    break;
    NR_W(1) NR_W(2) NR_W(3) NR_W(4) NR_W(5) NR_W(6) NR_W(7) NR_W(8)
#undef NR_W
    default: break;
  }
  return v;
}

Variant variant_alt_csr(int W, int NW) {
  switch (NW) {
    case 1:
    case 2: return alt_csr_nw<2>(W);
    case 3:
    case 4: return alt_csr_nw<4>(W);
    default: return Variant{nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr, 0, nullptr};
  }
}

}  // namespace nr
