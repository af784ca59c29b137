/*
 * nr_csr.hip - instantiations of the level and single-expansion kernels
 * with the CSR expander: graphs of up to 256 nodes and directed graphs
 * (items are edge-state planes and a mass; adjacency is read from the CSR
 * arcs).  Kept apart from nr_engine.hip so the two compile in parallel.
 */
#include "nr_level.h"

namespace nr {

template <int W, int NW>
struct KernCsr {
  static void* level() { return (void*)&k_expand_level<W, CsrExpander<W, NW>, 0>; }
  static void* items() { return (void*)&k_expand_items<W, CsrExpander<W, NW>>; }
  static void* small() { return (void*)&k_small_levels<W, CsrExpander<W, NW>>; }
  static void* ramp() { return (void*)&k_ramp_levels<W, CsrExpander<W, NW>>; }
};

template <int NW>
static Variant variant_csr_nw(int W) {
  switch (W) {
#define NR_W(X) \
  case X:      \
    {                                                                                     \
      Variant v{&KernCsr<X, NW>::level, &KernCsr<X, NW>::items, nullptr, nullptr, 0, NW,  \
                nullptr, nullptr, nullptr, nullptr, 0, nullptr};                          \
      v.small = &KernCsr<X, NW>::small;                                                   \
      v.ramp = &KernCsr<X, NW>::ramp;                                                     \
      return v;                                                                           \
    }
    NR_W(1) NR_W(2) NR_W(3) NR_W(4) NR_W(5) NR_W(6) NR_W(7) NR_W(8)
#undef NR_W
    default: return Variant{nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr, 0, nullptr};
  }
}

Variant variant_csr(int W, int NW) {
  switch (NW) {
    case 1: return variant_csr_nw<1>(W);
    case 2: return variant_csr_nw<2>(W);
    case 3:
    case 4: return variant_csr_nw<4>(W);
    default: return Variant{nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr};
  }
}

}  // namespace nr
