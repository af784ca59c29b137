/*
 * nr_learn_csr.hip - learned-order (MODE 2) instantiations of the level
 * kernel with the CSR expander (graphs of up to 256 nodes, directed graphs).
 */
#include "nr_level.h"

namespace nr {

template <int W, int NW>
struct KernLearnCsr {
  static void* learn() { return (void*)&k_expand_level<W, CsrExpander<W, NW>, 2>; }
};

template <int NW>
static void learn_csr_nw(Variant& v, int W) {
  switch (W) {
#define NR_W(X)                                   \
  case X:                                         \
    v.level_learn = &KernLearnCsr<X, NW>::learn;  \
    break;
    NR_W(1) NR_W(2) NR_W(3) NR_W(4) NR_W(5) NR_W(6) NR_W(7) NR_W(8)
#undef NR_W
    default: break;
  }
}

void add_learn_kernels_csr(Variant& v, int W) {
  switch (v.nw) {
    case 1: learn_csr_nw<1>(v, W); break;
    case 2: learn_csr_nw<2>(v, W); break;
    case 4: learn_csr_nw<4>(v, W); break;
    default: break;
  }
}

}  // namespace nr
