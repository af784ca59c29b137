/*
 * nr_mass.hip - max-mass class-round (MODE 4) instantiations of the level
 * kernel for every expander form: the class-round code is compiled only
 * here, so the level-order kernels carry none of it.
 */
#include "nr_level.h"

namespace nr {

template <int W, class EX>
struct KernMass {
  static void* mass() { return (void*)&k_expand_level<W, EX, 4>; }
};

template <template <int> class Ex>
static KernelGetter mass_for_w(int W) {
  switch (W) {
#define NR_W(X) \
  case X: return &KernMass<X, Ex<X>>::mass;
    NR_W(1) NR_W(2) NR_W(3) NR_W(4) NR_W(5) NR_W(6) NR_W(7) NR_W(8)
#undef NR_W
    default: return nullptr;
  }
}

template <int W> using Rows32 = Expander<W, u32>;
template <int W> using Rows64 = Expander<W, u64>;
template <int W> using Lean32 = LeanExpander<W, u32>;
template <int W> using Lean64 = LeanExpander<W, u64>;
template <int W> using Csr1 = CsrExpander<W, 1>;
template <int W> using Csr2 = CsrExpander<W, 2>;
template <int W> using Csr4 = CsrExpander<W, 4>;

template <int W, typename RowNS, bool kTag>
struct KernSplit {
  static void* split() { return (void*)&k_class_split<W, RowNS, kTag>; }
};

template <typename RowNS, bool kTag>
static KernelGetter split_for_w(int W) {
  switch (W) {
#define NR_W(X) \
  case X: return &KernSplit<X, RowNS, kTag>::split;
    NR_W(1) NR_W(2) NR_W(3) NR_W(4) NR_W(5) NR_W(6) NR_W(7) NR_W(8)
#undef NR_W
    default: return nullptr;
  }
}

KernelGetter split_kernel(int W, int row_bytes, bool tag_class) {
  if (tag_class)
    return row_bytes == 4   ? split_for_w<u32, true>(W)
           : row_bytes == 8 ? split_for_w<u64, true>(W)
                            : split_for_w<void, true>(W);
  return row_bytes == 4   ? split_for_w<u32, false>(W)
         : row_bytes == 8 ? split_for_w<u64, false>(W)
                          : split_for_w<void, false>(W);
}

void add_mass_kernels(Variant& v, int W) {
  if (v.nw == 1)
    v.level_mass = mass_for_w<Csr1>(W);
  else if (v.nw == 2)
    v.level_mass = mass_for_w<Csr2>(W);
  else if (v.nw == 4)
    v.level_mass = mass_for_w<Csr4>(W);
  else if (v.lean)
    v.level_mass = v.ns_bytes == 4 ? mass_for_w<Lean32>(W) : mass_for_w<Lean64>(W);
  else
    v.level_mass = v.ns_bytes == 4 ? mass_for_w<Rows32>(W) : mass_for_w<Rows64>(W);
}

}  // namespace nr
