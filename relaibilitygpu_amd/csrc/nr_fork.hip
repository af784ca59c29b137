/*
 * nr_fork.hip - forkBest-rule (MODE 3) instantiations of the level kernel,
 * row form and CSR form (workingPad/main.cu:304-430, 550-567: a confirmed
 * WORKING path is a success, otherwise a two-way branch on the path's most
 * reliable UNKNOWN edge).
 */
#include "nr_level.h"

namespace nr {

template <int W, class EX>
struct KernFork {
  static void* fork() { return (void*)&k_expand_level<W, EX, 3>; }
};

template <class Make>
static KernelGetter fork_for_w(int W) {
  switch (W) {
#define NR_W(X) \
  case X: return &KernFork<X, typename Make::template Ex<X>>::fork;
    NR_W(1) NR_W(2) NR_W(3) NR_W(4) NR_W(5) NR_W(6) NR_W(7) NR_W(8)
#undef NR_W
    default: return nullptr;
  }
}

template <typename NS>
struct RowsOf {
  template <int W>
  using Ex = Expander<W, NS>;
};
template <int NW>
struct CsrOf {
  template <int W>
  using Ex = CsrExpander<W, NW>;
};

void add_fork_kernels(Variant& v, int W) {
  if (v.nw == 1)
    v.level_fork = fork_for_w<CsrOf<1>>(W);
  else if (v.nw == 2)
    v.level_fork = fork_for_w<CsrOf<2>>(W);
  else if (v.nw == 4)
    v.level_fork = fork_for_w<CsrOf<4>>(W);
  else if (v.ns_bytes == 4)
    v.level_fork = fork_for_w<RowsOf<u32>>(W);
  else
    v.level_fork = fork_for_w<RowsOf<u64>>(W);
}

}  // namespace nr
