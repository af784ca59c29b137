/*
 * nr_big.h - the wave-per-item form for graphs past the per-lane forms
 * (nr_big.hip): LDS shape of one wave's slice and the kernel getters.
 */
#ifndef NR_BIG_H
#define NR_BIG_H

#include "nr_level.h"

namespace nr {

/* One wave's LDS slice: the item's planes (known[W], up[W]), the current
 * and next frontier bitsets (NW words each), node levels (bytes, 0xff =
 * unreached) and the hop stack (hop_cap u16).  Slices follow the
 * workgroup's shared region at off_waves. */
struct BigShape {
  int w, hop_cap;
  int off_planes, off_cur, off_nxt, off_lev, off_hops, wave_bytes;
  int off_waves;
};

inline BigShape big_shape(int n, int w, int off_waves) {
  BigShape b;
  b.w = w;
  b.hop_cap = 256; /* levels are bytes, so a path has at most 254 hops */
  int o = 0;
  b.off_planes = o;
  o += 16 * w;
  const int nw = (n + 63) / 64;
  b.off_cur = o;
  o += 8 * nw;
  b.off_nxt = o;
  o += 8 * nw;
  b.off_lev = o;
  o += (n + 7) / 8 * 8;
  b.off_hops = o;
  o += 2 * b.hop_cap;
  b.wave_bytes = (o + 15) / 16 * 16;
  b.off_waves = (off_waves + 15) / 16 * 16;
  return b;
}

void* big_expand_kernel(); /* k_expand_big(LevelArgs, int parity, BigShape) */
void* big_init_kernel();   /* k_init_big(LevelArgs, const u64* root, int W, u64 n_slots) */

}  // namespace nr

#endif
