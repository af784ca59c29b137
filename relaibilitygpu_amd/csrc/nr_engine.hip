/*
 * nr_engine.hip - HIP kernels (gfx950, wave64) and host runtime of the
 * two-terminal reliability engine; implements nr_solve / nr_expand_items of
 * include/nr_api.h.
 *
 * Level order (the reference's order, workingPad/main.cu:963-1008).  The
 * queue is a pair of HBM buffers in structure-of-arrays form (known planes,
 * up planes, allowed-adjacency rows, FP64 mass), each cut into one segment
 * per workgroup of the expand grid.  One launch of k_expand_level consumes
 * one tree level: every workgroup first scans the previous launch's
 * per-segment counts into a prefix table in LDS; lane i of a wave then takes
 * the level's item i, runs the BFS/path search of nr_device.h over the
 * item's rows, and the wave reserves room for all of its children in its
 * workgroup's own output segment with one LDS atomic; every lane then writes
 * its children at its offset.  No global atomics and one launch per level,
 * so consecutive levels run back to back with no host round trip; the host
 * reads the control block once per chunk of levels.
 *
 * Masses are FP64 products taken left to right along the path; R, D and ESP
 * are per-lane Neumaier sums, merged per wave into double-double slots, and
 * the counters go to per-wave slots; k_finalize reduces both in a fixed
 * order (replaces the single-address FP64 CAS of main.cu:68-79 and the
 * global counters of main.cu:58-63).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <vector>

#include "nr_device.h"
#include "nr_internal.h"
#include "nr_big.h"
#include "nr_level.h"

namespace nr {

/* Sum the per-segment class histograms into the control block: one
 * workgroup per class, a strided sum over the segments and a tree in LDS
 * (fixed order, so the totals do not depend on timing). */
__global__ void __launch_bounds__(256) k_hist_total(LevelArgs a) {
  __shared__ u64 part[256];
  const int cl = blockIdx.x;
  u64 sum = 0;
  for (int b = threadIdx.x; b < a.n_seg; b += blockDim.x) sum += a.hist[(u64)b * kClasses + cl];
  part[threadIdx.x] = sum;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.ctl->hist[cl] = part[0];
}

/* Fold every queued item of buf[parity] into ESP (time budget expiry;
 * original/Seq_16.cpp:44,227 PURGEQ restated on the device). */
__global__ void __launch_bounds__(256) k_fold(LevelArgs a, int parity) {
  extern __shared__ __align__(16) unsigned char smem[];
  u32* sp = (u32*)smem;
  u32* part = sp + a.n_seg + 2;
  block_prefix(a.seg_count[parity], a.n_seg, sp, part);
  const u32 n_items = sp[a.n_seg];
  const int lane = threadIdx.x & 63;
  const u64 wave_id = (u64)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const u64 n_waves = (u64)(gridDim.x * blockDim.x) >> 6;
  Acc2 e;
  e.init();
  u32 c[C_N] = {0, 0, 0, 0, 0, 0, 0, 0};
  const u64 per_wave = (n_items + n_waves - 1) / n_waves;
  const u64 chunk = per_wave >= 64 ? 64 : per_wave;
  for (u64 base = wave_id * chunk; base < n_items; base += n_waves * chunk) {
    const u32 i = (u32)(base + lane);
    if ((u64)lane < chunk && i < n_items) {
      int s = segment_of(sp, a.n_seg, (u32)base);
      while (sp[s + 1] <= i) ++s;
      e.add(a.buf[parity].mass[(u64)s * a.seg_cap + (i - sp[s])]);
      ++c[C_FOLD];
    }
  }
  Acc2 zr, zd;
  zr.init();
  zd.init();
  __syncthreads();
  block_flush(a, zr, e, zd, c, (unsigned char*)part);
}

/* State a resumed solve starts from (device copy). */
struct SeedMeta {
  DD sum[3];       /* R, ESP, D so far */
  u64 counts[C_N];
  u64 iterations, max_width;
  u64 count;       /* pending items */
};

/* Seed a solve: zero the control block, the partial-sum and counter slots,
 * set the first prefix table to one item, and put the root item
 * (main.cu:853-855) in slot 0 of buffer 0.  A kernel instead of memset /
 * memcpy calls keeps the whole solve on one stream. */
struct Root {
  u64 known[8], up[8];
  u64 rows[64];
};

__global__ void __launch_bounds__(256) k_init(LevelArgs a, Root root, int W, int n,
                                              int ns_bytes, u64 n_slots, const SeedMeta* seed) {
  const u64 tid = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 i = tid; i < n_slots; i += stride) {
    a.partR[i].hi = a.partR[i].lo = 0.0;
    a.partE[i].hi = a.partE[i].lo = 0.0;
    a.partD[i].hi = a.partD[i].lo = 0.0;
    for (int j = 0; j < C_N; ++j) a.cnt[i * C_N + j] = 0;
  }
  if (blockIdx.x == 0) { /* the control block, zeroed by the whole workgroup */
    static_assert(sizeof(LevelCtl) % 8 == 0, "LevelCtl is a whole number of words");
    for (u32 i = threadIdx.x; i < sizeof(LevelCtl) / 8; i += blockDim.x) ((u64*)a.ctl)[i] = 0;
    __syncthreads();
  }
  const u64 ns = (u64)a.n_seg;
  for (u64 s = tid; s < ns; s += stride) {
    a.seg_count[0][s] =
        seed ? (u32)(seed->count / ns + (s < seed->count % ns ? 1 : 0)) : (s == 0 ? 1u : 0u);
    a.seg_count[1][s] = 0u;
    if (a.seg_settle[0]) {
      a.seg_settle[0][s] = 0u;
      a.seg_settle[1][s] = 0u;
    }
  }
  if (seed) { /* a resumed solve: sums and counters so far in slot 0 */
    if (tid == 0) {
      a.ctl->start_tick = wall_clock64();
      a.ctl->last_items = seed->count;
      a.ctl->iterations = seed->iterations;
      a.ctl->max_width = seed->max_width;
      a.partR[0] = seed->sum[0];
      a.partE[0] = seed->sum[1];
      a.partD[0] = seed->sum[2];
      for (int j = 0; j < C_N; ++j) a.cnt[j] = seed->counts[j];
    }
    return;
  }
  if (tid == 0) {
    a.ctl->start_tick = wall_clock64();
    a.ctl->last_items = 1;
    for (int w = 0; w < W; ++w) {
      a.buf[0].known[(u64)w * a.stride] = root.known[w];
      a.buf[0].up[(u64)w * a.stride] = root.up[w];
    }
    for (int u = 0; u < n && ns_bytes; ++u) {
      if (ns_bytes == 4 && n <= 16) /* two 16-bit rows per word (Expander::packed) */
        ((u32*)a.buf[0].rows)[(u64)(u >> 1) * a.stride] =
            (u & 1) ? ((u32*)a.buf[0].rows)[(u64)(u >> 1) * a.stride] | ((u32)root.rows[u] << 16)
                    : (u32)root.rows[u];
      else if (ns_bytes == 4)
        ((u32*)a.buf[0].rows)[(u64)u * a.stride] = (u32)root.rows[u];
      else
        ((u64*)a.buf[0].rows)[(u64)u * a.stride] = root.rows[u];
    }
    a.buf[0].mass[0] = 1.0;
    if (a.tag[0]) a.tag[0][0] = a.mlp ? 0u : 0xffffffffu; /* class 0 / no parent record */
  }
}

/* The solve's totals (finalize_slots, min(kFinMax, ceil(n_slots / 64))
 * workgroups), when no empty level launch has computed them. */
__global__ void __launch_bounds__(256) k_finalize(LevelArgs a, u64 n_slots) {
  __shared__ u32 s_last;
  finalize_slots(a, n_slots, &s_last);
}

inline unsigned finalize_groups(u64 n_slots) {
  return (unsigned)std::min<u64>(kFinMax, (n_slots + 63) / 64);
}

/* ---- checkpoints: pending items out of, and back into, the queue ---- */


/* Compact copy of the items of buf[parity] (planes, mass, tag). */
__global__ void __launch_bounds__(256) k_gather(LevelArgs a, int parity, int W, u64* gk,
                                                u64* gu, double* gm, u32* gt) {
  extern __shared__ __align__(16) unsigned char smem[];
  u32* sp = (u32*)smem;
  u32* part = sp + a.n_seg + 2;
  block_prefix(a.seg_count[parity], a.n_seg, sp, part);
  const u32 n_items = sp[a.n_seg];
  const QBuf b = a.buf[parity];
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n_items;
       i += (u64)gridDim.x * blockDim.x) {
    int s = segment_of(sp, a.n_seg, (u32)i);
    const u64 slot = (u64)s * a.seg_cap + (i - sp[s]);
    for (int w = 0; w < W; ++w) {
      gk[i * W + w] = b.known[(u64)w * a.stride + slot];
      gu[i * W + w] = b.up[(u64)w * a.stride + slot];
    }
    gm[i] = b.mass[slot];
    if (gt) gt[i] = a.tag[parity] ? a.tag[parity][slot] : 0u;
  }
}

/* Seed item i of a resumed solve into buffer 0: segment i % n_seg, slot
 * i / n_seg (the counts k_init gives every segment). */
__device__ inline u64 seed_slot(const LevelArgs& a, u64 i) {
  return (i % (u64)a.n_seg) * a.seg_cap + i / (u64)a.n_seg;
}

__global__ void __launch_bounds__(256) k_seed_copy(LevelArgs a, int W, const u64* sk,
                                                   const u64* su, const double* sm,
                                                   const u32* st, u64 count) {
  const QBuf b = a.buf[0];
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (u64)gridDim.x * blockDim.x) {
    const u64 slot = seed_slot(a, i);
    for (int w = 0; w < W; ++w) {
      b.known[(u64)w * a.stride + slot] = sk[i * W + w];
      b.up[(u64)w * a.stride + slot] = su[i * W + w];
    }
    b.mass[slot] = sm[i];
    if (a.tag[0]) a.tag[0][slot] = st ? st[i] : 0u;
  }
}

/* Row form: each seeded item's adjacency rows, derived from its planes. */
template <int W, typename NS>
__global__ void __launch_bounds__(256) k_seed_rows(LevelArgs a, u64 count) {
  extern __shared__ __align__(16) unsigned char smem[];
  stage_tables(a.g, a.lay, smem);
  __syncthreads();
  Expander<W, NS> ex;
  ex.bind_graph(smem, a.lay, a.g, threadIdx.x);
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const u64 slot = seed_slot(a, i);
  u64 known[W], up[W];
  double mass;
  load_item<W>(a.buf[0], a.stride, slot, known, up, mass);
  ex.derive(known, up);
  ex.store_rows(a.buf[0], a.stride, slot);
}

/* ------------------------------------------------------------------ */
/* Batched instances: many graphs / terminal pairs in one level loop   */
/* ------------------------------------------------------------------ */

/* Per-instance exact accumulator: value * 2^96 as a 128-bit integer in two
 * u64 words (lo, hi), added with two atomics and an explicit carry, so the
 * total is exact (to 2^-96 per term) and independent of the order in which
 * lanes add.  Masses are in [0, 1]. */
__device__ inline void fx_add(u64* acc, double v) {
  const double s = v * 4294967296.0; /* v * 2^32 */
  const u64 hi = (u64)s;
  const double frac = s - (double)hi;
  const u64 lo = (u64)(frac * 18446744073709551616.0);
  if (lo) {
    const u64 old = atomicAdd((unsigned long long*)&acc[0], (unsigned long long)lo);
    if (old + lo < old) atomicAdd((unsigned long long*)&acc[1], 1ull);
  }
  if (hi) atomicAdd((unsigned long long*)&acc[1], (unsigned long long)hi);
}

/* The same fixed-point value split into three 32-bit limbs held in u64
 * lanes (value * 2^96 = l2 * 2^64 + l1 * 2^32 + l0), so a wave can add up to
 * 2^25 terms with plain integer sums before one flush. */
__device__ inline void fx_limbs_add(u64& l2, u64& l1, u64& l0, double v) {
  const double s = v * 4294967296.0;
  const u64 hi = (u64)s;
  const u64 lo = (u64)((s - (double)hi) * 18446744073709551616.0);
  l2 += hi;
  l1 += lo >> 32;
  l0 += lo & 0xffffffffull;
}

__device__ inline void fx_flush(u64* acc, u64 l2, u64 l1, u64 l0) {
  const u64 lo = l0 + ((l1 & 0xffffffffull) << 32);
  u64 hi = l2 + (l1 >> 32) + (lo < l0 ? 1ull : 0ull);
  if (lo) {
    const u64 old = atomicAdd((unsigned long long*)&acc[0], (unsigned long long)lo);
    if (old + lo < old) hi += 1;
  }
  if (hi) atomicAdd((unsigned long long*)&acc[1], (unsigned long long)hi);
}

/* Per-instance description (device). */
struct InstDesc {
  int n, e, src, dst;
  /* byte offsets of this instance's tables from BatchArgs::tab */
  u32 off_p, off_q, off_sadj, off_inc, off_ep, off_deg, off_eidm, pad;
};

/* Per-instance results (device): fixed-point sums and counters. */
enum { B_R = 0, B_E = 2, B_D = 4, B_NODES = 6, B_PATHS, B_DEAD, B_TRUNC, B_OVER, B_FOLD,
       B_LEVELS, B_ERR, B_N = 16 };

/* Item tag in a batch: instance id (low 22 bits) and tree depth. */
constexpr u32 kTagIdBits = 22;
constexpr u32 kTagIdMask = (1u << kTagIdBits) - 1u;
constexpr u32 kMaxBatch = 1u << kTagIdBits;
constexpr u32 kMaxDepth = (1u << (32 - kTagIdBits)) - 1u;

/* Stream state, double-buffered by the parity of the buffer a refill writes:
 * refill(out = b) reads top[b] / next[b] and writes top[b ^ 1] / next[b ^ 1];
 * the expand launch reading buffer b pushes its spill on top[b ^ 1]. */
struct StreamCtl {
  u64 top[2];  /* spill stack height */
  u32 next[2]; /* instances admitted so far (index into perm) */
};

struct BatchArgs {
  Layout lay;        /* per-lane LDS slices sized for the largest instance */
  QBuf buf[2];
  u32* tag[2];       /* instance id | depth << kTagIdBits, per slot */
  u32* seg_count[2]; /* [n_seg]: segments 0..grid-1 (seg_cap slots each)
                        belong to workgroups; segment grid (fill_step slots,
                        the last) is the refill segment */
  QBuf spill;        /* LIFO of children that did not fit their segment */
  u32* spill_tag;
  u64 spill_cap;
  const InstDesc* desc;
  const unsigned char* tab; /* every instance's tables (graph-table layout) */
  const u64* roots;  /* [count][2W]: known[W] | up[W] */
  const u32* perm;   /* admission order */
  StreamCtl* sc;
  u64* res;          /* [count][B_N] */
  LevelCtl* ctl;
  u64 stride;
  u32 seg_cap;
  int n_seg;         /* grid + 1 */
  u32 count;         /* instances in perm */
  u32 fill_step;     /* at most this many items refilled per level */
  u32 fill_to;       /* refill while the next level holds fewer items */
  u32 root_step;     /* at most this many new roots per refill */
  int roots_when_empty; /* admit roots only into an empty queue and stack */
  double eps;
};

template <int W>
__device__ inline void copy_item(const QBuf& from, u64 fstride, u64 fi, const QBuf& to,
                                 u64 tstride, u64 ti) {
#pragma unroll
  for (int w = 0; w < W; ++w) {
    to.known[(u64)w * tstride + ti] = from.known[(u64)w * fstride + fi];
    to.up[(u64)w * tstride + ti] = from.up[(u64)w * fstride + fi];
  }
  to.mass[ti] = from.mass[fi];
}

/* Between two expand launches: top up the level about to run (buffer
 * `out`) from the spill stack, deepest items first, then with new instance
 * roots, through the refill segment.  Every workgroup derives the same
 * counts from the same inputs; workgroup 0 publishes the new state. */
template <int W>
__global__ void __launch_bounds__(256) k_refill_batch(BatchArgs a, int out) {
  __shared__ u32 s_part[256];
  const int grid = a.n_seg - 1;
  u32 sum = 0;
  for (int b = threadIdx.x; b < grid; b += blockDim.x) sum += a.seg_count[out][b];
  s_part[threadIdx.x] = sum;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) s_part[threadIdx.x] += s_part[threadIdx.x + off];
    __syncthreads();
  }
  const u64 n_next = s_part[0];
  /* pushes past spill_cap were settled as overflow and never stored */
  const u64 top = a.sc->top[out] < a.spill_cap ? a.sc->top[out] : a.spill_cap;
  const u32 next = a.sc->next[out];
  u64 room = n_next < a.fill_to ? a.fill_to - n_next : 0;
  if (room > a.fill_step) room = a.fill_step;
  const u64 n_pop = top < room ? top : room;
  u64 n_new = a.count - next;
  if (n_new > room - n_pop) n_new = room - n_pop;
  if (n_new > a.root_step) n_new = a.root_step;
  if (a.roots_when_empty && (n_next || top)) n_new = 0;
  const u64 seg0 = (u64)grid * a.seg_cap;
  const QBuf dst = a.buf[out];
  for (u64 j = (u64)blockIdx.x * blockDim.x + threadIdx.x; j < n_pop + n_new;
       j += (u64)gridDim.x * blockDim.x) {
    const u64 slot = seg0 + j;
    if (j < n_pop) {
      const u64 si = top - n_pop + j;
      copy_item<W>(a.spill, a.spill_cap, si, dst, a.stride, slot);
      a.tag[out][slot] = a.spill_tag[si];
    } else {
      const u32 id = a.perm[next + (u32)(j - n_pop)];
      const u64* r = a.roots + (u64)id * (2 * W);
#pragma unroll
      for (int w = 0; w < W; ++w) {
        dst.known[(u64)w * a.stride + slot] = r[w];
        dst.up[(u64)w * a.stride + slot] = r[W + w];
      }
      dst.mass[slot] = 1.0;
      a.tag[out][slot] = id; /* depth 0 */
      u64* res = a.res + (u64)id * B_N;
      for (int k = 0; k < B_N; ++k) res[k] = 0;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.sc->top[out ^ 1] = top - n_pop;
    a.sc->next[out ^ 1] = next + (u32)n_new;
    a.seg_count[out][grid] = (u32)(n_pop + n_new);
    const u64 total = n_next + n_pop + n_new;
    a.ctl->last_items = total;
    if (total) {
      a.ctl->iterations += 1;
      if (total > a.ctl->max_width) a.ctl->max_width = total;
    }
  }
}

template <int W, typename NS>
__global__ void __launch_bounds__(256) k_expand_batch(BatchArgs a, int parity) {
  extern __shared__ __align__(16) unsigned char smem[];
  u32* sp = (u32*)(smem + a.lay.off_extra);
  u32* s_cnt = sp + a.n_seg + 1;
  u32* part = s_cnt + 1;
  block_prefix(a.seg_count[parity], a.n_seg, sp, part);
  const u32 n_items = sp[a.n_seg];
  if (n_items == 0) {
    if (threadIdx.x == 0) a.seg_count[parity ^ 1][blockIdx.x] = 0;
    return;
  }
  if (threadIdx.x == 0) *s_cnt = 0;
  __syncthreads();

  Expander<W, NS> ex;
  ex.rows = nullptr;
  ex.lvl = nullptr;
  ex.uh = (u16*)(smem + a.lay.off_uh) + threadIdx.x;
  ex.stride = a.lay.block;
  unsigned char* lvb = smem + a.lay.off_lvl + 4 * threadIdx.x;

  const QBuf in = a.buf[parity];
  const QBuf out = a.buf[parity ^ 1];
  const u32* tin = a.tag[parity];
  u32* tout = a.tag[parity ^ 1];
  u64* spill_top = &a.sc->top[parity ^ 1];
  const int lane = threadIdx.x & 63;
  const u64 wave_id = (u64)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const u64 n_waves = (u64)(gridDim.x * blockDim.x) >> 6;
  const u64 out_base = (u64)blockIdx.x * a.seg_cap;
  const u64 per_wave = (n_items + n_waves - 1) / n_waves;
  const u64 chunk = per_wave >= 64 ? 64 : per_wave;

  for (u64 base = wave_id * chunk; base < n_items; base += n_waves * chunk) {
    const u32 i = (u32)(base + lane);
    u64 known[W], up[W];
    double mass = 0.0;
    int k = 0;
    bool live = false, active = false;
    u32 id = 0, depth = 0;
    /* this lane's contributions, flushed per instance once per wave pass:
     * counts packed nodes | truncated << 16 | dead << 32 | overflowed << 48 */
    u64 cnt = 0, paths = 0;
    u64 r2 = 0, r1 = 0, r0 = 0, e2 = 0, e1 = 0, e0 = 0, d2 = 0, d1 = 0, d0 = 0;
    int stored = 0; /* children queued: those below the truncation are settled */
    if ((u64)lane < chunk && i < n_items) {
      int s = segment_of(sp, a.n_seg, (u32)base);
      while (sp[s + 1] <= i) ++s;
      const u64 slot_in = (u64)s * a.seg_cap + (i - sp[s]);
      load_item<W>(in, a.stride, slot_in, known, up, mass);
      const u32 tg = tin[slot_in];
      id = tg & kTagIdMask;
      depth = tg >> kTagIdBits;
      active = true;
      const InstDesc d = a.desc[id];
      ex.p = (const double*)(a.tab + d.off_p);
      ex.q = (const double*)(a.tab + d.off_q);
      ex.sadj = (const u64*)(a.tab + d.off_sadj);
      ex.inc = (const u64*)(a.tab + d.off_inc);
      ex.ep = (const u32*)(a.tab + d.off_ep);
      ex.deg = (const u32*)(a.tab + d.off_deg);
      ex.eidm = (const u16*)(a.tab + d.off_eidm);
      ex.n = d.n;
      ex.src = d.src;
      ex.dst = d.dst;
      ex.nterms = 0;
      cnt = 1;
      if (mass < a.eps) { /* main.cu:537-543 */
        fx_limbs_add(e2, e1, e0, mass);
        cnt |= 1ull << 16;
      } else {
        k = ex.find_lean(known, up, lvb);
        if (k >= 0) {
          live = true;
          /* children below the truncation are counted when created (as
           * the truncated nodes they would be when dequeued) and never
           * queued */
          double run = mass;
          for (int j = k - 1; j >= 0; --j) {
            const int e = ex.uh[j * ex.stride];
            if (run * ex.q[e] >= a.eps) ++stored;
            run = run * ex.p[e];
          }
        } else {
          if (k == -2) atomicAdd((unsigned long long*)&a.res[(u64)id * B_N + B_ERR], 1ull);
          fx_limbs_add(d2, d1, d0, mass); /* main.cu:579-582 */
          cnt |= 1ull << 32;
          k = 0;
        }
      }
    }
    const int incl = wave_incl_scan(stored);
    const int total = __shfl(incl, 63, 64);
    u32 wbase = 0;
    u64 sbase = 0;
    bool settled_any = false;
    if (total) {
      if (lane == 63) {
        wbase = atomicAdd(s_cnt, (u32)total);
        /* children past the end of the segment go to the spill stack */
        const u32 fit = wbase < a.seg_cap ? a.seg_cap - wbase : 0u;
        if ((u32)total > fit)
          sbase = atomicAdd((unsigned long long*)spill_top, (unsigned long long)(total - fit));
      }
      wbase = __shfl(wbase, 63, 64);
      sbase = __shfl(sbase, 63, 64);
    }
    if (live) {
      u32 off = wbase + (u32)(incl - stored);
      const u32 seg_end = wbase > a.seg_cap ? wbase : a.seg_cap;
      const u32 child_tag = id | ((depth + 1u) << kTagIdBits);
      double run = mass;
      u64 pre[W];
#pragma unroll
      for (int w = 0; w < W; ++w) pre[w] = 0;
      for (int j = k - 1; j >= 0; --j) {
        const int e = ex.uh[j * ex.stride];
        const int wi = e >> 6;
        const u64 bit = 1ull << (e & 63);
        const double cm = run * ex.q[e];
        if (cm < a.eps) { /* settled at creation (main.cu:537-543) */
          fx_limbs_add(e2, e1, e0, cm);
          cnt += 1ull | (1ull << 16);
          settled_any = true;
          run = run * ex.p[e];
#pragma unroll
          for (int w = 0; w < W; ++w)
            if (w == wi) pre[w] |= bit;
          continue;
        }
        u64 ck[W], cu[W];
#pragma unroll
        for (int w = 0; w < W; ++w) {
          ck[w] = known[w] | pre[w] | (w == wi ? bit : 0ull);
          cu[w] = up[w] | pre[w];
        }
        if (off < a.seg_cap) {
          store_item<W>(out, a.stride, out_base + off, ck, cu, cm);
          tout[out_base + off] = child_tag;
        } else {
          const u64 si = sbase + (off - seg_end);
          if (si < a.spill_cap) {
            store_item<W>(a.spill, a.spill_cap, si, ck, cu, cm);
            a.spill_tag[si] = child_tag;
          } else { /* spill full: main.cu:245-252 */
            fx_limbs_add(e2, e1, e0, cm);
            paths += 1;
            cnt += 1ull << 48;
          }
        }
        ++off;
        run = run * ex.p[e];
#pragma unroll
        for (int w = 0; w < W; ++w)
          if (w == wi) pre[w] |= bit;
      }
      fx_limbs_add(r2, r1, r0, run); /* success term, main.cu:254-262 */
      paths += 1;
    }
    /* flush: lanes of one instance form runs (a parent's children are
     * adjacent); a segmented inclusive scan sums every run at once and the
     * run's last lane adds its sums to the instance (integer sums: the
     * grouping does not change them) */
    {
      const u32 myid = active ? id : 0xffffffffu;
      const u32 prev_id = __shfl_up(myid, 1, 64);
      const u32 next_id = __shfl_down(myid, 1, 64);
      const bool head = lane == 0 || prev_id != myid;
      const bool tail = lane == 63 || next_id != myid;
      const u64 heads = __ballot(head);
      const u64 upto = lane == 63 ? heads : heads & ((2ull << lane) - 1ull);
      const int seg0 = 63 - __builtin_clzll(upto);
      const bool any_e = __ballot((e2 | e1 | e0) != 0) != 0;
      const bool any_d = __ballot((d2 | d1 | d0) != 0) != 0;
      u64 gc = cnt, gp = paths, g_r2 = r2, g_r1 = r1, g_r0 = r0;
      u64 g_e2 = e2, g_e1 = e1, g_e0 = e0, g_d2 = d2, g_d1 = d1, g_d0 = d0;
      u32 gdepth = active ? depth + (settled_any ? 2u : 1u) : 0u;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const bool take = lane - off >= seg0;
        u64 o;
        o = __shfl_up(gc, off, 64); if (take) gc += o;
        o = __shfl_up(gp, off, 64); if (take) gp += o;
        o = __shfl_up(g_r2, off, 64); if (take) g_r2 += o;
        o = __shfl_up(g_r1, off, 64); if (take) g_r1 += o;
        o = __shfl_up(g_r0, off, 64); if (take) g_r0 += o;
        if (any_e) {
          o = __shfl_up(g_e2, off, 64); if (take) g_e2 += o;
          o = __shfl_up(g_e1, off, 64); if (take) g_e1 += o;
          o = __shfl_up(g_e0, off, 64); if (take) g_e0 += o;
        }
        if (any_d) {
          o = __shfl_up(g_d2, off, 64); if (take) g_d2 += o;
          o = __shfl_up(g_d1, off, 64); if (take) g_d1 += o;
          o = __shfl_up(g_d0, off, 64); if (take) g_d0 += o;
        }
        const u32 od = __shfl_up(gdepth, off, 64);
        if (take) gdepth = max(gdepth, od);
      }
      if (tail && active) {
        u64* res = a.res + (u64)id * B_N;
        const u64 c_nodes = gc & 0xffff, c_tr = (gc >> 16) & 0xffff, c_dead = (gc >> 32) & 0xffff,
                  c_over = gc >> 48;
        atomicAdd((unsigned long long*)&res[B_NODES], (unsigned long long)c_nodes);
        atomicMax((unsigned long long*)&res[B_LEVELS], (unsigned long long)gdepth);
        if (c_tr) atomicAdd((unsigned long long*)&res[B_TRUNC], (unsigned long long)c_tr);
        if (c_dead) atomicAdd((unsigned long long*)&res[B_DEAD], (unsigned long long)c_dead);
        if (c_over) atomicAdd((unsigned long long*)&res[B_OVER], (unsigned long long)c_over);
        if (gp) atomicAdd((unsigned long long*)&res[B_PATHS], (unsigned long long)gp);
        fx_flush(&res[B_R], g_r2, g_r1, g_r0);
        if (any_e) fx_flush(&res[B_E], g_e2, g_e1, g_e0);
        if (any_d) fx_flush(&res[B_D], g_d2, g_d1, g_d0);
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const u32 n = *s_cnt;
    a.seg_count[parity ^ 1][blockIdx.x] = n < a.seg_cap ? n : a.seg_cap;
  }
}

/* Time-budget fold of a batch's pending items (the level about to run,
 * buffer `parity`, and the spill stack) into each instance's ESP. */
__global__ void __launch_bounds__(256) k_fold_batch(BatchArgs a, int parity) {
  extern __shared__ __align__(16) unsigned char smem[];
  u32* sp = (u32*)smem;
  u32* part = sp + a.n_seg + 2;
  block_prefix(a.seg_count[parity], a.n_seg, sp, part);
  const u32 n_items = sp[a.n_seg];
  const u64 top = a.sc->top[parity ^ 1] < a.spill_cap ? a.sc->top[parity ^ 1] : a.spill_cap;
  const u64 tid = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u64 nthreads = (u64)gridDim.x * blockDim.x;
  for (u64 i = tid; i < (u64)n_items + top; i += nthreads) {
    u32 tg;
    double m;
    if (i < n_items) {
      int s = segment_of(sp, a.n_seg, (u32)i);
      const u64 slot = (u64)s * a.seg_cap + (i - sp[s]);
      tg = a.tag[parity][slot];
      m = a.buf[parity].mass[slot];
    } else {
      tg = a.spill_tag[i - n_items];
      m = a.spill.mass[i - n_items];
    }
    u64* res = a.res + (u64)(tg & kTagIdMask) * B_N;
    fx_add(&res[B_E], m);
    atomicAdd((unsigned long long*)&res[B_FOLD], 1ull);
  }
}

/* Reset a batch pass: empty segments and spill, admission from perm[0]. */
__global__ void __launch_bounds__(256) k_init_batch(BatchArgs a) {
  const u64 tid = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u64 stride = (u64)gridDim.x * blockDim.x;
  for (u64 s = tid; s < (u64)a.n_seg; s += stride) {
    a.seg_count[0][s] = 0u;
    a.seg_count[1][s] = 0u;
  }
  if (tid == 0) {
    LevelCtl z = {};
    *a.ctl = z;
    StreamCtl c = {};
    *a.sc = c;
  }
}

/* ------------------------------------------------------------------ */
/* host runtime                                                        */
/* ------------------------------------------------------------------ */

hipError_t& last_hip_error() {
  static thread_local hipError_t e = hipSuccess;
  return e;
}

#define NR_HIP(call)                                   \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) {                            \
      last_hip_error() = _e;                           \
      return NR_ERR_DEVICE;                            \
    }                                                  \
  } while (0)

/* Per-device workspace kept across calls: queue buffers, graph tables,
 * partial-sum slots, control block (device + pinned host copy), a stream
 * and an event pool.  Buffers grow on demand and are reused, so repeated
 * solves pay no allocation.  One mutex per device serialises calls. */
/* Graph tables packed into one buffer: p, q (FP64), static neighbour sets,
 * incident-edge sets, edge endpoints, incident counts, the (u,v) -> edge id
 * matrix.  Offsets are 16-byte aligned. */
struct TableBlob {
  std::vector<unsigned char> bytes;
  size_t off_p, off_q, off_sadj, off_inc, off_ep, off_deg, off_eidm;
  size_t off_rp, off_arc, off_irp, off_iarc, off_terms, off_lk, off_lqp, off_lqq;
};

struct Workspace {
  bool ready = false;
  int device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t prop;
  void* q = nullptr;
  size_t q_bytes = 0;
  void* tab = nullptr;
  size_t tab_bytes = 0;
  void* meta = nullptr; /* per-wave slots, segment counts, prefix tables */
  size_t meta_bytes = 0;
  LevelCtl* ctl = nullptr;
  LevelCtl* hctl = nullptr; /* pinned */
  u64* hmirror = nullptr;   /* pinned [kMirWords], written by the level kernel and k_finalize */
  u64 solves = 0;           /* sequence number of the last solve */
  u64* dmirror = nullptr;   /* its device address */
  /* device copy of the last batch's instance tables, roots and admission
   * order, reused while the same graphs (pointer and stamp) come back */
  void* btab = nullptr;
  size_t btab_bytes = 0;
  std::vector<std::pair<const nr_graph*, unsigned long long>> bkey;
  int bW = 0;
  std::vector<u32> bperm; /* the admission order on the host */
  size_t bo_desc = 0, bo_tabs = 0, bo_roots = 0, bo_perm = 0, bo_perm2 = 0;
  std::vector<hipEvent_t> events;
  std::mutex mu;
  /* the graph whose tables ws.tab holds (pointer, stamp, plane words) and
   * their host image, kept so repeated solves of one graph upload nothing */
  TableBlob tab_blob;
  const nr_graph* tab_g = nullptr;
  unsigned long long tab_stamp = 0;
  int tab_W = 0;
  /* no-allocation mode (nr_solve_static): every device buffer is carved
   * from the file-scope scratch array (offsets restart with each solve), the
   * null stream is used, and no events or pinned host memory exist: the
   * counters are read back with plain copies into host_ctl */
  bool fixed = false;
  unsigned char* arena = nullptr;
  size_t arena_size = 0, arena_used = 0;
  LevelCtl host_ctl;

  int grow(void** p, size_t* have, size_t need) {
    if (fixed) {
      const size_t at = (arena_used + 255) / 256 * 256;
      if (at > arena_size || need > arena_size - at) return NR_ERR_SCRATCH;
      *p = arena + at;
      *have = need;
      arena_used = at + need;
      return NR_OK;
    }
    if (need <= *have) return NR_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *have = 0;
    if (hipMalloc(p, need) != hipSuccess) {
      *p = nullptr;
      return NR_ERR_NOMEM;
    }
    *have = need;
    return NR_OK;
  }
  /* per-kernel launch facts, looked up once: the dynamic-LDS limit set on a
   * kernel, and the occupancy query per (kernel, block, LDS bytes) */
  struct Occ {
    const void* k;
    int block, lds, per_cu;
  };
  std::vector<std::pair<const void*, int>> lds_attr;
  std::vector<Occ> occ;
  int allow_lds(const void* k, int bytes) {
    for (const auto& x : lds_attr)
      if (x.first == k && x.second == bytes) return NR_OK;
    NR_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    lds_attr.push_back({k, bytes});
    return NR_OK;
  }
  int blocks_per_cu(const void* k, int block, int lds, int* out) {
    for (const Occ& o : occ)
      if (o.k == k && o.block == block && o.lds == lds) {
        *out = o.per_cu;
        return NR_OK;
      }
    int n = 0;
    NR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, block, lds));
    occ.push_back({k, block, lds, n});
    *out = n;
    return NR_OK;
  }
  int event(size_t i, hipEvent_t* e) {
    while (events.size() <= i) {
      hipEvent_t x;
      NR_HIP(hipEventCreate(&x));
      events.push_back(x);
    }
    *e = events[i];
    return NR_OK;
  }
  int open(int dev) {
    if (ready) return NR_OK;
    device = dev;
    NR_HIP(hipSetDevice(dev));
    NR_HIP(hipGetDeviceProperties(&prop, dev));
    NR_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    NR_HIP(hipMalloc((void**)&ctl, sizeof(LevelCtl)));
    NR_HIP(hipHostMalloc((void**)&hctl, sizeof(LevelCtl), hipHostMallocDefault));
    NR_HIP(hipHostMalloc((void**)&hmirror, kMirWords * sizeof(u64), hipHostMallocDefault));
    NR_HIP(hipHostGetDevicePointer((void**)&dmirror, hmirror, 0));
    ready = true;
    return NR_OK;
  }
  void release() {
    if (!ready) return;
    (void)hipSetDevice(device);
    (void)hipStreamSynchronize(stream);
    for (auto e : events) (void)hipEventDestroy(e);
    events.clear();
    if (q) (void)hipFree(q);
    if (tab) (void)hipFree(tab);
    if (btab) (void)hipFree(btab);
    btab = nullptr;
    btab_bytes = 0;
    bkey.clear();
    tab_g = nullptr; /* the device tables are gone */
    lds_attr.clear();
    occ.clear();
    if (meta) (void)hipFree(meta);
    if (ctl) (void)hipFree(ctl);
    if (hctl) (void)hipHostFree(hctl);
    if (hmirror) (void)hipHostFree(hmirror);
    hmirror = dmirror = nullptr;
    (void)hipStreamDestroy(stream);
    q = tab = meta = nullptr;
    ctl = hctl = nullptr;
    q_bytes = tab_bytes = meta_bytes = 0;
    ready = false;
  }
};

constexpr int kMaxDevices = 64;
Workspace g_ws[kMaxDevices];

/* The no-allocation entry point's device memory: one fixed array per GPU
 * (each device holds its own copy at its own address until the process
 * ends).  Host code never passes the array's name to a kernel or a copy: it
 * uses the address hipGetSymbolAddress returns for the current device. */
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
Workspace g_ws_fixed[kMaxDevices];

void pack_tables(const nr_graph* g, int W, TableBlob& b) {
  const int n = g->n, e = std::max(g->e, 1);
  /* node-set tables exist for the row form only (n <= 64) */
  const int ns = n <= 64 ? n : 0;
  /* the head of the blob is the table part of the row form's LDS layout,
   * offset for offset, so a workgroup stages it with one flat copy */
  const Layout T = make_layout(ns, e, W, 64, 0, 0, true);
  size_t o = (size_t)T.table_bytes;
  auto take = [&](size_t bytes) {
    o = (o + 15) / 16 * 16;
    size_t at = o;
    o += bytes;
    return at;
  };
  b.off_p = (size_t)T.off_p;
  b.off_q = (size_t)T.off_q;
  b.off_sadj = (size_t)T.off_sadj;
  b.off_inc = (size_t)T.off_inc;
  b.off_ep = (size_t)T.off_ep;
  b.off_deg = (size_t)T.off_deg;
  b.off_eidm = (size_t)T.off_eidm;
  b.off_terms = take(2 * (size_t)std::max<size_t>(g->terms.size(), 1));
  b.off_rp = take(4 * (size_t)(n + 1));
  b.off_arc = take(4 * (size_t)std::max(g->nnz, 1));
  b.off_irp = take(4 * (size_t)(n + 1));
  b.off_iarc = take(4 * (size_t)std::max(g->nnz, 1));
  b.off_lk = take(4 * (size_t)e);
  b.off_lqp = take((size_t)e);
  b.off_lqq = take((size_t)e);
  b.bytes.assign((o + 15) / 16 * 16, 0);
  double* p = (double*)&b.bytes[b.off_p];
  double* q = (double*)&b.bytes[b.off_q];
  u64* sadj = (u64*)&b.bytes[b.off_sadj];
  u64* inc = (u64*)&b.bytes[b.off_inc];
  u32* ep = (u32*)&b.bytes[b.off_ep];
  u32* deg = (u32*)&b.bytes[b.off_deg];
  u16* eidm = (u16*)&b.bytes[b.off_eidm];
  float* lk = (float*)&b.bytes[b.off_lk];
  for (size_t i = 0; i < (size_t)ns * ns; ++i) eidm[i] = 0xffff;
  for (int id = 0; id < g->e; ++id) {
    const int a = g->ea[id], c = g->eb[id];
    p[id] = g->p[id];
    q[id] = g->q[id];
    /* forkBest's key, computed on the host as graph.h:138 does */
    const float pf = (float)g->p[id];
    lk[id] = pf > 0.0f ? logf(pf) : -1e30f;
    /* the learned chooser's probability steps: round(-4 log2 x) in 0..63 */
    auto lq = [](double x) -> unsigned char {
      if (!(x > 0.0)) return 63;
      const double v = std::nearbyint(-4.0 * std::log2(x));
      return (unsigned char)(v < 0.0 ? 0.0 : (v > 63.0 ? 63.0 : v));
    };
    b.bytes[b.off_lqp + id] = lq(g->p[id]);
    b.bytes[b.off_lqq + id] = lq(g->q[id]);
    ep[id] = (u32)a | ((u32)c << 16);
    if (!ns) continue;
    inc[(size_t)a * W + (id >> 6)] |= 1ull << (id & 63);
    inc[(size_t)c * W + (id >> 6)] |= 1ull << (id & 63);
    if (a != c) {
      sadj[a] |= 1ull << c;
      sadj[c] |= 1ull << a;
    }
    eidm[(size_t)a * n + c] = (u16)id;
    eidm[(size_t)c * n + a] = (u16)id;
  }
  for (int u = 0; u < ns; ++u) {
    int cnt = 0;
    for (int i = 0; i < W; ++i) cnt += __builtin_popcountll(inc[(size_t)u * W + i]);
    deg[u] = (u32)cnt;
  }
  /* CSR arcs for the CSR expander: out-rows sorted by (v, e), in-rows by
   * (u, e); an arc is packed node | edge << 16 */
  std::vector<std::vector<u32>> out(n), in(n);
  for (int u = 0; u < n; ++u)
    for (int j = g->row_ptr[u]; j < g->row_ptr[u + 1]; ++j) {
      const int v = g->col_idx[j], id = g->edge_id[j];
      out[u].push_back((u32)v | ((u32)id << 16));
      in[v].push_back((u32)u | ((u32)id << 16));
    }
  auto by_node_then_edge = [](u32 x, u32 y) {
    return (x & 0xffffu) != (y & 0xffffu) ? (x & 0xffffu) < (y & 0xffffu) : (x >> 16) < (y >> 16);
  };
  int* rp = (int*)&b.bytes[b.off_rp];
  int* irp = (int*)&b.bytes[b.off_irp];
  u32* arc = (u32*)&b.bytes[b.off_arc];
  u32* iarc = (u32*)&b.bytes[b.off_iarc];
  int k = 0, ik = 0;
  for (int u = 0; u < n; ++u) {
    std::sort(out[u].begin(), out[u].end(), by_node_then_edge);
    std::sort(in[u].begin(), in[u].end(), by_node_then_edge);
    rp[u] = k;
    irp[u] = ik;
    for (u32 x : out[u]) arc[k++] = x;
    for (u32 x : in[u]) iarc[ik++] = x;
  }
  rp[n] = k;
  irp[n] = ik;
  u16* tm = (u16*)&b.bytes[b.off_terms];
  for (size_t i = 0; i < g->terms.size(); ++i) tm[i] = (u16)g->terms[i];
}

int upload_tables(Workspace& ws, const nr_graph* g, int W, DevGraph& d) {
  /* the device copy stays valid while the same graph (pointer and stamp:
   * a new graph or new terminals get a new stamp) comes back with the same W */
  TableBlob& b = ws.tab_blob;
  if (!(ws.tab_g == g && ws.tab_stamp == g->stamp && ws.tab_W == W)) {
    ws.tab_g = nullptr;
    pack_tables(g, W, b);
    int rc = ws.grow(&ws.tab, &ws.tab_bytes, b.bytes.size());
    if (rc) return rc;
    NR_HIP(hipMemcpyAsync(ws.tab, b.bytes.data(), b.bytes.size(), hipMemcpyHostToDevice,
                          ws.stream));
    ws.tab_g = g;
    ws.tab_stamp = g->stamp;
    ws.tab_W = W;
  }
  unsigned char* base = (unsigned char*)ws.tab;
  d.p = (const double*)(base + b.off_p);
  d.q = (const double*)(base + b.off_q);
  d.sadj = (const u64*)(base + b.off_sadj);
  d.inc = (const u64*)(base + b.off_inc);
  d.ep = (const u32*)(base + b.off_ep);
  d.deg = (const u32*)(base + b.off_deg);
  d.eidm = (const u16*)(base + b.off_eidm);
  d.rp = (const int*)(base + b.off_rp);
  d.arc = (const u32*)(base + b.off_arc);
  d.irp = (const int*)(base + b.off_irp);
  d.iarc = (const u32*)(base + b.off_iarc);
  d.terms = (const u16*)(base + b.off_terms);
  d.lk = (const float*)(base + b.off_lk);
  d.lqp = base + b.off_lqp;
  d.lqq = base + b.off_lqq;
  d.nterms = (int)g->terms.size();
  for (int k = 0; k < 4; ++k) d.tmask[k] = 0;
  for (int t : g->terms) d.tmask[t >> 6] |= 1ull << (t & 63);
  d.n = g->n;
  d.e = g->e;
  d.w = W;
  d.src = g->src;
  d.dst = g->dst;
  d.period = 2;
  d.pivot = 0;
  return NR_OK;
}

/* Width classes: W 64-bit words per edge-state plane, node sets of 32 or 64
 * bits.  Each class is a separate instantiation of the kernels. */
template <int W, typename NS>
struct Kern {
  static void* level() { return (void*)&k_expand_level<W, Expander<W, NS>, 0>; }
  static void* items() { return (void*)&k_expand_items<W, Expander<W, NS>>; }
  static void* batch() { return (void*)&k_expand_batch<W, NS>; }
  static void* refill() { return (void*)&k_refill_batch<W>; }
  static void* seed_rows() { return (void*)&k_seed_rows<W, NS>; }
  static void* small() { return (void*)&k_small_levels<W, Expander<W, NS>>; }
  static void* ramp() { return (void*)&k_ramp_levels<W, Expander<W, NS>>; }
};

template <typename NS>
Variant variant_for_w(int W) {
  switch (W) {
#define NR_W(X)                                                                           \
  case X: {                                                                               \
    Variant v{&Kern<X, NS>::level, &Kern<X, NS>::items, &Kern<X, NS>::batch,              \
              &Kern<X, NS>::refill, (int)sizeof(NS), 0, nullptr, nullptr, nullptr,        \
              &Kern<X, NS>::seed_rows, 0, nullptr};                                       \
    v.small = &Kern<X, NS>::small;                                                        \
    v.ramp = &Kern<X, NS>::ramp;                                                          \
    v.settle = 1;                                                                         \
    return v;                                                                             \
  }
    NR_W(1) NR_W(2) NR_W(3) NR_W(4) NR_W(5) NR_W(6) NR_W(7) NR_W(8)
#undef NR_W
    default: return Variant{nullptr, nullptr, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr, 0, nullptr};
  }
}

/* LDS layout of a variant's expand kernels for graph g. */
Layout layout_for(const Variant& v, const nr_graph* g, int W, int block, int extra = 0) {
  if (v.nw) return make_layout_csr(g->n, block, extra);
  if (v.lean) return make_layout_lean(g->n, std::max(g->e, 1), W, block, extra);
  return make_layout(g->n, std::max(g->e, 1), W, block, v.ns_bytes, extra);
}

/* Largest block (<= 256, multiple of 64) whose LDS footprint fits in 64 KiB;
 * a requested size is honoured when it fits. */
int choose_block(const Variant& v, const nr_graph* g, int W, int requested, int extra = 0) {
  const int cands[4] = {requested, 256, 128, 64};
  for (int b : cands) { /* 64, 128 or 256: LDS slices are strided by a power of two */
    if (b != 64 && b != 128 && b != 256) continue;
    if (layout_for(v, g, W, b, extra).total_bytes <= 64 * 1024) return b;
  }
  return 0;
}

/* Row form (items carry adjacency rows) for undirected graphs of up to 64
 * nodes; CSR form for up to 256 nodes or one-way arcs.  Up to 512 edges. */
int pick_variant(const nr_graph* g, Variant& v) {
  const int W = mask_words(g);
  if (W > 8 || g->n > 256 || g->n < 1) return NR_ERR_UNSUPPORTED;
  if (g->symmetric && g->n <= 64)
    v = g->n <= 32 ? variant_for_w<u32>(W) : variant_for_w<u64>(W);
  else
    v = variant_csr(W, (g->n + 63) / 64);
  if (!v.level) return NR_ERR_UNSUPPORTED;
  add_learn_kernels(v, W);
  add_fork_kernels(v, W);
  add_mass_kernels(v, W);
  return NR_OK;
}

Root root_item(const nr_graph* g, int W) {
  /* main.cu:717-748: p <= 0 FAILED, p >= 1 WORKING, otherwise UNKNOWN */
  Root r;
  std::memset(&r, 0, sizeof r);
  for (int e = 0; e < g->e; ++e) {
    u64 bit = 1ull << (e & 63);
    if (g->p[e] <= 0.0) {
      r.known[e >> 6] |= bit;
    } else if (g->p[e] >= 1.0) {
      r.known[e >> 6] |= bit;
      r.up[e >> 6] |= bit;
    }
    /* rows (row form, n <= 64): neighbours over edges that are not FAILED */
    if (g->n <= 64 && g->p[e] > 0.0 && g->ea[e] != g->eb[e]) {
      r.rows[g->ea[e]] |= 1ull << g->eb[e];
      r.rows[g->eb[e]] |= 1ull << g->ea[e];
    }
  }
  (void)W;
  return r;
}

/* RAII bag of device allocations made for one call. */
struct DevMem {
  std::vector<void*> ptrs;
  ~DevMem() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  int alloc(T** out, size_t count) {
    void* p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) return NR_ERR_NOMEM;
    ptrs.push_back(p);
    *out = (T*)p;
    return NR_OK;
  }
};

}  // namespace nr

/* A solve's pending subproblems and sums, kept when its time budget
 * expired (host memory). */
struct nr_checkpoint {
  int n = 0, e = 0, w = 0, src = 0, dst = 0, rule = 0, order = 0;
  int period = 0; /* sweep period of NR_RULE_ALTERNATE (0 otherwise) */
  double eps = 0.0;
  unsigned long long graph_hash = 0;
  double sum_hi[3] = {0, 0, 0}, sum_lo[3] = {0, 0, 0}; /* R, ESP, D */
  unsigned long long counts[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  unsigned long long iterations = 0, max_width = 0;
  unsigned long long count = 0;
  std::vector<unsigned long long> known, up; /* [count][w] */
  std::vector<double> mass;
};

namespace nr {

/* FNV-1a over the graph's shape, arcs, probabilities and terminals. */
u64 graph_hash(const nr_graph* g) {
  u64 h = 1469598103934665603ull;
  auto mix = [&](const void* p, size_t bytes) {
    const unsigned char* c = (const unsigned char*)p;
    for (size_t i = 0; i < bytes; ++i) h = (h ^ c[i]) * 1099511628211ull;
  };
  mix(&g->n, sizeof g->n);
  mix(&g->e, sizeof g->e);
  mix(&g->src, sizeof g->src);
  mix(&g->dst, sizeof g->dst);
  mix(g->row_ptr.data(), sizeof(int) * g->row_ptr.size());
  mix(g->col_idx.data(), sizeof(int) * g->col_idx.size());
  mix(g->edge_id.data(), sizeof(int) * g->edge_id.size());
  mix(g->p.data(), sizeof(double) * g->p.size());
  mix(g->terms.data(), sizeof(int) * g->terms.size());
  return h;
}

/* Extras of a level-engine solve: the tree log (mode 1) or a learned
 * scorer (mode 2). */
struct LevelExtra {
  int mode = 0;
  float* log_feat = nullptr;
  u32* log_parent = nullptr;
  u32* log_level = nullptr;
  u64* log_count = nullptr;
  u64 log_cap = 0;
  double* log_mass = nullptr; /* conditional-reliability labels (optional) */
  double* log_succ = nullptr;
  const MlpDev* mlp = nullptr;
  const MlpDev* chooser = nullptr; /* NR_RULE_LEARNED: the path chooser */
  float* log_choice = nullptr;     /* tree log of the choices (mode 1) */
  const nr_checkpoint* resume = nullptr; /* start from these pending items */
  nr_checkpoint** save = nullptr;        /* on deadline keep them here */
};

/* The path rule in level order on graphs past the per-lane forms (more than
 * 256 nodes or 512 edges): one wave per subproblem (nr_big.hip).  Same
 * counting, capacity, truncation and time-budget semantics as the level
 * kernel; items are planes of W words stored item-major. */
int solve_big(Workspace& ws, const nr_graph* g, const nr_config* cfg, nr_result* out) {
  if (g->n > 65535 || g->e > 65535) return NR_ERR_UNSUPPORTED; /* arcs pack node | edge << 16 */
  if (!g->terms.empty() || cfg->rule != NR_RULE_PATH || cfg->order != NR_ORDER_LEVEL)
    return NR_ERR_UNSUPPORTED;
  const int W = mask_words(g);
  const int cus = ws.prop.multiProcessorCount;
  const void* kfun = big_expand_kernel();
  NR_HIP(hipFuncSetAttribute(kfun, hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024));
  /* workgroups of 4, 2 or 1 waves, the most whose slices fit in 64 KiB
   * beside the segment prefix table */
  int block = 0, grid = 0, per_cu = 0;
  Layout lay;
  BigShape bs;
  for (int waves = 4; waves >= 1 && !block; waves /= 2) {
    const int b = 64 * waves;
    for (int pass = 0; pass < 2; ++pass) { /* size the grid, then fit its table */
      const int g0 = pass ? grid : 4 * cus;
      const int extra = 4 * (g0 + 2 + b + kClasses);
      bs = big_shape(g->n, W, extra);
      const int total = bs.off_waves + waves * bs.wave_bytes;
      if (total > 64 * 1024) break;
      int pc = 0;
      NR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, kfun, b, total));
      if (pc < 1) break;
      if (pass == 0) {
        grid = cfg->grid_blocks > 0 ? cfg->grid_blocks : pc * cus;
        continue;
      }
      block = b;
      per_cu = pc;
      std::memset(&lay, 0, sizeof lay);
      lay.n = g->n;
      lay.e = g->e;
      lay.w = W;
      lay.block = b;
      lay.off_extra = 0;
      lay.total_bytes = total;
    }
  }
  if (!block) return NR_ERR_UNSUPPORTED;
  const u64 n_slots = (u64)grid;
  const u64 cap = cfg->queue_capacity > 0 ? (u64)cfg->queue_capacity : (1ull << 20);
  const u64 seg_cap64 = std::max<u64>(1, (cap + grid - 1) / grid);
  if (seg_cap64 * (u64)grid >= (1ull << 32)) return NR_ERR_ARG;
  const u32 seg_cap = (u32)seg_cap64;
  const u64 slots = seg_cap64 * (u64)grid;
  int rc;
  LevelArgs a;
  std::memset(&a, 0, sizeof a);
  if ((rc = upload_tables(ws, g, W, a.g))) return rc;
  a.lay = lay;
  a.seg_cap = seg_cap;
  a.n_seg = grid;
  a.stride = slots;
  a.eps = cfg->truncation_eps > 0 ? cfg->truncation_eps : 0.0;
  a.budget_ticks = 0;
  if (cfg->time_budget_s > 0) {
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ws.device) == hipSuccess &&
        khz > 0)
      a.budget_ticks = std::max<u64>(1, (u64)(cfg->time_budget_s * 1e3 * (double)khz));
  }
  /* buffers: planes [slots][2W] u64, then masses [slots] */
  const size_t per_buf = ((size_t)2 * W * slots + slots) * 8;
  if ((rc = ws.grow(&ws.q, &ws.q_bytes, 2 * per_buf + 16 * (size_t)W))) return rc;
  for (int b = 0; b < 2; ++b) {
    u64* base = (u64*)((unsigned char*)ws.q + b * per_buf);
    a.buf[b].known = base;
    a.buf[b].up = nullptr;
    a.buf[b].mass = (double*)(base + (size_t)2 * W * slots);
    a.buf[b].rows = nullptr;
  }
  u64* droot = (u64*)((unsigned char*)ws.q + 2 * per_buf);
  const size_t meta_bytes = n_slots * (3 * sizeof(DD) + C_N * sizeof(u64)) +
                            sizeof(u32) * (4 * (size_t)grid + 8);
  if ((rc = ws.grow(&ws.meta, &ws.meta_bytes, meta_bytes))) return rc;
  {
    unsigned char* m = (unsigned char*)ws.meta;
    a.partR = (DD*)m;
    a.partE = a.partR + n_slots;
    a.partD = a.partE + n_slots;
    a.cnt = (u64*)(a.partD + n_slots);
    u32* q = (u32*)(a.cnt + n_slots * C_N);
    a.seg_count[0] = q;
    a.seg_count[1] = q + grid;
    a.seg_settle[0] = q + 2 * grid;
    a.seg_settle[1] = q + 3 * grid;
  }
  /* children below the truncation are counted when they are created and
   * never queued (the level widths count them) */
  a.settle = (cfg->flags & NR_FLAG_QUEUE_ALL) ? 0 : kSettleTrunc;
  a.ctl = ws.ctl;
  hipStream_t st = ws.stream;
  {
    /* the root (main.cu:717-748): p <= 0 FAILED, p >= 1 WORKING */
    std::vector<u64> root((size_t)2 * W, 0);
    for (int e = 0; e < g->e; ++e) {
      const u64 bit = 1ull << (e & 63);
      if (g->p[e] <= 0.0) {
        root[e >> 6] |= bit;
      } else if (g->p[e] >= 1.0) {
        root[e >> 6] |= bit;
        root[W + (e >> 6)] |= bit;
      }
    }
    NR_HIP(hipMemcpyAsync(droot, root.data(), 16 * (size_t)W, hipMemcpyHostToDevice, st));
    NR_HIP(hipStreamSynchronize(st));
  }
  hipEvent_t ev_start, ev_stop;
  if ((rc = ws.event(0, &ev_start)) || (rc = ws.event(1, &ev_stop))) return rc;
  const size_t max_timed = (cfg->flags & NR_FLAG_TIME_LAUNCHES) ? 2048 : 0;
  auto t_begin = std::chrono::steady_clock::now();
  NR_HIP(hipEventRecord(ev_start, st));
  {
    const u64* dr = droot;
    int Wi = W;
    u64 ns = n_slots;
    void* args[] = {&a, &dr, &Wi, &ns};
    NR_HIP(hipLaunchKernel(big_init_kernel(), dim3(64), dim3(256), args, 0, st));
  }
  long long timed = 0;
  int level = 0, chunk = cfg->time_budget_s > 0 ? 1 : 8;
  bool finished = false, deadline = false;
  auto expired = [&]() {
    return cfg->time_budget_s > 0 &&
           std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count() >=
               cfg->time_budget_s;
  };
  while (!finished) {
    for (int c = 0; c < chunk; ++c, ++level) {
      hipEvent_t kb = nullptr, ke = nullptr;
      if ((size_t)timed < max_timed) {
        if ((rc = ws.event(2 + 2 * timed, &kb)) || (rc = ws.event(3 + 2 * timed, &ke))) return rc;
        NR_HIP(hipEventRecord(kb, st));
      }
      int par = level & 1;
      void* args[] = {&a, &par, &bs};
      NR_HIP(hipLaunchKernel(kfun, dim3(grid), dim3(block), args, lay.total_bytes, st));
      if (kb) {
        NR_HIP(hipEventRecord(ke, st));
        ++timed;
      }
    }
    NR_HIP(hipMemcpyAsync(ws.hctl, a.ctl, kCtlCopyBytes, hipMemcpyDeviceToHost, st));
    NR_HIP(hipStreamSynchronize(st));
    if (ws.hctl->counts[C_ERR] || ws.hctl->last_items == 0) finished = true;
    if (!finished && expired()) {
      int par = level & 1;
      void* args[] = {&a, &par};
      NR_HIP(hipLaunchKernel((const void*)&k_fold, dim3(grid), dim3(256), args,
                             4 * (grid + 2 + 256), st));
      deadline = true;
      finished = true;
    }
    chunk = std::min(chunk * 2, 16);
    if (ws.hctl->last_items < ws.hctl->max_width) chunk = std::min(chunk, 4);
    if (cfg->time_budget_s > 0) chunk = 1;
  }
  {
    u64 nw = n_slots;
    void* args[] = {&a, &nw};
    NR_HIP(hipLaunchKernel((const void*)&k_finalize, dim3(finalize_groups(n_slots)), dim3(256),
                           args, 0, st));
  }
  NR_HIP(hipEventRecord(ev_stop, st));
  NR_HIP(hipMemcpyAsync(ws.hctl, a.ctl, kCtlCopyBytes, hipMemcpyDeviceToHost, st));
  NR_HIP(hipStreamSynchronize(st));
  float ms = 0.f;
  NR_HIP(hipEventElapsedTime(&ms, ev_start, ev_stop));
  const LevelCtl& h = *ws.hctl;
  double kms = 0.0;
  for (long long i = 0; i < timed; ++i) {
    float xm = 0.f;
    NR_HIP(hipEventElapsedTime(&xm, ws.events[2 + 2 * i], ws.events[3 + 2 * i]));
    kms += xm;
  }
  if (h.counts[C_ERR]) return NR_ERR_UNSUPPORTED; /* a path longer than 254 hops */
  std::memset(out, 0, sizeof(*out));
  out->R = h.sum[0];
  out->ESP = h.sum[1];
  out->D = h.sum[2];
  out->paths = h.counts[C_PATHS];
  out->nodes = h.counts[C_NODES];
  out->dead = h.counts[C_DEAD];
  out->truncated = h.counts[C_TRUNC];
  out->overflowed = h.counts[C_OVER];
  out->folded = h.counts[C_FOLD];
  out->iterations = (long long)h.iterations;
  out->max_queue_items = (long long)h.max_width;
  out->item_bytes = 16LL * W + 8;
  out->elapsed_ms = ms;
  out->kernel_ms = kms;
  out->kernel_launches = timed;
  out->deadline_hit = deadline ? 1 : 0;
  out->overflowed_flag = h.counts[C_OVER] ? 1 : 0;
  out->grid_blocks = grid;
  out->block_size = block;
  out->lds_bytes = lay.total_bytes;
  out->blocks_per_cu = per_cu;
  out->compute_units = cus;
  return NR_OK;
}

/* The seeds of the static prefix partition (original/Rel_16M.cu:97-101,
 * where thread-id bits fix the first edges): the root's first t UNKNOWN
 * edges in id order, seed i with edge j WORKING when bit j of i is set and
 * FAILED otherwise, mass the product of the p and q so chosen (FP64, edges
 * in id order); only seeds i % shards == shard when shards > 1. */
int prefix_seeds(const nr_graph* g, int W, const nr_config* cfg,
                 std::unique_ptr<nr_checkpoint>& out) {
  const int t = cfg->prefix_edges;
  const int shards = cfg->shards > 1 ? cfg->shards : 1;
  if (t < 0 || t > 24 || cfg->shard < 0 || cfg->shard >= shards) return NR_ERR_ARG;
  const Root root = root_item(g, W);
  std::vector<int> free_e;
  for (int e = 0; e < g->e && (int)free_e.size() < t; ++e)
    if (!((root.known[e >> 6] >> (e & 63)) & 1ull)) free_e.push_back(e);
  const int tt = (int)free_e.size(); /* fewer when the root has fewer UNKNOWN edges */
  out.reset(new (std::nothrow) nr_checkpoint);
  if (!out) return NR_ERR_NOMEM;
  nr_checkpoint& c = *out;
  const u64 all = 1ull << tt;
  for (u64 i = (u64)cfg->shard; i < all; i += (u64)shards) {
    double m = 1.0;
    std::vector<u64> k(root.known, root.known + W), u(root.up, root.up + W);
    for (int j = 0; j < tt; ++j) {
      const int e = free_e[j];
      const u64 b = 1ull << (e & 63);
      k[e >> 6] |= b;
      if ((i >> j) & 1ull) {
        u[e >> 6] |= b;
        m *= g->p[e];
      } else {
        m *= g->q[e];
      }
    }
    c.known.insert(c.known.end(), k.begin(), k.end());
    c.up.insert(c.up.end(), u.begin(), u.end());
    c.mass.push_back(m);
  }
  c.count = c.mass.size();
  c.w = W;
  return NR_OK;
}

int solve_level(Workspace& ws, const nr_graph* g, const nr_config* cfg, nr_result* out,
                const LevelExtra* x = nullptr) {
  Variant var;
  int rc = pick_variant(g, var);
  if (rc) return rc;
  const int mode = x ? x->mode : 0;
  /* the plain path rule on a row-form graph whose rows would exceed 64 bytes
   * per item runs with items that carry none, unless the caller asks for
   * the row-carrying kernel (measured on MI355X, level order: K8 0.207 ms
   * with rows vs 0.241 ms without, K10 0.758 vs 1.24 ms, K15 at 1e-12 8.8
   * vs 10.3 ms; K25, 100 B of rows, 54.9 vs 30.6 ms) */
  if (mode == 0 && cfg->rule == NR_RULE_PATH && g->terms.empty() && !var.nw &&
      (g->n * var.ns_bytes > 64 || (cfg->flags & NR_FLAG_LEAN)) && !(cfg->flags & NR_FLAG_ROWS)) {
    var = variant_lean(mask_words(g), var.ns_bytes);
    add_mass_kernels(var, mask_words(g));
  }
  if (cfg->rule == NR_RULE_ALTERNATE || cfg->rule == NR_RULE_LEARNED) {
    if (!g->symmetric || !g->terms.empty()) return NR_ERR_UNSUPPORTED;
    if (cfg->rule == NR_RULE_LEARNED) {
      /* the chooser scores directions, except in a tree log, which explores */
      if (mode != 1 && !(x && x->chooser)) return NR_ERR_ARG;
      var = g->n <= 64 ? variant_choose(mask_words(g), var.ns_bytes)
                       : variant_choose_csr(mask_words(g), (g->n + 63) / 64);
    } else {
      var = g->n <= 64 ? variant_alt(mask_words(g), var.ns_bytes)
                       : variant_alt_csr(mask_words(g), (g->n + 63) / 64);
    }
    if (!var.level) return NR_ERR_UNSUPPORTED;
  }
  /* per-wave feature tiles (learned order, learned choice) */
  const int tile_bytes = (mode == 2 || var.choose) ? 32 * 256 + 16 : 0;
  const int W = mask_words(g);
  const int cus = ws.prop.multiProcessorCount;
  /* The prefix table of the segment counts (one per workgroup) lives in
   * LDS.  Size the grid from the occupancy without it, then add the table
   * for that grid; occupancy can only drop, so the table stays large
   * enough for the final grid. */
  const bool mass_order = cfg->order == NR_ORDER_MASS;
  KernelGetter getter = mode == 1                       ? var.level_log
                        : mode == 2                     ? var.level_learn
                        : cfg->rule == NR_RULE_FORKBEST || cfg->rule == NR_RULE_PIVOT
                            ? var.level_fork
                        : mass_order                    ? var.level_mass
                                                        : var.level;
  if (!getter) return NR_ERR_UNSUPPORTED;
  const void* kfun = (const void*)getter();
  /* the level kernel's LDS class counts: class rounds and the learned order */
  const int hist_words = (mass_order || mode == 2) ? kClasses : 0;
  int block =
      choose_block(var, g, W, cfg->block_size, 4 * (1027 + 256 + hist_words) + tile_bytes);
  if (block <= 0) return NR_ERR_UNSUPPORTED;
  Layout lay0 = layout_for(var, g, W, block, 4 * (block + hist_words) + tile_bytes);
  if ((rc = ws.allow_lds(kfun, 64 * 1024))) return rc;
  int per_cu0 = 0;
  if ((rc = ws.blocks_per_cu(kfun, block, lay0.total_bytes, &per_cu0))) return rc;
  if (per_cu0 < 1) return NR_ERR_UNSUPPORTED;
  int grid = cfg->grid_blocks > 0 ? cfg->grid_blocks : per_cu0 * cus;
  /* k_ramp_levels keeps 4 + 84 + kRampMax + 1 words there whatever the grid */
  const int extra = 4 * std::max(grid + 3 + block + hist_words, 4 + 84 + kRampMax + 1) + tile_bytes;
  Layout lay = layout_for(var, g, W, block, extra);
  if (lay.total_bytes > 64 * 1024) return NR_ERR_UNSUPPORTED;
  int per_cu = 0;
  if ((rc = ws.blocks_per_cu(kfun, block, lay.total_bytes, &per_cu))) return rc;
  if (per_cu < 1) return NR_ERR_UNSUPPORTED;
  if (cfg->grid_blocks <= 0) grid = std::min(grid, per_cu * cus);
  const u64 n_slots = (u64)grid; /* partial sums and counters per workgroup */
  const u64 cap = cfg->queue_capacity > 0 ? (u64)cfg->queue_capacity : (1ull << 20);
  const bool mass = cfg->order == NR_ORDER_MASS || mode == 2;
  /* class rounds trim the pool to `cap` between rounds, and a round's
   * output (the children of the classes it expands plus the items it
   * carries) lands in buffers of four times that when the device has the
   * room (twice otherwise): near capacity a round can then expand more of
   * the pool and carry less of it (grid 10x10, plain rule: rounds expanding
   * 10-20 M items carried 60-78 M with twice the capacity) */
  u64 phys = cap;
  if (mass) {
    const int ns_est = var.lean || var.nw ? 0 : var.ns_bytes;
    const size_t item = (size_t)(2 * W + 1) * 8 + (size_t)g->n * ns_est + 4;
    size_t avail = 0;
    if (ws.fixed) {
      avail = ws.arena_size > ws.arena_used ? ws.arena_size - ws.arena_used : 0;
    } else {
      size_t free_b = 0, total_b = 0;
      NR_HIP(hipMemGetInfo(&free_b, &total_b));
      avail = free_b + ws.q_bytes; /* the queue buffers are given back first */
    }
    const size_t need4 = 2 * 4 * (size_t)(cap + (u64)grid) * item;
    const size_t slack = ws.fixed ? (1ull << 20) : (64ull << 20); /* tables, slots */
    phys = need4 + slack <= avail ? 4 * cap : 2 * cap;
  }
  /* capacity is split evenly over the workgroup segments */
  const u64 seg_cap64 = std::max<u64>(1, (phys + grid - 1) / grid);
  if (seg_cap64 * (u64)grid >= (1ull << 32)) return NR_ERR_ARG;
  const u32 seg_cap = (u32)seg_cap64;
  const u64 slots = seg_cap64 * (u64)grid;

  LevelArgs a;
  std::memset(&a, 0, sizeof a);
  if ((rc = upload_tables(ws, g, W, a.g))) return rc;
  a.g.period = cfg->sweep_period > 0 ? cfg->sweep_period : 2;
  a.g.pivot = cfg->rule == NR_RULE_PIVOT ? 1 : 0;
  a.lay = lay;
  a.seg_cap = seg_cap;
  a.n_seg = grid;
  a.stride = slots;
  a.eps = cfg->truncation_eps > 0 ? cfg->truncation_eps : 0.0;
  /* Under a time budget (and no checkpoint to keep whole levels for) the
   * kernels compare the GPU wall clock with the deadline before each batch
   * of items, so a long level cannot overrun the budget; the host folds what
   * is left once it sees the deadline. */
  a.budget_ticks = 0;
  if (cfg->time_budget_s > 0 && !(x && x->save)) {
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ws.device) == hipSuccess &&
        khz > 0)
      a.budget_ticks = std::max<u64>(1, (u64)(cfg->time_budget_s * 1e3 * (double)khz));
  }
  /* SoA buffer: known[W], up[W], mass, rows[n] (each rows plane holds
   * node sets of ns_bytes) */
  const int item_ns = var.lean ? 0 : var.ns_bytes; /* lean items carry no rows */
  const size_t rows_bytes = ((size_t)g->n * slots * item_ns + 255) / 256 * 256;
  const size_t tag_bytes = mode ? (slots * 4 + 255) / 256 * 256 : 0;
  const size_t per_buf = (size_t)(2 * W + 1) * slots * 8 + rows_bytes + tag_bytes;
  if ((rc = ws.grow(&ws.q, &ws.q_bytes, 2 * per_buf))) return rc;
  for (int b = 0; b < 2; ++b) {
    u64* base = (u64*)((unsigned char*)ws.q + b * per_buf);
    a.buf[b].known = base;
    a.buf[b].up = base + (size_t)W * slots;
    a.buf[b].mass = (double*)(base + (size_t)2 * W * slots);
    a.buf[b].rows = (void*)(base + (size_t)(2 * W + 1) * slots);
    a.tag[b] = mode ? (u32*)((unsigned char*)a.buf[b].rows + rows_bytes) : nullptr;
  }
  if (mode == 1) {
    a.log_feat = x->log_feat;
    a.log_parent = x->log_parent;
    a.log_level = x->log_level;
    a.log_count = x->log_count;
    a.log_cap = x->log_cap;
    a.log_mass = x->log_mass;
    a.log_succ = x->log_succ;
  }
  if (mode == 2) a.mlp = x->mlp;
  if (x) {
    a.chooser = x->chooser;
    a.log_choice = x->log_choice;
  }
  /* per-workgroup slots: 3 DD + C_N counters; segment counts and settled
   * children per buffer; class histograms */
  const size_t meta_bytes = n_slots * (3 * sizeof(DD) + C_N * sizeof(u64)) +
                            sizeof(u32) * (4 * (size_t)grid + 8) +
                            sizeof(u32) * (size_t)grid * kClasses;
  if ((rc = ws.grow(&ws.meta, &ws.meta_bytes, meta_bytes))) return rc;
  {
    unsigned char* m = (unsigned char*)ws.meta;
    a.partR = (DD*)m;
    a.partE = a.partR + n_slots;
    a.partD = a.partE + n_slots;
    a.cnt = (u64*)(a.partD + n_slots);
    u32* q = (u32*)(a.cnt + n_slots * C_N);
    a.seg_count[0] = q;
    a.seg_count[1] = q + grid;
    a.seg_settle[0] = q + 2 * grid;
    a.seg_settle[1] = q + 3 * grid;
    a.hist = q + 4 * grid + 8;
  }
  a.ctl = ws.ctl;
  a.hmirror = ws.dmirror; /* null in no-allocation mode */
  a.seq = ++ws.solves;
  /* until this solve's first launch writes it, the level word reads as
   * unfinished: a chunk boundary never sees the last solve's empty level */
  if (ws.hmirror) {
    ws.hmirror[0] = ~0ull;
    ws.hmirror[kMirSeq] = 0;
  }
  a.mass_mode = mass ? 1 : 0;
  a.thr_class = kClasses;
  a.drop_class = kClasses;
  a.fin_on_empty = (!mass && ws.hmirror) ? 1 : 0;
  /* level order (MODE 0 kernels) outside checkpoint chains: children below
   * the truncation are counted when created, not queued, and in the row form
   * of the path rule with two terminals on a simple graph so are dead
   * children (a bridge hop); the counts, levels and widest level are those
   * of queueing them.  Under a time budget a settled child is resolved (ESP
   * or D) where a queued one could be folded into ESP at the deadline. */
  a.settle = 0;
  if (mode == 0 && !mass && cfg->rule != NR_RULE_FORKBEST && cfg->rule != NR_RULE_PIVOT &&
      !(x && (x->save || x->resume)) && !(cfg->flags & NR_FLAG_QUEUE_ALL)) {
    a.settle = kSettleTrunc;
    if (var.settle && cfg->rule == NR_RULE_PATH && g->terms.empty() && g->simple)
      a.settle |= kSettleDead;
  }

  hipStream_t st = ws.stream;
  const nr_checkpoint* ck = x ? x->resume : nullptr;
  /* static prefix partition: the seeds enter as a resumed solve's pending
   * items, with no sums or counts yet */
  std::unique_ptr<nr_checkpoint> seeds;
  if (!ck && cfg->prefix_edges > 0) {
    if ((rc = prefix_seeds(g, W, cfg, seeds))) return rc;
    ck = seeds.get();
  }
  DevMem seedmem;
  if (ck) { /* resume: pending items spread over the segments of buffer 0 */
    if (ck->count > seg_cap64 * (u64)grid) return NR_ERR_ARG; /* capacity below the checkpoint */
    SeedMeta sm;
    std::memset(&sm, 0, sizeof sm);
    for (int k = 0; k < 3; ++k) {
      sm.sum[k].hi = ck->sum_hi[k];
      sm.sum[k].lo = ck->sum_lo[k];
    }
    for (int j = 0; j < C_N; ++j) sm.counts[j] = ck->counts[j];
    sm.iterations = mass ? 0 : ck->iterations;
    sm.max_width = mass ? 0 : ck->max_width;
    sm.count = ck->count;
    SeedMeta* dmeta;
    u64 *dk, *du;
    double* dm;
    if ((rc = seedmem.alloc(&dmeta, 1)) || (rc = seedmem.alloc(&dk, ck->count * W)) ||
        (rc = seedmem.alloc(&du, ck->count * W)) || (rc = seedmem.alloc(&dm, ck->count)))
      return rc;
    NR_HIP(hipMemcpyAsync(dmeta, &sm, sizeof sm, hipMemcpyHostToDevice, st));
    if (ck->count) {
      NR_HIP(hipMemcpyAsync(dk, ck->known.data(), 8 * ck->count * W, hipMemcpyHostToDevice, st));
      NR_HIP(hipMemcpyAsync(du, ck->up.data(), 8 * ck->count * W, hipMemcpyHostToDevice, st));
      NR_HIP(hipMemcpyAsync(dm, ck->mass.data(), 8 * ck->count, hipMemcpyHostToDevice, st));
    }
    Root root;
    std::memset(&root, 0, sizeof root);
    int Wi = W, ni = g->n, nsb = item_ns;
    u64 nw = n_slots;
    const SeedMeta* dmc = dmeta;
    void* args[] = {&a, &root, &Wi, &ni, &nsb, &nw, &dmc};
    NR_HIP(hipLaunchKernel((const void*)&k_init, dim3(64), dim3(256), args, 0, st));
    if (ck->count) {
      const u64* dkc = dk;
      const u64* duc = du;
      const double* dmcc = dm;
      const u32* dt = nullptr;
      u64 cnt = ck->count;
      void* cargs[] = {&a, &Wi, &dkc, &duc, &dmcc, &dt, &cnt};
      NR_HIP(hipLaunchKernel((const void*)&k_seed_copy, dim3(1024), dim3(256), cargs, 0, st));
      if (var.seed_rows) {
        const void* kseed = (const void*)var.seed_rows();
        const Layout ls = layout_for(var, g, W, block);
        NR_HIP(hipFuncSetAttribute(kseed, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   64 * 1024));
        void* rargs[] = {&a, &cnt};
        NR_HIP(hipLaunchKernel(kseed, dim3((unsigned)((cnt + block - 1) / block)), dim3(block),
                               rargs, ls.total_bytes, st));
      }
    }
  }
  /* a fresh solve: the root in buffer 0, segment 0 */
  auto init_root = [&]() -> int {
    Root root = root_item(g, W);
    int Wi = W, ni = g->n, nsb = item_ns;
    u64 nw = n_slots;
    const SeedMeta* none = nullptr;
    void* args[] = {&a, &root, &Wi, &ni, &nsb, &nw, &none};
    NR_HIP(hipLaunchKernel((const void*)&k_init, dim3(64), dim3(256), args, 0, st));
    return NR_OK;
  };
  if (!ck && (rc = init_root())) return rc;
  nr_checkpoint** save = x ? x->save : nullptr;
  int pend_parity = -1; /* buffer holding the pending items to keep */
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  if (!ws.fixed && ((rc = ws.event(0, &ev_start)) || (rc = ws.event(1, &ev_stop)))) return rc;
  const size_t max_timed = (cfg->flags & NR_FLAG_TIME_LAUNCHES) && !ws.fixed ? 2048 : 0;

  auto t_begin = std::chrono::steady_clock::now();
  const bool has_budget = cfg->time_budget_s > 0;
  if (ev_start) NR_HIP(hipEventRecord(ev_start, st));
  int level = 0, chunk = has_budget ? 1 : 8;
  bool finished = false, deadline = false;
  /* the queue was seen empty at a chunk boundary (ev_chunk): the launch that
   * found it empty has computed the totals, so the launches queued after it
   * (empty, and leaving the mirror alone) need not be waited for */
  hipEvent_t ev_done = nullptr;
  long long timed = 0, rounds = 0;
  u64 max_pool = 1;
  auto launch_level = [&](int parity, int lv) -> int {
    hipEvent_t kb = nullptr, ke = nullptr;
    if ((size_t)timed < max_timed) {
      int r = ws.event(2 + 2 * timed, &kb);
      if (!r) r = ws.event(3 + 2 * timed, &ke);
      if (r) return r;
      NR_HIP(hipEventRecord(kb, st));
    }
    void* args[] = {&a, (void*)&parity, (void*)&lv};
    NR_HIP(hipLaunchKernel(kfun, dim3(grid), dim3(block), args, lay.total_bytes, st));
    if (kb) {
      NR_HIP(hipEventRecord(ke, st));
      ++timed;
    }
    return NR_OK;
  };
  auto expired = [&]() {
    return has_budget &&
           std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count() >=
               cfg->time_budget_s;
  };
  auto budget_chunk = [&](std::chrono::steady_clock::time_point c0, int launched, int want) {
    const auto now = std::chrono::steady_clock::now();
    const double per = std::chrono::duration<double>(now - c0).count() / std::max(launched, 1);
    const double left =
        cfg->time_budget_s - std::chrono::duration<double>(now - t_begin).count();
    const int fit = per > 0 ? (int)(0.25 * left / per) : want;
    return std::max(1, std::min(want, fit));
  };
  auto fold = [&](int parity) -> int {
    void* args[] = {&a, (void*)&parity};
    NR_HIP(hipLaunchKernel((const void*)&k_fold, dim3(grid), dim3(block), args,
                           4 * (grid + 2 + block), st));
    return NR_OK;
  };
  /* the first, small levels in one workgroup (fresh level-order solves of
   * the path rule): a level of at most max_in items, whose children fit
   * segment 0 (k <= n - 1 hops each), never needs a second workgroup; then
   * the ramp's levels, each workgroup on its own slice's subtrees, up to
   * about one item per lane of the grid.  When a ramp region overflows (the
   * level launches report it at the first chunk boundary) the solve starts
   * again without the ramp, so counts never depend on it. */
  const bool small_ok = var.small && !ck && mode == 0 && !mass &&
                        (cfg->rule == NR_RULE_PATH || cfg->rule == NR_RULE_ALTERNATE);
  bool ramp_used = false;
  int ramp_restarts = 0;
  auto launch_small = [&](bool ramp) -> int {
    const u32 per = (u32)std::max(1, g->n - 1);
    const u32 max_in = std::min<u32>(256u, std::min<u32>((u32)block, seg_cap / per));
    if (max_in < 1) return NR_OK;
    auto timed_launch = [&](const void* k, unsigned blocks, void** args) -> int {
      int r = ws.allow_lds(k, 64 * 1024);
      if (r) return r;
      hipEvent_t kb = nullptr, ke = nullptr;
      if ((size_t)timed < max_timed) {
        if ((r = ws.event(2 + 2 * timed, &kb)) || (r = ws.event(3 + 2 * timed, &ke))) return r;
        NR_HIP(hipEventRecord(kb, st));
      }
      NR_HIP(hipLaunchKernel(k, dim3(blocks), dim3(block), args, lay.total_bytes, st));
      if (kb) {
        NR_HIP(hipEventRecord(ke, st));
        ++timed;
      }
      return NR_OK;
    };
    u32 mi = max_in;
    int ml = 64;
    void* sargs[] = {&a, &mi, &ml};
    int r = timed_launch((const void*)var.small(), 1, sargs);
    if (r || !ramp || !var.ramp || grid < 2 || cfg->ramp_items < 0) return r;
    /* four workgroups per CU, each running its own subtrees level after
     * level with the segments 1.. of the grid split evenly; by default the
     * ramp ends near 1024 items per CU (MI355X: K10 0.745-0.754 ms per
     * solve against 0.762 ms with one workgroup per CU and 512 items, 0.80
     * ms without the ramp; the alternating rule on K15 at 1e-12 2.62 against
     * 2.70 ms; K25 at 1e-12 within 0.5%) */
    const int groups = std::min(grid - 1, 4 * cus);
    int span = (grid - 1) / groups;
    /* by default the last ramp level's children fit the workgroups'
     * regions with room to spare (a region that fills restarts the solve
     * without the ramp); a requested target is taken as it is */
    u64 target64 = cfg->ramp_items > 0 ? (u64)cfg->ramp_items
                                       : std::min<u64>(1024ull * (u64)cus,
                                                       (u64)groups * (u64)span * seg_cap / 4);
    u32 target = (u32)std::min<u64>(target64, 0xffffffffull);
    void* rargs[] = {&a, &span, &target};
    if ((r = timed_launch((const void*)var.ramp(), (unsigned)groups, rargs))) return r;
    ramp_used = true;
    return NR_OK;
  };
  if (ws.hmirror) ws.hmirror[kMirRamp] = 0;
  if (small_ok && (rc = launch_small(true))) return rc;
  if (!mass) {
    /* level order: chunks of back-to-back level launches; under a time
     * budget a chunk is sized from the measured time per launch so the
     * deadline is checked at least every quarter of the remaining budget */
    /* without a budget or a checkpoint to honour, the next chunk's first two
     * launches are queued before the host waits for this chunk, so the GPU
     * keeps working while the host decides (at most two launches run past
     * the last level); an empty level in the mirror, whichever launch wrote
     * it, means the queue is empty for good */
    const int ahead = (!has_budget && !save && !ws.fixed) ? 2 : 0;
    hipEvent_t ev_chunk = nullptr;
    if (!ws.fixed && (rc = ws.event(2 + 2 * max_timed, &ev_chunk))) return rc;
    int queued = 0; /* launches of the coming chunk already queued */
    while (!finished) {
      const auto c0 = std::chrono::steady_clock::now();
      const int launched = chunk;
      for (int c = queued; c < chunk; ++c, ++level)
        if ((rc = launch_level(level & 1, level))) return rc;
      u64 w = 0, max_w = 0;
      bool ramp_over = false;
      if (ev_chunk) {
        NR_HIP(hipEventRecord(ev_chunk, st));
        for (queued = 0; queued < ahead; ++queued, ++level)
          if ((rc = launch_level(level & 1, level))) return rc;
        NR_HIP(hipEventSynchronize(ev_chunk)); /* its last launch wrote ws.hmirror */
        w = ws.hmirror[0];
        max_w = ws.hmirror[2];
        ramp_over = ws.hmirror[kMirRamp] != 0;
        if (w == 0) ev_done = ev_chunk;
      } else { /* no-allocation mode: the counters by a plain copy */
        NR_HIP(hipMemcpyAsync(ws.hctl, a.ctl, kCtlCopyBytes, hipMemcpyDeviceToHost, st));
        NR_HIP(hipStreamSynchronize(st));
        w = ws.hctl->last_items;
        max_w = ws.hctl->max_width;
        ramp_over = ws.hctl->ramp_over != 0;
      }
      if (ramp_used && ramp_over) { /* a ramp region filled: again without the ramp */
        NR_HIP(hipStreamSynchronize(st));
        if (ws.hmirror) { /* nothing of the discarded run may look like a result */
          ws.hmirror[0] = ~0ull;
          ws.hmirror[kMirSeq] = 0;
        }
        if ((rc = init_root()) || (rc = launch_small(false))) return rc;
        ramp_used = false;
        ++ramp_restarts;
        timed = 0; /* kernel_ms and kernel_launches: the solve that counts */
        level = 0;
        queued = 0;
        chunk = has_budget ? 1 : 8;
        continue;
      }
      if (w == 0) finished = true;
      /* past the widest level: shorter chunks as the levels thin out, so
       * few launches run after the last level */
      const bool shrinking = w < max_w;
      if (!finished && expired()) {
        if (save)
          pend_parity = level & 1;
        else if ((rc = fold(level & 1)))
          return rc;
        deadline = true;
        finished = true;
      }
      /* at most 8 launches a chunk: a chunk that runs past the last level
       * leaves empty launches behind it */
      chunk = std::min(chunk * 2, 8);
      if (shrinking) chunk = std::min(chunk, 4);
      if (shrinking && 16 * w < max_w) chunk = std::min(chunk, 2);
      if (has_budget && !finished) chunk = budget_chunk(c0, launched, chunk);
    }
  } else {
    /* class rounds: each launch expands every queued item of the heaviest
     * classes (at least `target` items, or all) and carries the rest; when
     * the output would exceed the capacity the lightest classes are given
     * up first. */
    const bool trace = std::getenv("NR_TRACE_ROUNDS") != nullptr;
    std::vector<u64> hist(kClasses, 0);
    u64 pool = 1;
    long long launches_here = 0;
    if (ck) {
      for (u64 i = 0; i < ck->count; ++i) hist[mass_class(ck->mass[i])] += 1;
      pool = ck->count;
      rounds = (long long)ck->iterations; /* reported rounds continue the chain */
      max_pool = std::max<u64>(ck->max_width, ck->count);
    } else {
      hist[mass_class(1.0)] = 1;
    }
    double ratio = 1.0; /* queued children per expanded item, running estimate */
    /* a round expands at least a quarter of the pool (heaviest classes
     * first): carrying the rest then costs at most three times the useful
     * work, against ~pool/1M times with a fixed-size round */
    const u64 target0 = std::max<u64>(1ull << 20, (u64)grid * block);
    /* a round that leaves much of the pool unexpanded is split first
     * (k_class_split), so its waves expand only; learned order classes by
     * the items' tags */
    const void* split_fn = nullptr;
    if (mode == 0 || mode == 2) {
      const KernelGetter sg = split_kernel(W, var.lean ? 0 : var.ns_bytes, mode == 2);
      if (sg) {
        split_fn = (const void*)sg();
        if ((rc = ws.allow_lds(split_fn, 64 * 1024))) return rc;
      }
    }
    while (pool > 0) {
      const u64 target = std::max<u64>(target0, pool / 4);
      /* trim: when the pool exceeds the capacity, give up its lightest
       * classes (at least the heaviest class is kept) */
      int drop = kClasses;
      u64 keep = 0;
      for (int cl = 0; cl < kClasses; ++cl) {
        if (keep + hist[cl] > cap && keep > 0) {
          drop = cl;
          break;
        }
        keep += hist[cl];
      }
      /* expand the heaviest kept classes: at least `target` items (or all),
       * while the expected output fits the physical headroom */
      u64 cum = 0;
      int thr = -1;
      for (int cl = 0; cl < drop; ++cl) {
        if (!hist[cl]) continue;
        const double out_if = ratio * (double)(cum + hist[cl]) + (double)(keep - cum - hist[cl]);
        if (cum > 0 && (cum >= target || out_if > (double)phys)) break;
        cum += hist[cl];
        thr = cl;
      }
      const u64 carried = keep - cum;
      a.thr_class = thr;
      a.drop_class = drop;
      /* buffer parity counts this call's launches (seeds are in buffer 0) */
      const int parity = (int)(launches_here & 1);
      /* a quarter or more of the pool carried or given up: the split pass
       * leaves the round only items to expand (a wave whose lanes mostly
       * carry would run its few expansions at the cost of a full wave) */
      const bool split = split_fn && (pool - cum) * 4 >= pool;
      if (split) {
        int sp_parity = parity;
        void* sargs[] = {&a, (void*)&sp_parity};
        NR_HIP(hipLaunchKernel(split_fn, dim3(grid), dim3(block), sargs, lay.total_bytes, st));
        a.append = 1;
        a.thr_class = kClasses - 1;
        a.drop_class = kClasses;
      }
      rc = launch_level(parity, (int)rounds);
      a.append = 0;
      if (rc) return rc;
      void* hargs[] = {&a};
      NR_HIP(hipLaunchKernel((const void*)&k_hist_total, dim3(kClasses), dim3(256), hargs, 0, st));
      NR_HIP(hipMemcpyAsync(ws.hctl, a.ctl, kCtlCopyBytes, hipMemcpyDeviceToHost, st));
      NR_HIP(hipStreamSynchronize(st));
      ++rounds;
      ++launches_here;
      pool = 0;
      for (int cl = 0; cl < kClasses; ++cl) {
        hist[cl] = ws.hctl->hist[cl];
        pool += hist[cl];
      }
      max_pool = std::max(max_pool, pool);
      if (trace)
        std::fprintf(stderr, "round %lld thr %d drop %d expand %llu carried %llu -> pool %llu\n",
                     rounds, thr, drop, (unsigned long long)cum, (unsigned long long)carried,
                     (unsigned long long)pool);
      if (cum > 0) {
        const double kids = pool > carried ? (double)(pool - carried) : 0.0;
        ratio = std::max(1.0, 0.5 * ratio + 0.5 * kids / (double)cum);
      }
      if (pool > 0 && expired()) {
        if (save)
          pend_parity = (int)(launches_here & 1);
        else if ((rc = fold((int)(launches_here & 1))))
          return rc;
        deadline = true;
        break;
      }
    }
  }
  std::unique_ptr<nr_checkpoint> kept;
  if (pend_parity >= 0) { /* keep the pending items instead of folding them */
    std::vector<u32> segs(grid);
    NR_HIP(hipMemcpyAsync(segs.data(), a.seg_count[pend_parity], 4 * (size_t)grid,
                          hipMemcpyDeviceToHost, st));
    NR_HIP(hipStreamSynchronize(st));
    u64 count = 0;
    for (u32 c2 : segs) count += c2;
    kept.reset(new (std::nothrow) nr_checkpoint);
    if (!kept) return NR_ERR_NOMEM;
    kept->count = count;
    kept->known.resize(count * W);
    kept->up.resize(count * W);
    kept->mass.resize(count);
    if (count) {
      DevMem gm;
      u64 *gk, *gu;
      double* gmass;
      if ((rc = gm.alloc(&gk, count * W)) || (rc = gm.alloc(&gu, count * W)) ||
          (rc = gm.alloc(&gmass, count)))
        return rc;
      int par = pend_parity, Wi = W;
      u32* gt = nullptr;
      void* args[] = {&a, &par, &Wi, &gk, &gu, &gmass, &gt};
      NR_HIP(hipLaunchKernel((const void*)&k_gather, dim3(grid), dim3(block), args,
                             4 * (grid + 2 + block), st));
      NR_HIP(hipMemcpyAsync(kept->known.data(), gk, 8 * count * W, hipMemcpyDeviceToHost, st));
      NR_HIP(hipMemcpyAsync(kept->up.data(), gu, 8 * count * W, hipMemcpyDeviceToHost, st));
      NR_HIP(hipMemcpyAsync(kept->mass.data(), gmass, 8 * count, hipMemcpyDeviceToHost, st));
      NR_HIP(hipStreamSynchronize(st));
    }
  }
  /* a queue that ran empty had its totals computed by the first empty level
   * launch (fin_on_empty); otherwise (a fold, kept items, no mirror) by
   * k_finalize */
  bool have_totals = false;
  if (a.fin_on_empty && !deadline && pend_parity < 0) {
    if (ev_done && max_timed == 0 && !ramp_restarts &&
        __atomic_load_n(&ws.hmirror[kMirSeq], __ATOMIC_ACQUIRE) == a.seq) {
      ev_stop = ev_done; /* the solve's GPU time ends there */
      have_totals = true;
    } else {
      if (ev_stop) NR_HIP(hipEventRecord(ev_stop, st));
      NR_HIP(hipStreamSynchronize(st));
      have_totals = __atomic_load_n(&ws.hmirror[kMirSeq], __ATOMIC_ACQUIRE) == a.seq;
    }
  }
  if (!have_totals) {
    u64 nw = n_slots;
    void* args[] = {&a, &nw};
    NR_HIP(hipLaunchKernel((const void*)&k_finalize, dim3(finalize_groups(n_slots)), dim3(256),
                           args, 0, st));
    if (ev_stop) NR_HIP(hipEventRecord(ev_stop, st));
    NR_HIP(hipStreamSynchronize(st));
  }
  float ms = 0.f;
  if (ev_stop)
    NR_HIP(hipEventElapsedTime(&ms, ev_start, ev_stop));
  else /* no-allocation mode: host clock around the synchronised solve */
    ms = (float)(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                           t_begin)
                     .count());
  /* the result as k_finalize mirrored it into pinned host memory; should
   * its stamp not be there yet, the device copy is read instead */
  LevelCtl& h = *ws.hctl;
  if (!ws.hmirror || __atomic_load_n(&ws.hmirror[kMirSeq], __ATOMIC_ACQUIRE) != a.seq) {
    NR_HIP(hipMemcpy(ws.hctl, a.ctl, kCtlCopyBytes, hipMemcpyDeviceToHost));
  } else {
    const u64* m = ws.hmirror;
    h.iterations = m[1];
    h.max_width = m[2];
    std::memcpy(h.sum, m + kMirSum, sizeof h.sum);
    std::memcpy(h.sum_lo, m + kMirSumLo, sizeof h.sum_lo);
    for (int j = 0; j < C_N; ++j) h.counts[j] = m[kMirCounts + j];
  }
  /* summed time of every expand launch, including the few empty launches
   * enqueued past the last level of the final chunk */
  double kms = 0.0;
  for (long long i = 0; i < timed; ++i) {
    float x = 0.f;
    NR_HIP(hipEventElapsedTime(&x, ws.events[2 + 2 * i], ws.events[3 + 2 * i]));
    kms += x;
  }
  if (h.counts[C_ERR]) return NR_ERR_DEVICE;

  std::memset(out, 0, sizeof(*out));
  out->R = h.sum[0];
  out->ESP = h.sum[1];
  out->D = h.sum[2];
  out->paths = h.counts[C_PATHS];
  out->nodes = h.counts[C_NODES];
  out->dead = h.counts[C_DEAD];
  out->truncated = h.counts[C_TRUNC];
  out->overflowed = h.counts[C_OVER];
  out->folded = h.counts[C_FOLD];
  out->iterations = mass ? rounds : (long long)h.iterations;
  out->max_queue_items = mass ? (long long)max_pool : (long long)h.max_width;
  out->item_bytes = 16LL * W + 8;
  out->elapsed_ms = ms;
  out->kernel_ms = kms;
  out->kernel_launches = timed;
  out->deadline_hit = deadline ? 1 : 0;
  out->overflowed_flag = h.counts[C_OVER] ? 1 : 0;
  out->grid_blocks = grid;
  out->block_size = block;
  out->lds_bytes = lay.total_bytes;
  out->blocks_per_cu = per_cu;
  out->compute_units = cus;
  out->rerun = ramp_restarts ? 1 : 0;
  if (save) *save = nullptr;
  if (kept) { /* the bracket stays certified: pending mass counts in ESP */
    kept->n = g->n;
    kept->e = g->e;
    kept->w = W;
    kept->src = g->src;
    kept->dst = g->dst;
    kept->rule = cfg->rule;
    kept->period = cfg->rule == NR_RULE_ALTERNATE ? a.g.period : 0;
    kept->order = mass ? NR_ORDER_MASS : NR_ORDER_LEVEL;
    kept->eps = a.eps;
    kept->graph_hash = graph_hash(g);
    for (int k = 0; k < 3; ++k) {
      kept->sum_hi[k] = h.sum[k];
      kept->sum_lo[k] = h.sum_lo[k];
    }
    for (int j = 0; j < C_N; ++j) kept->counts[j] = h.counts[j];
    kept->iterations = mass ? (u64)rounds : h.iterations;
    kept->max_width = mass ? max_pool : h.max_width;
    double ps = 0.0, pc = 0.0;
    for (double m : kept->mass) {
      const double t = ps + m;
      pc += std::fabs(ps) >= std::fabs(m) ? (ps - t) + m : (m - t) + ps;
      ps = t;
    }
    out->ESP = h.sum[1] + (ps + pc);
    *save = kept.release();
  }
  return NR_OK;
}

/* Batched solve.  All instances share one level-synchronous queue; an item
 * is its planes, its mass and a tag (instance id, tree depth); adjacency is
 * derived from the planes and the instance's tables in HBM, and
 * per-instance sums are exact fixed-point integers, added once per wave and
 * instance.  Children that do not fit their workgroup's segment go to a
 * spill stack; between two levels a refill launch tops the next level up to
 * a quarter of the capacity, first from the top of the stack (the deepest
 * items, so the stack stays shallow), then with new instance roots.  Nothing
 * is dropped unless the stack itself is full; an instance that overflowed
 * then is solved again alone, so its counts are those of an unbounded queue
 * unless it overflows on its own. */
/* The HIP graphs of one batch pass's chunks (8 and 16 level pairs). */
struct ChunkGraphs {
  hipGraphExec_t g8 = nullptr, g16 = nullptr;
  ~ChunkGraphs() {
    if (g8) (void)hipGraphExecDestroy(g8);
    if (g16) (void)hipGraphExecDestroy(g16);
  }
};

int solve_batch(Workspace& ws, const nr_graph* const* gs, int count, const nr_config* cfg,
                nr_result* outs) {
  int Wmax = 1, Nmax = 1;
  for (int i = 0; i < count; ++i) {
    const nr_graph* g = gs[i];
    if (!g) return NR_ERR_ARG;
    if (!g->symmetric || g->n > 64 || mask_words(g) > 8) return NR_ERR_UNSUPPORTED;
    if ((u32)g->e >= kMaxDepth) return NR_ERR_UNSUPPORTED; /* depth <= E fits the tag */
    if (!g->terms.empty()) return NR_ERR_UNSUPPORTED;     /* two-terminal instances only */
    Wmax = std::max(Wmax, mask_words(g));
    Nmax = std::max(Nmax, g->n);
  }
  Variant var = Nmax <= 32 ? variant_for_w<u32>(Wmax) : variant_for_w<u64>(Wmax);
  if (!var.batch) return NR_ERR_UNSUPPORTED;
  const int W = Wmax;
  const void* kfun = (const void*)var.batch();
  const int cus = ws.prop.multiProcessorCount;
  const int lds_max = 64 * 1024;
  NR_HIP(hipFuncSetAttribute(kfun, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max));
  /* block: requested, else the largest of 256/128/64 whose per-lane slices
   * leave room for 2 workgroups per CU */
  int block = cfg->block_size > 0 ? cfg->block_size : 256;
  if (block % 64 || block > 256) return NR_ERR_ARG;
  auto lay_for = [&](int blk, int grid) { /* level bytes + hop stack per lane */
    return make_layout_csr(Nmax, blk, 4 * (grid + 3 + blk));
  };
  if (cfg->block_size <= 0)
    while (block > 64 && lay_for(block, 4 * cus).total_bytes > lds_max) block /= 2;
  int per_cu = 0;
  NR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfun, block,
                                                      lay_for(block, 4 * cus).total_bytes));
  if (per_cu < 1) return NR_ERR_UNSUPPORTED;
  const int grid = cfg->grid_blocks > 0 ? cfg->grid_blocks : per_cu * cus;
  const Layout lay = lay_for(block, grid);
  if (lay.total_bytes > lds_max) return NR_ERR_UNSUPPORTED;
  const int n_seg = grid + 1;
  const u64 cap0 = cfg->queue_capacity > 0 ? (u64)cfg->queue_capacity : (1ull << 20);
  const u64 seg_cap64 = std::max<u64>(64, (cap0 + grid - 1) / grid);
  const u64 cap = seg_cap64 * (u64)grid;
  const u64 fill_cap = std::max<u64>(seg_cap64, cap / 4); /* refill segment */
  const u64 slots = cap + fill_cap;
  if (slots >= (1ull << 32)) return NR_ERR_ARG;
  const u32 seg_cap = (u32)seg_cap64;
  if ((u64)count > kMaxBatch) return NR_ERR_ARG;

  /* one buffer: desc | instance tables | roots | perm | perm2, packed and
   * uploaded once for a given list of graphs */
  int rc;
  hipStream_t st = ws.stream;
  std::vector<std::pair<const nr_graph*, unsigned long long>> key((size_t)count);
  for (int i = 0; i < count; ++i) key[(size_t)i] = {gs[i], gs[i]->stamp};
  if (!(ws.btab && ws.bW == W && ws.bkey == key)) {
    ws.bkey.clear();
    /* one buffer: desc | instance tables | roots | perm | perm2 */
    std::vector<InstDesc> desc(count);
    std::vector<unsigned char> tabs;
    std::vector<u64> roots((size_t)count * 2 * W, 0);
    for (int i = 0; i < count; ++i) {
      const nr_graph* g = gs[i];
      TableBlob tb;
      pack_tables(g, W, tb);
      const size_t base = (tabs.size() + 255) / 256 * 256;
      if (base + tb.bytes.size() >= (1ull << 32)) return NR_ERR_ARG;
      tabs.resize(base + tb.bytes.size(), 0);
      std::memcpy(&tabs[base], tb.bytes.data(), tb.bytes.size());
      InstDesc d;
      std::memset(&d, 0, sizeof d);
      d.n = g->n;
      d.e = g->e;
      d.src = g->src;
      d.dst = g->dst;
      d.off_p = (u32)(base + tb.off_p);
      d.off_q = (u32)(base + tb.off_q);
      d.off_sadj = (u32)(base + tb.off_sadj);
      d.off_inc = (u32)(base + tb.off_inc);
      d.off_ep = (u32)(base + tb.off_ep);
      d.off_deg = (u32)(base + tb.off_deg);
      d.off_eidm = (u32)(base + tb.off_eidm);
      desc[i] = d;
      Root r = root_item(g, mask_words(g));
      u64* dst = &roots[(size_t)i * 2 * W];
      for (int w = 0; w < W; ++w) {
        dst[w] = w < mask_words(g) ? r.known[w] : 0;
        dst[W + w] = w < mask_words(g) ? r.up[w] : 0;
      }
    }
    /* admission order: larger instances first, so the longest subtrees start
     * early and the tail of the stream holds the small ones */
    std::vector<u32> perm(count);
    for (int i = 0; i < count; ++i) perm[i] = (u32)i;
    std::stable_sort(perm.begin(), perm.end(), [&](u32 x, u32 y) {
      return (long long)gs[x]->e * gs[x]->n > (long long)gs[y]->e * gs[y]->n;
    });
    size_t o = 0;
    auto take = [&](size_t bytes) {
      o = (o + 255) / 256 * 256;
      size_t at = o;
      o += bytes;
      return at;
    };
    const size_t o_desc = take(sizeof(InstDesc) * count), o_tabs = take(tabs.size()),
                 o_roots = take(8 * roots.size()), o_perm = take(4 * (size_t)count),
                 o_perm2 = take(4 * (size_t)count);
    std::vector<unsigned char> blob(o, 0);
    std::memcpy(&blob[o_desc], desc.data(), sizeof(InstDesc) * count);
    std::memcpy(&blob[o_tabs], tabs.data(), tabs.size());
    std::memcpy(&blob[o_roots], roots.data(), 8 * roots.size());
    std::memcpy(&blob[o_perm], perm.data(), 4 * perm.size());

    if ((rc = ws.grow(&ws.btab, &ws.btab_bytes, blob.size()))) return rc;
    NR_HIP(hipMemcpyAsync(ws.btab, blob.data(), blob.size(), hipMemcpyHostToDevice, st));
    NR_HIP(hipStreamSynchronize(st)); /* the host blob goes out of scope */
    ws.bkey = key;
    ws.bW = W;
    ws.bperm = perm;
    ws.bo_desc = o_desc;
    ws.bo_tabs = o_tabs;
    ws.bo_roots = o_roots;
    ws.bo_perm = o_perm;
    ws.bo_perm2 = o_perm2;
  }
  const size_t o_desc = ws.bo_desc, o_tabs = ws.bo_tabs, o_roots = ws.bo_roots,
               o_perm = ws.bo_perm, o_perm2 = ws.bo_perm2;
  unsigned char* tb = (unsigned char*)ws.btab;

  BatchArgs a;
  std::memset(&a, 0, sizeof a);
  a.lay = lay;
  a.desc = (const InstDesc*)(tb + o_desc);
  a.tab = tb + o_tabs;
  a.roots = (const u64*)(tb + o_roots);
  a.seg_cap = seg_cap;
  a.n_seg = n_seg;
  a.stride = slots;
  a.eps = cfg->truncation_eps > 0 ? cfg->truncation_eps : 0.0;
  a.spill_cap = 4 * cap; /* the stack holds what the levels cannot */
  const size_t tag_bytes = (slots * 4 + 255) / 256 * 256;
  const size_t per_buf = (size_t)(2 * W + 1) * slots * 8 + tag_bytes;
  const size_t spill_bytes =
      (size_t)(2 * W + 1) * a.spill_cap * 8 + (a.spill_cap * 4 + 255) / 256 * 256;
  if ((rc = ws.grow(&ws.q, &ws.q_bytes, 2 * per_buf + spill_bytes))) return rc;
  for (int b = 0; b < 2; ++b) {
    unsigned char* base = (unsigned char*)ws.q + b * per_buf;
    a.buf[b].known = (u64*)base;
    a.buf[b].up = (u64*)base + (size_t)W * slots;
    a.buf[b].mass = (double*)((u64*)base + (size_t)2 * W * slots);
    a.buf[b].rows = nullptr;
    a.tag[b] = (u32*)((u64*)base + (size_t)(2 * W + 1) * slots);
  }
  {
    unsigned char* base = (unsigned char*)ws.q + 2 * per_buf;
    a.spill.known = (u64*)base;
    a.spill.up = (u64*)base + (size_t)W * a.spill_cap;
    a.spill.mass = (double*)((u64*)base + (size_t)2 * W * a.spill_cap);
    a.spill.rows = nullptr;
    a.spill_tag = (u32*)((u64*)base + (size_t)(2 * W + 1) * a.spill_cap);
  }
  const size_t meta_bytes = sizeof(u64) * (size_t)count * B_N + sizeof(StreamCtl) +
                            sizeof(u32) * (2 * (size_t)n_seg + 8);
  if ((rc = ws.grow(&ws.meta, &ws.meta_bytes, meta_bytes))) return rc;
  a.res = (u64*)ws.meta;
  a.sc = (StreamCtl*)(a.res + (size_t)count * B_N);
  a.seg_count[0] = (u32*)(a.sc + 1);
  a.seg_count[1] = a.seg_count[0] + n_seg;
  a.ctl = ws.ctl;
  const void* kref = (const void*)var.refill();
  const int refill_grid = 2 * cus;

  hipEvent_t ev_start, ev_stop;
  if ((rc = ws.event(0, &ev_start)) || (rc = ws.event(1, &ev_stop))) return rc;
  const bool timing = (cfg->flags & NR_FLAG_TIME_LAUNCHES) != 0;
  double kernel_ms = 0.0;
  for (int c = 0; timing && c < 16; ++c) { /* events 2.. bracket one chunk's launches */
    hipEvent_t e0, e1;
    if ((rc = ws.event(2 + 2 * c, &e0)) || (rc = ws.event(3 + 2 * c, &e1))) return rc;
  }
  auto t_begin = std::chrono::steady_clock::now();
  NR_HIP(hipEventRecord(ev_start, st));
  std::vector<u64> res((size_t)count * B_N, 0);
  std::vector<char> admitted(count, 0), rerun(count, 0);
  bool deadline = false;
  long long launches = 0, max_width = 0;
  std::vector<u32> todo(ws.bperm);
  /* pass 0 streams every instance; an instance that still overflowed (spill
   * stack full) is solved again in a pass of its own */
  for (int pass = 0; !todo.empty() && !deadline; ++pass) {
    const bool alone = pass >= 1;
    a.perm = (const u32*)(tb + (pass ? o_perm2 : o_perm));
    if (pass)
      NR_HIP(hipMemcpyAsync((void*)a.perm, todo.data(), 4 * todo.size(), hipMemcpyHostToDevice,
                            st));
    a.count = (u32)todo.size();
    a.fill_step = (u32)fill_cap;
    /* the queue chunk: refills top the next level up to this many items */
    a.fill_to = cfg->chunk_items > 0 ? (u32)std::min<u64>((u64)cfg->chunk_items, cap)
                                     : (u32)std::max<u64>(cap / 4, 1);
    a.root_step = alone ? 1u : (u32)fill_cap;
    a.roots_when_empty = alone ? 1 : 0;
    {
      void* args[] = {&a};
      NR_HIP(hipLaunchKernel((const void*)&k_init_batch, dim3(64), dim3(256), args, 0, st));
      int out0 = 0;
      void* rargs[] = {&a, &out0};
      NR_HIP(hipLaunchKernel(kref, dim3(refill_grid), dim3(256), rargs, 0, st));
    }
    int level = 0, chunk = 8;
    bool finished = false;
    StreamCtl hsc;
    /* A chunk is a fixed sequence of (expand, refill) launches starting on
     * buffer 0 (chunks have even lengths): without per-launch timing it is
     * captured once per pass and length into a HIP graph and replayed. */
    auto enqueue_chunk = [&](int len) -> int {
      for (int c = 0; c < len; ++c) {
        int parity = c & 1, nxt = parity ^ 1;
        void* args[] = {&a, &parity};
        if (timing) NR_HIP(hipEventRecord(ws.events[2 + 2 * c], st));
        NR_HIP(hipLaunchKernel(kfun, dim3(grid), dim3(block), args, lay.total_bytes, st));
        if (timing) NR_HIP(hipEventRecord(ws.events[3 + 2 * c], st));
        void* rargs[] = {&a, &nxt};
        NR_HIP(hipLaunchKernel(kref, dim3(refill_grid), dim3(256), rargs, 0, st));
      }
      return NR_OK;
    };
    ChunkGraphs graphs;
    bool use_graphs = !timing;
    /* the chunk's graph, captured on first use; false (and eager launches
     * from then on) if the runtime cannot capture or instantiate it */
    auto chunk_graph = [&](int len, hipGraphExec_t& ge) -> bool {
      if (ge) return true;
      if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return false;
      }
      const int erc = enqueue_chunk(len);
      hipGraph_t gr = nullptr;
      const hipError_t ec = hipStreamEndCapture(st, &gr);
      bool ok = erc == NR_OK && ec == hipSuccess && gr;
      if (ok && hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0) != hipSuccess) {
        ge = nullptr;
        ok = false;
      }
      if (gr) (void)hipGraphDestroy(gr);
      if (!ok) (void)hipGetLastError();
      return ok;
    };
    while (!finished) {
      hipGraphExec_t& ge = chunk == 8 ? graphs.g8 : graphs.g16;
      if (use_graphs && !chunk_graph(chunk, ge)) use_graphs = false;
      if (use_graphs) {
        NR_HIP(hipGraphLaunch(ge, st));
      } else if ((rc = enqueue_chunk(chunk))) {
        return rc;
      }
      level += chunk;
      launches += chunk;
      NR_HIP(hipMemcpyAsync(ws.hctl, a.ctl, kCtlCopyBytes, hipMemcpyDeviceToHost, st));
      NR_HIP(hipMemcpyAsync(&hsc, a.sc, sizeof hsc, hipMemcpyDeviceToHost, st));
      NR_HIP(hipStreamSynchronize(st));
      for (int c = 0; timing && c < chunk; ++c) {
        float x = 0.f;
        NR_HIP(hipEventElapsedTime(&x, ws.events[2 + 2 * c], ws.events[3 + 2 * c]));
        kernel_ms += x;
      }
      max_width = std::max<long long>(max_width, (long long)ws.hctl->max_width);
      if (ws.hctl->last_items == 0) finished = true;
      if (!finished && cfg->time_budget_s > 0 &&
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count() >=
              cfg->time_budget_s) {
        int parity = level & 1;
        void* args[] = {&a, &parity};
        NR_HIP(hipLaunchKernel((const void*)&k_fold_batch, dim3(grid), dim3(block), args,
                               4 * (n_seg + 2 + block), st));
        deadline = true;
        finished = true;
      }
      chunk = std::min(chunk * 2, 16);
    }
    const u32 next_seen = hsc.next[(level & 1) ^ 1];
    std::vector<u64> r2((size_t)count * B_N);
    NR_HIP(hipMemcpyAsync(r2.data(), a.res, sizeof(u64) * r2.size(), hipMemcpyDeviceToHost, st));
    NR_HIP(hipStreamSynchronize(st));
    std::vector<u32> again;
    for (u32 j = 0; j < next_seen && j < todo.size(); ++j) {
      const u32 id = todo[j];
      admitted[id] = 1;
      std::memcpy(&res[(size_t)id * B_N], &r2[(size_t)id * B_N], sizeof(u64) * B_N);
      if (pass) rerun[id] = 1;
      if (!alone && !deadline && r2[(size_t)id * B_N + B_OVER]) again.push_back(id);
    }
    todo.swap(again);
  }
  NR_HIP(hipEventRecord(ev_stop, st));
  NR_HIP(hipEventSynchronize(ev_stop));
  float ms = 0.f;
  NR_HIP(hipEventElapsedTime(&ms, ev_start, ev_stop));
  auto fx = [](const u64* w) {
    return (double)w[1] * 2.3283064365386963e-10 + (double)w[0] * 1.2621774483536189e-29;
  };
  for (int i = 0; i < count; ++i) {
    nr_result& o2 = outs[i];
    std::memset(&o2, 0, sizeof o2);
    const u64* r = &res[(size_t)i * B_N];
    if (admitted[i] && r[B_ERR]) return NR_ERR_DEVICE;
    if (admitted[i]) {
      o2.R = fx(&r[B_R]);
      o2.ESP = fx(&r[B_E]);
      o2.D = fx(&r[B_D]);
      o2.nodes = r[B_NODES];
      o2.paths = r[B_PATHS];
      o2.dead = r[B_DEAD];
      o2.truncated = r[B_TRUNC];
      o2.overflowed = r[B_OVER];
      o2.folded = r[B_FOLD];
      o2.iterations = (long long)r[B_LEVELS];
    } else { /* never started before the deadline: the whole mass is the gap */
      o2.ESP = 1.0;
    }
    o2.max_queue_items = max_width;
    o2.item_bytes = 16LL * W + 12;
    o2.elapsed_ms = ms;
    o2.kernel_launches = launches;
    o2.kernel_ms = timing ? kernel_ms : 0.0;
    o2.deadline_hit = deadline ? 1 : 0;
    o2.overflowed_flag = o2.overflowed ? 1 : 0;
    o2.rerun = rerun[i];
    o2.grid_blocks = grid;
    o2.block_size = block;
    o2.lds_bytes = lay.total_bytes;
    o2.blocks_per_cu = per_cu;
    o2.compute_units = cus;
  }
  return NR_OK;
}


int expand_items(Workspace& ws, const nr_graph* g, int rule, int period, int count,
                 const u64* known, const u64* up, const double* mass, int max_children,
                 int* nchild, double* succ, u64* ck, u64* cu, double* cm) {
  Variant var;
  int rc = pick_variant(g, var);
  if (rc) return rc;
  if (rule == NR_RULE_ALTERNATE) {
    if (!g->symmetric || !g->terms.empty()) return NR_ERR_UNSUPPORTED;
    var = g->n <= 64 ? variant_alt(mask_words(g), var.ns_bytes)
                     : variant_alt_csr(mask_words(g), (g->n + 63) / 64);
    if (!var.items) return NR_ERR_UNSUPPORTED;
  } else if (rule != NR_RULE_PATH) {
    return NR_ERR_UNSUPPORTED;
  }
  if (count <= 0) return NR_OK;
  const int W = mask_words(g);
  const int block = choose_block(var, g, W, 0);
  if (block <= 0) return NR_ERR_UNSUPPORTED;
  Layout lay = layout_for(var, g, W, block);
  const void* kfun = (const void*)var.items();
  if (lay.total_bytes > 48 * 1024)
    NR_HIP(hipFuncSetAttribute(kfun, hipFuncAttributeMaxDynamicSharedMemorySize,
                               lay.total_bytes));
  DevMem mem;
  ItemsArgs a;
  std::memset(&a, 0, sizeof a);
  if ((rc = upload_tables(ws, g, W, a.g))) return rc;
  a.g.period = period > 0 ? period : 2;
  a.lay = lay;
  a.count = count;
  a.max_children = std::max(max_children, 1);
  const size_t nc = (size_t)count * a.max_children;
  u64 *dk, *du;
  double* dm;
  if ((rc = mem.alloc(&dk, (size_t)count * W)) || (rc = mem.alloc(&du, (size_t)count * W)) ||
      (rc = mem.alloc(&dm, count)) || (rc = mem.alloc(&a.nchild, count)) ||
      (rc = mem.alloc(&a.succ, count)) || (rc = mem.alloc(&a.ck, nc * W)) ||
      (rc = mem.alloc(&a.cu, nc * W)) || (rc = mem.alloc(&a.cm, nc)))
    return rc;
  hipStream_t st = ws.stream;
  NR_HIP(hipMemcpyAsync(dk, known, sizeof(u64) * count * W, hipMemcpyHostToDevice, st));
  NR_HIP(hipMemcpyAsync(du, up, sizeof(u64) * count * W, hipMemcpyHostToDevice, st));
  NR_HIP(hipMemcpyAsync(dm, mass, sizeof(double) * count, hipMemcpyHostToDevice, st));
  NR_HIP(hipMemsetAsync(a.ck, 0, sizeof(u64) * nc * W, st));
  NR_HIP(hipMemsetAsync(a.cu, 0, sizeof(u64) * nc * W, st));
  NR_HIP(hipMemsetAsync(a.cm, 0, sizeof(double) * nc, st));
  a.known = dk;
  a.up = du;
  a.mass = dm;
  const int grid = (count + block - 1) / block;
  void* args[] = {&a};
  NR_HIP(hipLaunchKernel(kfun, dim3(grid), dim3(block), args, lay.total_bytes, st));
  NR_HIP(hipMemcpyAsync(nchild, a.nchild, sizeof(int) * count, hipMemcpyDeviceToHost, st));
  NR_HIP(hipMemcpyAsync(succ, a.succ, sizeof(double) * count, hipMemcpyDeviceToHost, st));
  if (ck) NR_HIP(hipMemcpyAsync(ck, a.ck, sizeof(u64) * nc * W, hipMemcpyDeviceToHost, st));
  if (cu) NR_HIP(hipMemcpyAsync(cu, a.cu, sizeof(u64) * nc * W, hipMemcpyDeviceToHost, st));
  if (cm) NR_HIP(hipMemcpyAsync(cm, a.cm, sizeof(double) * nc, hipMemcpyDeviceToHost, st));
  NR_HIP(hipStreamSynchronize(st));
  return NR_OK;
}

/* ---- tree log: subtree sizes of every logged child ---- */
__global__ void __launch_bounds__(256) k_log_sizes(u64* size, u64 n) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (u64)gridDim.x * blockDim.x)
    size[i] = 1;
}

/* records are appended launch by launch, so each depth is one contiguous
 * range; first[d] receives its first record */
__global__ void __launch_bounds__(256) k_log_bounds(const u32* level, u64 n, u64* first) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (u64)gridDim.x * blockDim.x)
    if (i == 0 || level[i] != level[i - 1]) first[level[i]] = i;
}

/* Conditional reliability of record r once its subtree is accumulated: its
 * own success mass over its mass, plus its children's shares, which acc[r]
 * holds in fixed point (units of 2^-62 of r's mass).  Integer sums do not
 * depend on the order the device adds in, so the labels repeat exactly;
 * each share is truncated by at most 2^-62, so the label's error is below
 * depth * children * 2^-62. */
__device__ inline double log_cond(const double* mass, const double* succ, const u64* acc,
                                  u64 r) {
  const double m = mass[r];
  if (!(m > 0.0)) return 0.0;
  const double c = succ[r] / m + (double)acc[r] * 0x1p-62;
  return c < 1.0 ? c : 1.0;
}

/* one depth, deepest first: each child's finished subtree size is added to
 * its parent's (the root has no record); with acc, the child's conditional
 * reliability times its share of the parent's mass is added to the
 * parent's fixed-point sum */
__global__ void __launch_bounds__(256) k_log_accumulate(const u32* parent, u64* size, u64 r0,
                                                        u64 r1, const double* mass,
                                                        const double* succ, u64* acc) {
  for (u64 r = r0 + (u64)blockIdx.x * blockDim.x + threadIdx.x; r < r1;
       r += (u64)gridDim.x * blockDim.x) {
    const u32 p = parent[r];
    if (p == 0xffffffffu) continue;
    atomicAdd((unsigned long long*)&size[p], (unsigned long long)size[r]);
    if (acc && mass[p] > 0.0) {
      const double x = log_cond(mass, succ, acc, r) * (mass[r] / mass[p]);
      const u64 v = (u64)((x < 1.0 ? x : 1.0) * 0x1p62);
      if (v) atomicAdd((unsigned long long*)&acc[p], (unsigned long long)v);
    }
  }
}

/* conditional reliabilities of the records ids[0 .. m-1] */
__global__ void __launch_bounds__(256) k_log_cond(const u64* ids, u64 m, const double* mass,
                                                  const double* succ, const u64* acc, double* C) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (u64)gridDim.x * blockDim.x)
    C[i] = log_cond(mass, succ, acc, ids[i]);
}

__global__ void __launch_bounds__(256) k_log_gather(const float* feat, const u64* size,
                                                    const u64* idx, u64 m, float* X, float* Y,
                                                    const double* mass, const double* succ,
                                                    const u64* acc, double* C) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < m;
       i += (u64)gridDim.x * blockDim.x) {
    const u64 r = idx[i];
    for (int k = 0; k < kFeat; ++k) X[i * kFeat + k] = feat[r * kFeat + k];
    Y[i] = (float)log2((double)size[r]);
    if (C) C[i] = log_cond(mass, succ, acc, r);
  }
}

/* Seeded hash of a record's content (features and subtree size): the
 * sample of a log depends only on the tree, not on the order in which the
 * device appended its records (splitmix64 finalizer over the words). */
__device__ inline u64 mix64(u64 z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256) k_log_select(const float* feat, const u64* size, u64 n,
                                                    u64 seed, u64 thr, u64* keys, u64* ids,
                                                    u64* found, u64 cap, int need_const) {
  for (u64 r = (u64)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (u64)gridDim.x * blockDim.x) {
    /* choice logs: only records that were expanded carry features (their
     * constant feature is 1; the rest of the log is zero) */
    if (need_const && feat[r * kFeat + kFeat - 1] != 1.0f) continue;
    u64 h = mix64(seed ^ mix64(size[r]));
    const u32* f = (const u32*)(feat + r * kFeat);
    for (int k = 0; k < kFeat; ++k) h = mix64(h ^ (u64)f[k] ^ ((u64)k << 32));
    if (h < thr) {
      const u64 at = atomicAdd((unsigned long long*)found, 1ull);
      if (at < cap) {
        keys[at] = h;
        ids[at] = r;
      }
    }
  }
}

/* Tree log of a level-order solve.  choices = 0: the children's features
 * at creation (the learned order's samples); choices = 1: the learned
 * choice's samples, from a solve of the alternating rule's path whose sweep
 * direction explores (a reproducible pseudo-random bit per item), each
 * expanded item's features of the direction taken.  Labels: log2 subtree
 * nodes of the record. */
int log_subproblems(Workspace& ws, const nr_graph* g, const nr_config* cfg, long long max_out,
                    unsigned long long seed, float* X, float* Y, long long* n_out,
                    nr_result* res, int choices = 0, double* C = nullptr) {
  size_t free_b = 0, total_b = 0;
  NR_HIP(hipMemGetInfo(&free_b, &total_b));
  const size_t rec_bytes = kFeat * 4 * (choices ? 2 : 1) + 4 + 4 + 8 + (C ? 24 : 0);
  const u64 log_cap = std::min<u64>((u64)(free_b / 4 / rec_bytes), 1ull << 31);
  DevMem mem;
  LevelExtra x;
  x.mode = 1;
  x.log_cap = log_cap;
  u64* dsize;
  int rc;
  if ((rc = mem.alloc(&x.log_feat, (size_t)log_cap * kFeat)) ||
      (rc = mem.alloc(&x.log_parent, log_cap)) || (rc = mem.alloc(&x.log_level, log_cap)) ||
      (rc = mem.alloc(&x.log_count, 1)) || (rc = mem.alloc(&dsize, log_cap)))
    return rc;
  hipStream_t st = ws.stream;
  if (choices) {
    if ((rc = mem.alloc(&x.log_choice, (size_t)log_cap * kChoiceFeat))) return rc;
    NR_HIP(hipMemsetAsync(x.log_choice, 0, sizeof(float) * kChoiceFeat * log_cap, st));
  }
  u64* dacc = nullptr;
  if (C) {
    if ((rc = mem.alloc(&x.log_mass, log_cap)) || (rc = mem.alloc(&x.log_succ, log_cap)) ||
        (rc = mem.alloc(&dacc, log_cap)))
      return rc;
    NR_HIP(hipMemsetAsync(x.log_succ, 0, sizeof(double) * log_cap, st));
    NR_HIP(hipMemsetAsync(dacc, 0, sizeof(u64) * log_cap, st));
  }
  NR_HIP(hipMemsetAsync(x.log_count, 0, sizeof(u64), st));
  nr_config c2 = *cfg;
  c2.order = NR_ORDER_LEVEL;
  c2.time_budget_s = 0.0; /* a log needs whole subtrees */
  if (choices) c2.rule = NR_RULE_LEARNED;
  if ((rc = solve_level(ws, g, &c2, res, &x))) return rc;
  u64 count = 0;
  NR_HIP(hipMemcpy(&count, x.log_count, sizeof(u64), hipMemcpyDeviceToHost));
  if (count > log_cap) return NR_ERR_NOMEM;   /* tree larger than the log */
  if (res->overflowed) return NR_ERR_ARG;     /* capacity too small for a whole tree */
  const int depth = (int)res->iterations;     /* records have depth 1 .. depth - 1 */
  std::vector<u64> first((size_t)depth + 2, count);
  if (count) {
    u64* dfirst;
    if ((rc = mem.alloc(&dfirst, first.size()))) return rc;
    NR_HIP(hipMemcpyAsync(dfirst, first.data(), 8 * first.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_log_sizes, dim3(1024), dim3(256), 0, st, dsize, count);
    hipLaunchKernelGGL(k_log_bounds, dim3(1024), dim3(256), 0, st, (const u32*)x.log_level,
                       count, dfirst);
    NR_HIP(hipMemcpyAsync(first.data(), dfirst, 8 * first.size(), hipMemcpyDeviceToHost, st));
    NR_HIP(hipStreamSynchronize(st));
    /* a depth with no record keeps `count`; ends run to the next depth's start */
    for (int d = depth; d >= 2; --d) {
      const u64 r0 = first[d];
      u64 r1 = count;
      for (int e = d + 1; e <= depth + 1; ++e)
        if (first[e] < r1 && first[e] >= r0) r1 = first[e];
      if (r0 >= r1) continue;
      hipLaunchKernelGGL(k_log_accumulate, dim3(1024), dim3(256), 0, st,
                         (const u32*)x.log_parent, dsize, r0, r1, (const double*)x.log_mass,
                         (const double*)x.log_succ, dacc);
    }
  }
  /* the sample: records whose seeded content hash falls below a threshold
   * set for ~m hits, the m smallest hashes kept, in hash order; identical
   * for identical trees whatever the order the device wrote them in */
  const u64 want = std::min<u64>((u64)std::max<long long>(max_out, 0), count);
  std::vector<u64> idx;
  if (want > 0) {
    /* identical records share a hash, so a threshold aimed at ~want hits
     * can fall short: widen it until it catches want records (records of
     * equal content are interchangeable, so which copies come is moot) */
    double frac = std::min(1.0, ((double)want + 6.0 * std::sqrt((double)want) + 16.0) /
                                    (double)count);
    u64 *dkeys = nullptr, *dids = nullptr, *dfound;
    if ((rc = mem.alloc(&dkeys, count)) || (rc = mem.alloc(&dids, count)) ||
        (rc = mem.alloc(&dfound, 1)))
      return rc;
    u64 found = 0;
    for (;;) {
      const u64 thr = frac >= 1.0 ? ~0ull : (u64)(frac * 18446744073709551615.0);
      NR_HIP(hipMemsetAsync(dfound, 0, sizeof(u64), st));
      hipLaunchKernelGGL(k_log_select, dim3(1024), dim3(256), 0, st,
                         (const float*)(choices ? x.log_choice : x.log_feat), (const u64*)dsize,
                         count, (u64)seed, thr, dkeys, dids, dfound, count, choices);
      NR_HIP(hipMemcpyAsync(&found, dfound, sizeof(u64), hipMemcpyDeviceToHost, st));
      NR_HIP(hipStreamSynchronize(st));
      if (found >= want || frac >= 1.0) break;
      frac = std::min(1.0, frac * 2.0);
    }
    found = std::min(found, count);
    std::vector<u64> keys(found), ids(found);
    NR_HIP(hipMemcpyAsync(keys.data(), dkeys, 8 * found, hipMemcpyDeviceToHost, st));
    NR_HIP(hipMemcpyAsync(ids.data(), dids, 8 * found, hipMemcpyDeviceToHost, st));
    NR_HIP(hipStreamSynchronize(st));
    /* records of equal content (features and size) share a hash and are
     * interchangeable, except for their conditional reliabilities: when
     * those are returned they order equal hashes before the record id
     * (the order the device wrote them in), so the labels repeat too */
    std::vector<u64> cbits(C ? found : 0, 0);
    if (C && found) {
      double* dcf;
      if ((rc = mem.alloc(&dcf, found))) return rc;
      hipLaunchKernelGGL(k_log_cond, dim3(1024), dim3(256), 0, st, (const u64*)dids, found,
                         (const double*)x.log_mass, (const double*)x.log_succ, (const u64*)dacc,
                         dcf);
      NR_HIP(hipMemcpyAsync(cbits.data(), dcf, 8 * found, hipMemcpyDeviceToHost, st));
      NR_HIP(hipStreamSynchronize(st));
    }
    std::vector<u64> order(found);
    for (u64 i = 0; i < found; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](u64 x1, u64 x2) {
      if (keys[x1] != keys[x2]) return keys[x1] < keys[x2];
      if (C && cbits[x1] != cbits[x2]) return cbits[x1] < cbits[x2];
      return ids[x1] < ids[x2];
    });
    const u64 m0 = std::min<u64>(want, found);
    idx.resize((size_t)m0);
    for (u64 i = 0; i < m0; ++i) idx[i] = ids[order[i]];
  }
  const u64 m = (u64)idx.size();
  if (m) {
    u64* didx;
    float *dX, *dY;
    double* dC = nullptr;
    if ((rc = mem.alloc(&didx, m)) || (rc = mem.alloc(&dX, (size_t)m * kFeat)) ||
        (rc = mem.alloc(&dY, m)) || (C && (rc = mem.alloc(&dC, m))))
      return rc;
    NR_HIP(hipMemcpyAsync(didx, idx.data(), 8 * m, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_log_gather, dim3(1024), dim3(256), 0, st,
                       (const float*)(choices ? x.log_choice : x.log_feat), (const u64*)dsize,
                       (const u64*)didx, m, dX, dY, (const double*)x.log_mass,
                       (const double*)x.log_succ, (const u64*)dacc, dC);
    NR_HIP(hipMemcpyAsync(X, dX, sizeof(float) * m * kFeat, hipMemcpyDeviceToHost, st));
    NR_HIP(hipMemcpyAsync(Y, dY, sizeof(float) * m, hipMemcpyDeviceToHost, st));
    if (C) NR_HIP(hipMemcpyAsync(C, dC, sizeof(double) * m, hipMemcpyDeviceToHost, st));
    NR_HIP(hipStreamSynchronize(st));
  }
  *n_out = (long long)m;
  return NR_OK;
}

const MlpDev* model_dev(const nr_model* m);     /* nr_learn.hip: the order scorer or null */
const MlpDev* model_chooser(const nr_model* m); /* the path chooser or null */

/* A solve that uses a model: the learned order (cfg->order ==
 * NR_ORDER_LEARNED, the model's order scorer) and/or the learned path
 * choice (cfg->rule == NR_RULE_LEARNED, its chooser). */
int solve_learned(Workspace& ws, const nr_graph* g, const nr_config* cfg, const nr_model* model,
                  nr_result* out) {
  const MlpDev* order = model_dev(model);
  const MlpDev* chooser = model_chooser(model);
  const bool learned_order = cfg->order == NR_ORDER_LEARNED;
  const bool learned_rule = cfg->rule == NR_RULE_LEARNED;
  if ((learned_order && !order) || (learned_rule && !chooser)) return NR_ERR_ARG;
  DevMem mem;
  MlpDev* dm;
  int rc;
  if ((rc = mem.alloc(&dm, 2))) return rc;
  LevelExtra x;
  if (learned_order) {
    NR_HIP(hipMemcpyAsync(dm, order, sizeof(MlpDev), hipMemcpyHostToDevice, ws.stream));
    x.mode = 2;
    x.mlp = dm;
  }
  if (learned_rule) {
    NR_HIP(hipMemcpyAsync(dm + 1, chooser, sizeof(MlpDev), hipMemcpyHostToDevice, ws.stream));
    x.chooser = dm + 1;
  }
  return solve_level(ws, g, cfg, out, &x);
}

/* Locks and opens the workspace of a device. */
struct WsLock {
  Workspace* ws = nullptr;
  std::unique_lock<std::mutex> lk;
  int rc = NR_OK;
  explicit WsLock(int device) {
    if (device < 0 || device >= kMaxDevices) {
      rc = NR_ERR_ARG;
      return;
    }
    ws = &g_ws[device];
    lk = std::unique_lock<std::mutex>(ws->mu);
    rc = ws->open(device);
    if (rc == NR_OK && hipSetDevice(device) != hipSuccess) rc = NR_ERR_DEVICE;
  }
};

}  // namespace nr

extern "C" {

int nr_solve(const nr_graph* g, const nr_config* cfg, nr_result* out) {
  return nr::guarded([&]() -> int {
    if (!g || !out) return NR_ERR_ARG;
    nr_config def;
    if (!cfg) {
      nr_config_init(&def);
      cfg = &def;
    }
    if (cfg->queue_capacity < 0 || cfg->truncation_eps < 0 || cfg->time_budget_s < 0)
      return NR_ERR_ARG;
    if (cfg->order == NR_ORDER_LEARNED) return NR_ERR_ARG; /* needs a model: nr_solve_learned */
    if (cfg->order != NR_ORDER_LEVEL && cfg->order != NR_ORDER_MASS) return NR_ERR_ARG;
    if (cfg->rule != NR_RULE_PATH && cfg->rule != NR_RULE_FORKBEST &&
        cfg->rule != NR_RULE_ALTERNATE && cfg->rule != NR_RULE_PIVOT)
      return NR_ERR_ARG;
    if (cfg->sweep_period < 0) return NR_ERR_ARG;
    nr::Variant var;
    int rc = nr::pick_variant(g, var);
    if (rc == NR_ERR_UNSUPPORTED) { /* past the per-lane forms: one wave per item */
      if (cfg->prefix_edges > 0) return NR_ERR_UNSUPPORTED;
      nr::WsLock L(cfg->device);
      if (L.rc) return L.rc;
      return nr::solve_big(*L.ws, g, cfg, out);
    }
    if (rc) return rc;
    nr::WsLock L(cfg->device);
    if (L.rc) return L.rc;
    return nr::solve_level(*L.ws, g, cfg, out);
  });
}

int nr_solve_static(const nr_graph* g, const nr_config* cfg, nr_result* out) {
  return nr::guarded([&]() -> int {
    if (!g || !out) return NR_ERR_ARG;
    nr_config def;
    if (!cfg) {
      nr_config_init(&def);
      cfg = &def;
    }
    if (cfg->queue_capacity < 0 || cfg->truncation_eps < 0 || cfg->time_budget_s < 0)
      return NR_ERR_ARG;
    if (cfg->order != NR_ORDER_LEVEL && cfg->order != NR_ORDER_MASS) return NR_ERR_ARG;
    if (cfg->rule != NR_RULE_PATH && cfg->rule != NR_RULE_FORKBEST &&
        cfg->rule != NR_RULE_ALTERNATE && cfg->rule != NR_RULE_PIVOT)
      return NR_ERR_ARG;
    if (cfg->sweep_period < 0) return NR_ERR_ARG;
    if (cfg->device < 0 || cfg->device >= nr::kMaxDevices) return NR_ERR_ARG;
    nr::Variant var;
    int rc = nr::pick_variant(g, var);
    if (rc) return rc; /* the one-wave-per-item form keeps its own buffers */
    nr::Workspace& ws = nr::g_ws_fixed[cfg->device];
    std::lock_guard<std::mutex> lk(ws.mu);
    if (hipSetDevice(cfg->device) != hipSuccess) return NR_ERR_DEVICE;
    if (!ws.ready) { /* once per device, with the device current */
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(nr::scratch)) != hipSuccess || !p)
        return NR_ERR_DEVICE;
      if (hipGetDeviceProperties(&ws.prop, cfg->device) != hipSuccess) return NR_ERR_DEVICE;
      ws.device = cfg->device;
      ws.fixed = true;
      ws.arena = (unsigned char*)p;
      ws.arena_size = nr::kScratchBytes;
      ws.stream = nullptr; /* the null stream */
      ws.hctl = &ws.host_ctl;
      ws.hmirror = ws.dmirror = nullptr;
      ws.ready = true;
    }
    /* every buffer is carved afresh: offsets restart, tables are uploaded
     * again, and k_init clears what must start at zero */
    ws.arena_used = 0;
    ws.q = ws.tab = ws.meta = nullptr;
    ws.q_bytes = ws.tab_bytes = ws.meta_bytes = 0;
    ws.tab_g = nullptr;
    void* c = nullptr;
    size_t have = 0;
    if ((rc = ws.grow(&c, &have, sizeof(nr::LevelCtl)))) return rc;
    ws.ctl = (nr::LevelCtl*)c;
    return nr::solve_level(ws, g, cfg, out);
  });
}

int nr_tune_sweep_period(const nr_graph* g, const nr_config* cfg, double pilot_eps,
                         int max_period, int* best, double* pilot_esp) {
  return nr::guarded([&]() -> int {
    if (!g || !best || max_period < 1 || max_period > 64 || !(pilot_eps > 0.0)) return NR_ERR_ARG;
    nr_config c;
    if (cfg)
      c = *cfg;
    else
      nr_config_init(&c);
    c.rule = NR_RULE_ALTERNATE;
    c.order = NR_ORDER_LEVEL;
    c.truncation_eps = pilot_eps;
    c.time_budget_s = 0.0;
    *best = 0;
    double best_esp = 0.0;
    unsigned long long best_nodes = 0;
    for (int per = 1; per <= max_period; ++per) {
      c.sweep_period = per;
      nr_result r;
      const int rc = nr_solve(g, &c, &r);
      if (rc) return rc;
      if (pilot_esp) pilot_esp[per - 1] = r.ESP;
      if (*best == 0 || r.ESP < best_esp || (r.ESP == best_esp && r.nodes < best_nodes)) {
        *best = per;
        best_esp = r.ESP;
        best_nodes = r.nodes;
      }
    }
    return NR_OK;
  });
}

int nr_solve_learned(const nr_graph* g, const nr_config* cfg, const nr_model* model,
                     nr_result* out) {
  return nr::guarded([&]() -> int {
    if (!g || !out || !model) return NR_ERR_ARG;
    nr_config def;
    if (!cfg) {
      nr_config_init(&def);
      cfg = &def;
    }
    if (cfg->queue_capacity < 0 || cfg->truncation_eps < 0 || cfg->time_budget_s < 0)
      return NR_ERR_ARG;
    if (cfg->order != NR_ORDER_LEVEL && cfg->order != NR_ORDER_MASS &&
        cfg->order != NR_ORDER_LEARNED)
      return NR_ERR_ARG;
    if (cfg->rule != NR_RULE_PATH && cfg->rule != NR_RULE_ALTERNATE &&
        cfg->rule != NR_RULE_LEARNED)
      return NR_ERR_UNSUPPORTED; /* scores path children / path sweeps */
    if (cfg->sweep_period < 0) return NR_ERR_ARG;
    nr::Variant var;
    int rc = nr::pick_variant(g, var);
    if (rc) return rc;
    nr::WsLock L(cfg->device);
    if (L.rc) return L.rc;
    return nr::solve_learned(*L.ws, g, cfg, model, out);
  });
}

int nr_solve_checkpointed(const nr_graph* g, const nr_config* cfg, const nr_checkpoint* resume,
                          nr_checkpoint** kept, nr_result* out) {
  return nr::guarded([&]() -> int {
    if (!g || !out || !kept) return NR_ERR_ARG;
    *kept = nullptr;
    nr_config def;
    if (!cfg) {
      nr_config_init(&def);
      cfg = &def;
    }
    if (cfg->queue_capacity < 0 || cfg->truncation_eps < 0 || cfg->time_budget_s < 0)
      return NR_ERR_ARG;
    if (cfg->order != NR_ORDER_LEVEL && cfg->order != NR_ORDER_MASS) return NR_ERR_ARG;
    if (cfg->rule != NR_RULE_PATH && cfg->rule != NR_RULE_FORKBEST &&
        cfg->rule != NR_RULE_ALTERNATE && cfg->rule != NR_RULE_PIVOT)
      return NR_ERR_ARG;
    if (cfg->sweep_period < 0) return NR_ERR_ARG;
    if (resume && (resume->graph_hash != nr::graph_hash(g) || resume->w != nr::mask_words(g) ||
                   resume->rule != cfg->rule || resume->order != cfg->order ||
                 resume->eps != (cfg->truncation_eps > 0 ? cfg->truncation_eps : 0.0) ||
                   (cfg->rule == NR_RULE_ALTERNATE &&
                    resume->period != (cfg->sweep_period > 0 ? cfg->sweep_period : 2))))
      return NR_ERR_ARG; /* another graph, rule, sweep period, order or eps */
    nr::Variant var;
    int rc = nr::pick_variant(g, var);
    if (rc) return rc;
    nr::WsLock L(cfg->device);
    if (L.rc) return L.rc;
    nr::LevelExtra x;
    x.resume = resume;
    x.save = kept;
    return nr::solve_level(*L.ws, g, cfg, out, &x);
  });
}

int nr_checkpoint_info(const nr_checkpoint* c, long long* pending, double* pending_mass,
                       long long* nodes_done) {
  return nr::guarded([&]() -> int {
    if (!c) return NR_ERR_ARG;
    if (pending) *pending = (long long)c->count;
    if (pending_mass) {
      double s = 0.0;
      for (double m : c->mass) s += m;
      *pending_mass = s;
    }
    if (nodes_done) *nodes_done = (long long)c->counts[nr::C_NODES];
    return NR_OK;
  });
}

void nr_checkpoint_destroy(nr_checkpoint* c) { delete c; }

static const char kCkptMagic[8] = {'N', 'R', 'C', 'K', 'P', 'T', '0', '1'};

int nr_checkpoint_save(const nr_checkpoint* c, const char* path) {
  return nr::guarded([&]() -> int {
    if (!c || !path) return NR_ERR_ARG;
    FILE* f = std::fopen(path, "wb");
    if (!f) return NR_ERR_IO;
    const int hdr[8] = {c->n, c->e, c->w, c->src, c->dst, c->rule, c->order, c->period};
    const unsigned long long tail[3] = {c->iterations, c->max_width, c->count};
    bool ok = std::fwrite(kCkptMagic, 1, 8, f) == 8 && std::fwrite(hdr, sizeof hdr, 1, f) == 1 &&
              std::fwrite(&c->eps, 8, 1, f) == 1 && std::fwrite(&c->graph_hash, 8, 1, f) == 1 &&
              std::fwrite(c->sum_hi, 8, 3, f) == 3 && std::fwrite(c->sum_lo, 8, 3, f) == 3 &&
              std::fwrite(c->counts, 8, 8, f) == 8 && std::fwrite(tail, 8, 3, f) == 3;
    const size_t nw = (size_t)c->count * c->w;
    ok = ok && std::fwrite(c->known.data(), 8, nw, f) == nw &&
         std::fwrite(c->up.data(), 8, nw, f) == nw &&
         std::fwrite(c->mass.data(), 8, c->count, f) == c->count;
    ok = (std::fclose(f) == 0) && ok;
    return ok ? NR_OK : NR_ERR_IO;
  });
}

int nr_checkpoint_load(const char* path, nr_checkpoint** out) {
  return nr::guarded([&]() -> int {
    if (!path || !out) return NR_ERR_ARG;
    *out = nullptr;
    FILE* f = std::fopen(path, "rb");
    if (!f) return NR_ERR_IO;
    std::unique_ptr<nr_checkpoint> c(new (std::nothrow) nr_checkpoint);
    char magic[8];
    int hdr[8];
    unsigned long long tail[3];
    bool ok = c && std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kCkptMagic, 8) == 0 &&
              std::fread(hdr, sizeof hdr, 1, f) == 1 && std::fread(&c->eps, 8, 1, f) == 1 &&
              std::fread(&c->graph_hash, 8, 1, f) == 1 && std::fread(c->sum_hi, 8, 3, f) == 3 &&
              std::fread(c->sum_lo, 8, 3, f) == 3 && std::fread(c->counts, 8, 8, f) == 8 &&
              std::fread(tail, 8, 3, f) == 3;
    if (ok) {
      c->n = hdr[0];
      c->e = hdr[1];
      c->w = hdr[2];
      c->src = hdr[3];
      c->dst = hdr[4];
      c->rule = hdr[5];
      c->order = hdr[6];
      c->period = hdr[7];
      c->iterations = tail[0];
      c->max_width = tail[1];
      c->count = tail[2];
      ok = c->w >= 1 && c->w <= 16 && c->count < (1ull << 34);
    }
    if (ok) { /* the items must be in the file before they size any buffer */
      const long here = std::ftell(f);
      ok = here >= 0 && std::fseek(f, 0, SEEK_END) == 0;
      const long end = ok ? std::ftell(f) : -1;
      ok = ok && end >= here && std::fseek(f, here, SEEK_SET) == 0 &&
           (unsigned long long)(end - here) == c->count * (16ull * (unsigned)c->w + 8ull);
    }
    if (ok) {
      const size_t nw = (size_t)c->count * c->w;
      c->known.resize(nw);
      c->up.resize(nw);
      c->mass.resize(c->count);
      ok = std::fread(c->known.data(), 8, nw, f) == nw && std::fread(c->up.data(), 8, nw, f) == nw &&
           std::fread(c->mass.data(), 8, c->count, f) == c->count;
    }
    std::fclose(f);
    if (!ok) return c ? NR_ERR_PARSE : NR_ERR_NOMEM;
    *out = c.release();
    return NR_OK;
  });
}

int nr_log_subproblems(const nr_graph* g, const nr_config* cfg, long long max_records,
                       unsigned long long seed, float* features, float* labels,
                       long long* n_out, nr_result* result) {
  return nr::guarded([&]() -> int {
    if (!g || !cfg || max_records < 0 || !n_out || !result ||
        (max_records > 0 && (!features || !labels)))
      return NR_ERR_ARG;
    if (cfg->queue_capacity < 0 || cfg->truncation_eps < 0 || cfg->sweep_period < 0)
      return NR_ERR_ARG;
    if (cfg->rule != NR_RULE_PATH && cfg->rule != NR_RULE_ALTERNATE) return NR_ERR_UNSUPPORTED;
    nr::Variant var;
    int rc = nr::pick_variant(g, var);
    if (rc) return rc;
    if (!var.level_log) return NR_ERR_UNSUPPORTED;
    nr::WsLock L(cfg->device);
    if (L.rc) return L.rc;
    return nr::log_subproblems(*L.ws, g, cfg, max_records, seed, features, labels, n_out, result);
  });
}

int nr_log_subproblems_rel(const nr_graph* g, const nr_config* cfg, long long max_records,
                           unsigned long long seed, float* features, float* labels,
                           double* cond_rel, long long* n_out, nr_result* result) {
  return nr::guarded([&]() -> int {
    if (!g || !cfg || max_records < 0 || !n_out || !result ||
        (max_records > 0 && (!features || !labels || !cond_rel)))
      return NR_ERR_ARG;
    if (cfg->queue_capacity < 0 || cfg->truncation_eps < 0 || cfg->sweep_period < 0)
      return NR_ERR_ARG;
    if (cfg->rule != NR_RULE_PATH && cfg->rule != NR_RULE_ALTERNATE) return NR_ERR_UNSUPPORTED;
    nr::Variant var;
    int rc = nr::pick_variant(g, var);
    if (rc) return rc;
    if (!var.level_log) return NR_ERR_UNSUPPORTED;
    nr::WsLock L(cfg->device);
    if (L.rc) return L.rc;
    /* cond_rel is null only with max_records == 0: no labels to return */
    return nr::log_subproblems(*L.ws, g, cfg, max_records, seed, features, labels, n_out, result,
                               0, cond_rel);
  });
}

int nr_log_choices(const nr_graph* g, const nr_config* cfg, long long max_records,
                   unsigned long long seed, float* features, float* labels, long long* n_out,
                   nr_result* result) {
  return nr::guarded([&]() -> int {
    if (!g || !cfg || max_records < 0 || !n_out || !result ||
        (max_records > 0 && (!features || !labels)))
      return NR_ERR_ARG;
    if (cfg->queue_capacity < 0 || cfg->truncation_eps < 0 || cfg->sweep_period < 0)
      return NR_ERR_ARG;
    if (!g->symmetric || !g->terms.empty()) return NR_ERR_UNSUPPORTED;
    nr::Variant var;
    int rc = nr::pick_variant(g, var);
    if (rc) return rc;
    nr::WsLock L(cfg->device);
    if (L.rc) return L.rc;
    return nr::log_subproblems(*L.ws, g, cfg, max_records, seed, features, labels, n_out, result,
                               1);
  });
}

int nr_tune_batch(const nr_graph* const* graphs, int count, const nr_config* cfg, int sample,
                  nr_config* tuned, double* best_rate) {
  return nr::guarded([&]() -> int {
    if (!graphs || count <= 0 || !tuned || sample < 0) return NR_ERR_ARG;
    nr_config c;
    if (cfg)
      c = *cfg;
    else
      nr_config_init(&c);
    const int n = sample > 0 && sample < count ? sample : count;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, c.device) != hipSuccess) return NR_ERR_DEVICE;
    const int cus = prop.multiProcessorCount;
    std::vector<nr_result> res((size_t)n);
    *tuned = c;
    double best = -1.0;
    /* one trial: the subproblems per second of a warm batch, or < 0 when the
     * schedule does not fit */
    auto trial = [&](const nr_config& t, double* rate) -> int {
      *rate = -1.0;
      for (int rep = 0; rep < 2; ++rep) { /* the first run warms up */
        const int rc = nr_solve_batch(graphs, n, &t, res.data());
        if (rc == NR_ERR_UNSUPPORTED || rc == NR_ERR_NOMEM) return NR_OK; /* does not fit */
        if (rc) return rc;
      }
      double nodes = 0.0;
      for (const nr_result& r : res) nodes += (double)r.nodes;
      *rate = res[0].elapsed_ms > 0 ? nodes / (res[0].elapsed_ms * 1e-3) : 0.0;
      return NR_OK;
    };
    const int blocks[3] = {256, 128, 64};
    const int per_cu[4] = {0, 1, 2, 4}; /* 0: the occupancy limit */
    for (int b : blocks)
      for (int m : per_cu) {
        nr_config t = c;
        t.block_size = b;
        t.grid_blocks = m * cus;
        double rate = 0.0;
        if (int rc = trial(t, &rate)) return rc;
        if (rate > best) {
          best = rate;
          *tuned = t;
        }
      }
    if (best < 0) return NR_ERR_UNSUPPORTED;
    /* the queue chunk for the fastest launch shape */
    const long long cap = c.queue_capacity > 0 ? c.queue_capacity : (1ll << 20);
    const nr_config shape = *tuned;
    for (long long div : {8ll, 2ll}) {
      nr_config t = shape;
      t.chunk_items = (int)std::min<long long>(std::max<long long>(cap / div, 1), 1ll << 30);
      double rate = 0.0;
      if (int rc = trial(t, &rate)) return rc;
      if (rate > best) {
        best = rate;
        *tuned = t;
      }
    }
    if (best_rate) *best_rate = best;
    return NR_OK;
  });
}

int nr_solve_batch(const nr_graph* const* graphs, int count, const nr_config* cfg,
                   nr_result* results) {
  return nr::guarded([&]() -> int {
    if (!graphs || count < 0 || !results) return NR_ERR_ARG;
    if (count == 0) return NR_OK;
    nr_config def;
    if (!cfg) {
      nr_config_init(&def);
      cfg = &def;
    }
    if (cfg->queue_capacity < 0 || cfg->truncation_eps < 0 || cfg->time_budget_s < 0)
      return NR_ERR_ARG;
    if (cfg->order != NR_ORDER_LEVEL || cfg->rule != NR_RULE_PATH) return NR_ERR_UNSUPPORTED;
    nr::WsLock L(cfg->device);
    if (L.rc) return L.rc;
    return nr::solve_batch(*L.ws, graphs, count, cfg, results);
  });
}

int nr_expand_items(const nr_graph* g, int device, int count,
                    const unsigned long long* known, const unsigned long long* up,
                    const double* mass, int max_children, int* nchild, double* succ,
                    unsigned long long* child_known, unsigned long long* child_up,
                    double* child_mass) {
  return nr::guarded([&]() -> int {
    if (!g || count < 0 || (count > 0 && (!known || !up || !mass || !nchild || !succ)))
      return NR_ERR_ARG;
    nr::Variant var;
    int rc = nr::pick_variant(g, var);
    if (rc) return rc;
    nr::WsLock L(device);
    if (L.rc) return L.rc;
    return nr::expand_items(*L.ws, g, NR_RULE_PATH, 0, count, known, up, mass, max_children,
                            nchild, succ, child_known, child_up, child_mass);
  });
}

int nr_expand_items_rule(const nr_graph* g, const nr_config* cfg, int count,
                         const unsigned long long* known, const unsigned long long* up,
                         const double* mass, int max_children, int* nchild, double* succ,
                         unsigned long long* child_known, unsigned long long* child_up,
                         double* child_mass) {
  return nr::guarded([&]() -> int {
    if (!g || !cfg || count < 0 || (count > 0 && (!known || !up || !mass || !nchild || !succ)))
      return NR_ERR_ARG;
    if (cfg->rule != NR_RULE_PATH && cfg->rule != NR_RULE_ALTERNATE) return NR_ERR_UNSUPPORTED;
    nr::Variant var;
    int rc = nr::pick_variant(g, var);
    if (rc) return rc;
    nr::WsLock L(cfg->device);
    if (L.rc) return L.rc;
    return nr::expand_items(*L.ws, g, cfg->rule, cfg->sweep_period, count, known, up, mass,
                            max_children, nchild, succ, child_known, child_up, child_mass);
  });
}

int nr_device_release(int device) {
  return nr::guarded([&]() -> int {
    if (device < 0 || device >= nr::kMaxDevices) return NR_ERR_ARG;
    nr::Workspace& ws = nr::g_ws[device];
    std::lock_guard<std::mutex> lk(ws.mu);
    ws.release();
    return NR_OK;
  });
}

int nr_device_query(int device, char* name, int name_len, int* compute_units,
                    int* wave_size, long long* total_mem) {
  return nr::guarded([&]() -> int {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return NR_ERR_DEVICE;
    if (name && name_len > 0) {
      std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (wave_size) *wave_size = prop.warpSize;
    if (total_mem) *total_mem = (long long)prop.totalGlobalMem;
    return NR_OK;
  });
}

}  // extern "C"
