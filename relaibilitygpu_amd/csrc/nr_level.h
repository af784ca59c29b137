/*
 * nr_level.h - the level-order expand kernel and the single-expansion
 * kernel, templated on the per-lane expander (nr_device.h): Expander for
 * undirected graphs of up to 64 nodes (items carry adjacency rows),
 * CsrExpander for larger or directed graphs (items are planes only).  Shared
 * by the translation units that instantiate them.
 */
#ifndef NR_LEVEL_H
#define NR_LEVEL_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <type_traits>

#include "nr_device.h"
#include "nr_learn.h"

namespace nr {

/* Mass classes: two per binade, class 0 = heaviest (mass >= 0.75).  The
 * class-round order expands the heaviest classes first. */
constexpr int kClasses = 128;

__host__ __device__ inline int mass_class(double m) {
  unsigned long long bits;
  memcpy(&bits, &m, sizeof bits);
  const int e = (int)((bits >> 52) & 0x7ff) - 1023; /* m in [2^e, 2^(e+1)) */
  const int heavy = (int)((bits >> 51) & 1);         /* m >= 1.5 * 2^e */
  int c = 2 * (-e - 1) + (heavy ? 0 : 1);
  return c < 0 ? 0 : (c >= kClasses ? kClasses - 1 : c);
}

/* LevelArgs::settle bits. */
constexpr int kSettleTrunc = 1, kSettleDead = 2;

/* Per-wave counter slots. */
enum { C_NODES = 0, C_PATHS, C_DEAD, C_TRUNC, C_OVER, C_FOLD, C_ERR, C_N = 8 };

/* Most levels one k_ramp_levels launch runs. */
constexpr int kRampMax = 24;

/* Partial totals of one finalize_slots() workgroup, and the most such
 * workgroups. */
struct FinPart {
  DD s[3]; /* R, ESP, D */
  u64 n[C_N];
};
constexpr int kFinMax = 64;

/* Control block of a level-order solve (device memory). */
struct LevelCtl {
  u64 hist[kClasses]; /* class histogram of the pool (class rounds) */
  u64 last_items; /* items of the most recently launched level */
  u64 iterations; /* non-empty levels processed */
  u64 max_width;  /* widest level */
  u64 counts[C_N];
  double sum[3];  /* R, ESP, D after k_finalize */
  double sum_lo[3]; /* their compensation terms (double-double low parts) */
  u64 start_tick;   /* wall clock (s_memrealtime) when the solve was seeded */
  /* k_ramp_levels: items entering each of its levels over all workgroups
   * (the last entry: the level it hands over), finished workgroups, and
   * whether any workgroup's region overflowed */
  u64 ramp_width[kRampMax + 1];
  u32 ramp_done, ramp_over;
  /* finalize_slots(): partials of its workgroups, finished workgroups, and
   * whether this solve's totals are final */
  FinPart fin[kFinMax];
  u32 fin_ticket, fin_done;
};


/* Bytes of the control block the host reads back (not the partials). */
constexpr size_t kCtlCopyBytes = offsetof(LevelCtl, fin);

/* Words of the pinned host mirror: [0] last_items, [1] iterations,
 * [2] max_width (written by each level launch), then the final sums, their
 * low parts, the counters and the solve's sequence number (written by
 * k_finalize, the sequence number last), and the ramp's overflow flag
 * (written by k_ramp_levels). */
constexpr int kMirSum = 3, kMirSumLo = 6, kMirCounts = 9, kMirSeq = 9 + C_N,
              kMirRamp = 10 + C_N, kMirWords = 11 + C_N;

struct LevelArgs {
  DevGraph g;
  Layout lay;
  QBuf buf[2];
  u32* seg_count[2]; /* [n_seg] items written by each workgroup */
  /* [n_seg] children each workgroup settled at creation (dead, or below
   * the truncation) instead of writing them: they belong to the width of
   * the level the buffer holds (read when `settle` is set) */
  u32* seg_settle[2];
  /* level order: children counted when they are created and never queued
   * (kSettleTrunc: those below the truncation; kSettleDead: the dead ones,
   * row form with two terminals) */
  int settle;
  u32* hist;         /* [n_seg][kClasses] class histogram of each segment */
  LevelCtl* ctl;
  /* pinned host mirror (kMir* words) of the counters the host reads at a
   * chunk boundary and at the end, written by workgroup 0 of each level
   * launch and by k_finalize; read after a stream synchronisation, so no
   * copy is queued for them; or null */
  u64* hmirror;
  u64 seq; /* this solve's number, stamped into the mirror by k_finalize */
  DD* partR;
  DD* partE;
  DD* partD;
  u64* cnt; /* [n_seg][C_N] per-workgroup counters */
  u64 stride; /* n_seg * seg_cap: distance between planes */
  u32 seg_cap;
  int n_seg;
  double eps;
  /* class rounds (max-mass order): items of class <= thr_class are
   * expanded, lighter ones are carried to the output unchanged; queued
   * items of class >= drop_class are given up to ESP */
  int mass_mode;
  int thr_class;
  int drop_class;
  /* class rounds after k_class_split: the output segments and histograms
   * already hold the carried items; the launch appends children to them */
  int append;
  /* tree log (level order): one record per child, written by its parent */
  u32* tag[2];        /* per slot: record id (log) or class (learned order) */
  float* log_feat;    /* [cap][kFeat] child features (bf16-rounded) */
  u32* log_parent;    /* [cap] parent record, 0xffffffff for the root */
  u32* log_level;     /* [cap] tree depth of the child */
  u64* log_count;
  u64 log_cap;
  /* optional (null: not logged): the child's FP64 mass, and the success
   * mass its own expansion finds (zero-filled beforehand; a dead or
   * truncated record keeps zero) */
  double* log_mass;
  double* log_succ;
  /* learned order */
  const MlpDev* mlp;
  /* learned path choice: the chooser's exact inference form; in a tree log
   * (MODE 1) the direction explores instead and each expanded record's
   * features of the direction taken go to log_choice[record][16] */
  const MlpDev* chooser;
  float* log_choice;
  /* time budget in wall-clock ticks from ctl->start_tick (0: none): items a
   * wave reaches after it are carried to the output unexpanded */
  u64 budget_ticks;
  /* level order: the first launch that finds the queue empty computes the
   * totals (finalize_slots over n_seg slots), so no launch follows the last
   * level but the empty ones already queued */
  int fin_on_empty;
};

template <int W>
__device__ inline void load_item(const QBuf& b, u64 stride, u64 i, u64 (&known)[W],
                                 u64 (&up)[W], double& mass) {
#pragma unroll
  for (int w = 0; w < W; ++w) {
    known[w] = b.known[w * stride + i];
    up[w] = b.up[w * stride + i];
  }
  mass = b.mass[i];
}

template <int W>
__device__ inline void store_item(const QBuf& b, u64 stride, u64 i, const u64 (&known)[W],
                                  const u64 (&up)[W], double mass) {
#pragma unroll
  for (int w = 0; w < W; ++w) {
    b.known[w * stride + i] = known[w];
    b.up[w * stride + i] = up[w];
  }
  b.mass[i] = mass;
}

/* Segment holding global item index `base`: the largest s with
 * sp[s] <= base (sp[0] = 0 <= base < sp[n_seg]). */
__device__ inline int segment_of(const u32* sp, int n_seg, u32 base) {
  int lo = 0, hi = n_seg;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (sp[mid] <= base)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

/* segment_of() by a whole wave (all 64 lanes active, `base` wave-uniform,
 * n_seg <= 4096): each lane probes one of 64 evenly spaced boundaries, the
 * ballot's highest hit picks the stretch, and one more probe per lane
 * inside it finds the segment — two LDS rounds instead of log2(n_seg).
 * The lanes that see sp[.] <= base form a prefix (sp is non-decreasing). */
__device__ inline int wave_segment_of(const u32* sp, int n_seg, u32 base) {
  const int lane = threadIdx.x & 63;
  const int step = (n_seg + 63) >> 6;
  const int p = lane * step;
  u64 m = __ballot(p < n_seg && sp[p] <= base);
  const int lo = (63 - __builtin_clzll(m)) * step;
  const int q = lo + lane;
  m = __ballot(lane < step && q < n_seg && sp[q] <= base);
  return lo + 63 - __builtin_clzll(m);
}

/* Exclusive prefix of the segment counts into sp[0..n_seg] (sp[n_seg] is
 * the total).  The counts are loaded first and `between` runs while they are
 * in flight (the level kernels stage their graph tables there); then a
 * shuffle scan per wave and one barrier to combine the waves. */
template <class F>
__device__ inline void block_prefix(const u32* counts, int n_seg, u32* sp, u32* part,
                                    F&& between) {
  const int t = threadIdx.x, T = blockDim.x, lane = t & 63, wave = t >> 6;
  const int n_w = (T + 63) >> 6;
  const int per = (n_seg + T - 1) / T;
  const int b0 = min(t * per, n_seg), b1 = min(b0 + per, n_seg);
  constexpr int kReg = 8; /* counts held in registers; more are read twice */
  u32 v[kReg];
#pragma unroll
  for (int j = 0; j < kReg; ++j) v[j] = b0 + j < b1 ? counts[b0 + j] : 0u;
  between();
  u32 sum = 0;
#pragma unroll
  for (int j = 0; j < kReg; ++j) sum += v[j];
  for (int b = b0 + kReg; b < b1; ++b) sum += counts[b];
  u32 incl = sum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const u32 x = __shfl_up(incl, off, 64);
    if (lane >= off) incl += x;
  }
  if (lane == 63) part[wave] = incl;
  __syncthreads();
  u32 run = incl - sum, total = 0;
  for (int w = 0; w < n_w; ++w) {
    const u32 x = part[w];
    if (w < wave) run += x;
    total += x;
  }
#pragma unroll
  for (int j = 0; j < kReg; ++j)
    if (b0 + j < b1) {
      sp[b0 + j] = run;
      run += v[j];
    }
  for (int b = b0 + kReg; b < b1; ++b) {
    sp[b] = run;
    run += counts[b];
  }
  if (t == 0) sp[n_seg] = total;
  __syncthreads();
}

__device__ inline void block_prefix(const u32* counts, int n_seg, u32* sp, u32* part) {
  block_prefix(counts, n_seg, sp, part, [] {});
}

/* Wave totals for block_flush without LDS traffic: DPP moves sum within
 * each 16-lane row (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row half
 * mirror, row mirror: after each step every lane holds its group's sum),
 * then the four row sums are read with readlane.  All 64 lanes must be
 * active; the result is wave-uniform. */
template <int CTRL>
__device__ inline u32 dpp_u32(u32 x) {
  return (u32)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}

template <int CTRL>
__device__ inline double dpp_f64(double x) {
  const u64 b = (u64)__double_as_longlong(x);
  const u64 lo = dpp_u32<CTRL>((u32)b), hi = dpp_u32<CTRL>((u32)(b >> 32));
  return __longlong_as_double((long long)(hi << 32 | lo));
}

__device__ inline double readlane_f64(double x, int l) {
  const u64 b = (u64)__double_as_longlong(x);
  const u64 lo = (u32)__builtin_amdgcn_readlane((int)(u32)b, l);
  const u64 hi = (u32)__builtin_amdgcn_readlane((int)(u32)(b >> 32), l);
  return __longlong_as_double((long long)(hi << 32 | lo));
}

__device__ inline u32 wave_total_u32(u32 x) {
  x += dpp_u32<0xB1>(x);
  x += dpp_u32<0x4E>(x);
  x += dpp_u32<0x141>(x);
  x += dpp_u32<0x140>(x);
  return (u32)__builtin_amdgcn_readlane((int)x, 0) + (u32)__builtin_amdgcn_readlane((int)x, 16) +
         (u32)__builtin_amdgcn_readlane((int)x, 32) + (u32)__builtin_amdgcn_readlane((int)x, 48);
}

/* (s, c) pairs combined with error-free additions, as wave_sum_dd() does;
 * lanes paired by a step hold equal pairs afterwards (two_sum is symmetric) */
__device__ inline void wave_total_dd(double& s, double& c) {
  auto add = [&](double os, double oc) {
    double t, e;
    two_sum(s, os, t, e);
    s = t;
    c = c + oc + e;
  };
  add(dpp_f64<0xB1>(s), dpp_f64<0xB1>(c));
  add(dpp_f64<0x4E>(s), dpp_f64<0x4E>(c));
  add(dpp_f64<0x141>(s), dpp_f64<0x141>(c));
  add(dpp_f64<0x140>(s), dpp_f64<0x140>(c));
  const double s1 = readlane_f64(s, 16), c1 = readlane_f64(c, 16);
  const double s2 = readlane_f64(s, 32), c2 = readlane_f64(c, 32);
  const double s3 = readlane_f64(s, 48), c3 = readlane_f64(c, 48);
  s = readlane_f64(s, 0);
  c = readlane_f64(c, 0);
  add(s1, c1);
  add(s2, c2);
  add(s3, c3);
}

__device__ inline void flush_counts(u64* slot, const u32 (&c)[C_N]) {
  bool any = false;
#pragma unroll
  for (int j = 0; j < C_N; ++j) any |= c[j] != 0;
  if (!any) return;
#pragma unroll
  for (int j = 0; j < C_N; ++j) slot[j] += c[j];
}

/* Flush a lane's accumulators into its workgroup's slot: wave sums, then
 * wave 0's lane 0 merges the workgroup's waves in wave order and updates the
 * block's double-double and counter slots (owned by this workgroup for the
 * launch; launches on one stream do not overlap).  `scratch` is LDS of at
 * least 80 bytes per wave. */
/* The workgroup's slot values (three double-doubles, then C_N u64 counters:
 * 28 dwords) loaded one dword per lane by lanes 0..27 of wave 0 when the
 * launch starts, so block_flush() need not wait for them at the end. */
__device__ inline u32 prefetch_slot(const LevelArgs& a) {
  const int t = threadIdx.x;
  if (t >= 4 * 3 + 2 * C_N) return 0u;
  const u32* src = t < 4    ? (const u32*)&a.partR[blockIdx.x] + t
                   : t < 8  ? (const u32*)&a.partE[blockIdx.x] + (t - 4)
                   : t < 12 ? (const u32*)&a.partD[blockIdx.x] + (t - 8)
                            : (const u32*)&a.cnt[(u64)blockIdx.x * C_N] + (t - 12);
  return *src;
}

__device__ inline double dword_pair(u32 lo, u32 hi) {
  return __longlong_as_double((long long)(((u64)hi << 32) | lo));
}

__device__ inline void block_flush(const LevelArgs& a, Acc2& accR, Acc2& accE, Acc2& accD,
                                   const u32 (&c)[C_N], unsigned char* scratch,
                                   const u32* pre = nullptr) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n_w = (blockDim.x + 63) >> 6;
  wave_total_dd(accR.s, accR.c);
  wave_total_dd(accE.s, accE.c);
  wave_total_dd(accD.s, accD.c);
  double* sd = (double*)scratch;      /* [n_w][6] */
  u32* sc = (u32*)(sd + 6 * n_w);     /* [n_w][C_N] */
  u32 t[C_N];
#pragma unroll
  for (int j = 0; j < C_N; ++j) t[j] = wave_total_u32(c[j]);
  if (lane == 0) {
    sd[wave * 6 + 0] = accR.s;
    sd[wave * 6 + 1] = accR.c;
    sd[wave * 6 + 2] = accE.s;
    sd[wave * 6 + 3] = accE.c;
    sd[wave * 6 + 4] = accD.s;
    sd[wave * 6 + 5] = accD.c;
#pragma unroll
    for (int j = 0; j < C_N; ++j) sc[wave * C_N + j] = t[j];
  }
  __syncthreads();
  if (wave != 0) return;
  /* wave 0 merges the waves' totals in wave order: lane k < 3 the double-
   * double k, lane j < C_N counter j (a few registers per lane); the slot's
   * prefetched dwords (lane i of wave 0 holds dword i) arrive by shuffles */
  const u32 pw = pre ? *pre : 0u;
  const int kd = lane < 3 ? lane : 0, kc = lane < C_N ? lane : 0;
  const u32 d0 = __shfl(pw, 4 * kd, 64), d1 = __shfl(pw, 4 * kd + 1, 64);
  const u32 d2 = __shfl(pw, 4 * kd + 2, 64), d3 = __shfl(pw, 4 * kd + 3, 64);
  const u32 c0 = __shfl(pw, 12 + 2 * kc, 64), c1 = __shfl(pw, 12 + 2 * kc + 1, 64);
  constexpr int kMaxW = 4; /* launch bounds: at most 256 threads */
  if (lane < 3) {
    double v[kMaxW][2];
#pragma unroll
    for (int w = 0; w < kMaxW; ++w) {
      v[w][0] = w < n_w ? sd[w * 6 + 2 * kd] : 0.0; /* a missing wave adds exact zeros */
      v[w][1] = w < n_w ? sd[w * 6 + 2 * kd + 1] : 0.0;
    }
    double s = 0.0, cc = 0.0;
#pragma unroll
    for (int w = 0; w < kMaxW; ++w) {
      double x, e;
      two_sum(s, v[w][0], x, e);
      s = x;
      cc += v[w][1] + e;
    }
    if (s != 0.0 || cc != 0.0) {
      DD* dst = kd == 0 ? &a.partR[blockIdx.x] : kd == 1 ? &a.partE[blockIdx.x] : &a.partD[blockIdx.x];
      if (pre) { /* dd_merge_slot() on the prefetched value */
        const double ohi = dword_pair(d0, d1);
        const double olo = dword_pair(d2, d3);
        double t, e;
        two_sum(ohi, s, t, e);
        const double lo = olo + cc + e;
        double hi2, e2;
        two_sum(t, lo, hi2, e2);
        dst->hi = hi2;
        dst->lo = e2;
      } else {
        dd_merge_slot(dst, s, cc);
      }
    }
  }
  u32 tot = 0;
  if (lane < C_N) {
    u32 cv[kMaxW];
#pragma unroll
    for (int w = 0; w < kMaxW; ++w) cv[w] = w < n_w ? sc[w * C_N + kc] : 0u;
#pragma unroll
    for (int w = 0; w < kMaxW; ++w) tot += cv[w];
  }
  if (__ballot(tot != 0) && lane < C_N) {
    u64* slot = &a.cnt[(u64)blockIdx.x * C_N + kc];
    *slot = (pre ? ((u64)c1 << 32 | c0) : *slot) + tot;
  }
}

__device__ inline u64 wave_total_u64(u64 x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

/* The solve's totals from the per-workgroup slots, in an order fixed by the
 * slot count alone (replaces the single-address FP64 CAS of main.cu:68-79 and
 * the global counters of main.cu:58-63): workgroup w < n_fin = min(kFinMax,
 * ceil(n_slots / 64)) takes slots w * 64 + lane (+ 64 n_fin, ...); part
 * k < 3 (R, ESP, D) is added up lane by lane with error-free additions and
 * then over the wave, part 3 (the counters) likewise, wave k of the
 * workgroup taking parts k, k + waves, ...; the partials go to ctl->fin[w].
 * The last of these workgroups to finish (ctl->fin_ticket) adds the partials
 * up the same way into ctl->sum, sum_lo and counts and the host mirror, each
 * wave fencing its mirror words before thread 0 stamps the sequence number,
 * then sets ctl->fin_done and clears the ticket.  Whole workgroups of any
 * multiple of 64 threads; s_last is one word of LDS. */
__device__ inline void finalize_slots(const LevelArgs& a, u64 n_slots, u32* s_last) {
  const int t = threadIdx.x, lane = t & 63, n_w = (int)(blockDim.x >> 6);
  const u64 n_fin = min((u64)kFinMax, (n_slots + 63) / 64);
  if (blockIdx.x >= n_fin) return;
  LevelCtl* ctl = a.ctl;
  FinPart& mine = ctl->fin[blockIdx.x];
  for (int k = t >> 6; k < 4; k += n_w) {
    if (k < 3) {
      const DD* src = k == 0 ? a.partR : (k == 1 ? a.partE : a.partD);
      double s = 0.0, c = 0.0;
      for (u64 i = (u64)blockIdx.x * 64 + lane; i < n_slots; i += 64 * n_fin) {
        const DD v = src[i];
        double x, e;
        two_sum(s, v.hi, x, e);
        s = x;
        c += v.lo + e;
      }
      wave_total_dd(s, c);
      if (lane == 0) {
        mine.s[k].hi = s;
        mine.s[k].lo = c;
      }
    } else {
      u64 n[C_N];
#pragma unroll
      for (int j = 0; j < C_N; ++j) n[j] = 0;
      for (u64 i = (u64)blockIdx.x * 64 + lane; i < n_slots; i += 64 * n_fin)
#pragma unroll
        for (int j = 0; j < C_N; ++j) n[j] += a.cnt[i * C_N + j];
#pragma unroll
      for (int j = 0; j < C_N; ++j) n[j] = wave_total_u64(n[j]);
      if (lane == 0)
#pragma unroll
        for (int j = 0; j < C_N; ++j) mine.n[j] = n[j];
    }
  }
  __syncthreads();
  if (t == 0) {
    __threadfence();
    *s_last = atomicAdd(&ctl->fin_ticket, 1u) == (u32)(n_fin - 1) ? 1u : 0u;
  }
  __syncthreads();
  if (!*s_last) return;
  __threadfence(); /* the other workgroups' partials */
  u64* m = a.hmirror;
  for (int k = t >> 6; k < 4; k += n_w) {
    if (k < 3) {
      double s = 0.0, c = 0.0;
      if ((u64)lane < n_fin) {
        s = ctl->fin[lane].s[k].hi;
        c = ctl->fin[lane].s[k].lo;
      }
      wave_total_dd(s, c);
      if (lane == 0) {
        double hi, lo;
        two_sum(s, c, hi, lo);
        ctl->sum[k] = hi;
        ctl->sum_lo[k] = lo;
        if (m) {
          m[kMirSum + k] = (u64)__double_as_longlong(hi);
          m[kMirSumLo + k] = (u64)__double_as_longlong(lo);
          __threadfence_system();
        }
      }
    } else {
      u64 n[C_N];
#pragma unroll
      for (int j = 0; j < C_N; ++j)
        n[j] = wave_total_u64((u64)lane < n_fin ? ctl->fin[lane].n[j] : 0);
      if (lane == 0) {
#pragma unroll
        for (int j = 0; j < C_N; ++j) {
          ctl->counts[j] = n[j];
          if (m) m[kMirCounts + j] = n[j];
        }
        if (m) __threadfence_system();
      }
    }
  }
  __syncthreads();
  if (t == 0) {
    ctl->fin_ticket = 0;
    ctl->fin_done = 1;
    if (m) {
      m[1] = ctl->iterations;
      m[2] = ctl->max_width;
      __threadfence_system(); /* the result words before the stamp */
      m[kMirSeq] = a.seq;
    }
  }
}

/* A wave's children written child-major (path rule): child c of the wave,
 * counted in lane-prefix order, goes to slot wbase + c of the workgroup's
 * segment and is written by lane c % 64, which finds its owner lane by a
 * binary search over the wave's inclusive prefix, takes the owner's planes
 * and mass by shuffles and its hops and rows from the owner's LDS slice, and
 * forms the child mass with the same left-to-right products as the owner
 * would (main.cu:214-266): the queue receives the same items in the same
 * slots, every store instruction covers consecutive slots, and no lane idles
 * while another works through a long path.  Every lane of the wave calls
 * it. */
/* Orders one wave's LDS accesses across its lanes (a wave barrier between
 * wavefront-scope fences): lanes that read what other lanes of the wave
 * stored in LDS call it in between. */
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* `k` is the lane's number of stored children, `hops` its UNKNOWN hops on
 * uh[] and bit j of `lmask` marks child j (src side first) as stored; when
 * `sel` is false every child is stored (k == hops) and lmask is unused.
 * The r-th stored child of an owner is its child j for the r-th set bit j
 * of its mask. */
template <int W, class EX>
__device__ inline void emit_children_wave(const EX& ex, const QBuf& out, u64 qstride,
                                          u64 out_base, u32 seg_cap, u32 wbase, int incl, int k,
                                          int hops, u64 lmask, bool sel, const u64 (&known)[W],
                                          const u64 (&up)[W], double mass, Acc2& accE,
                                          u32 (&c)[C_N]) {
  const int lane = threadIdx.x & 63;
  const int total = __shfl(incl, 63, 64);
  wave_lds_sync(); /* the owners' hop stacks and rows, stored by their lanes */
  for (int c0 = 0; c0 < total; c0 += 64) {
    const int cc = c0 + lane;
    const int target = cc < total ? cc : total - 1;
    int lo = 0, hi = 63; /* first lane whose inclusive prefix exceeds target */
#pragma unroll
    for (int step = 0; step < 6; ++step) {
      const int mid = (lo + hi) >> 1;
      const int im = __shfl(incl, mid, 64);
      if (im > target)
        hi = mid;
      else
        lo = mid + 1;
    }
    const int o = lo;
    const int so = __shfl(k, o, 64);
    const int ko = sel ? __shfl(hops, o, 64) : so;
    const u64 lmo = sel ? __shfl(lmask, o, 64) : 0ull;
    const int io = __shfl(incl, o, 64);
    u64 ok[W], ou[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      ok[w] = __shfl(known[w], o, 64);
      ou[w] = __shfl(up[w], o, 64);
    }
    const double om = __shfl(mass, o, 64);
    if (cc < total) {
      int j = cc - (io - so); /* child index from the src side */
      if (sel) {
        u64 m = lmo;
        for (int r = 0; r < j; ++r) m &= m - 1;
        j = __builtin_ctzll(m);
      }
      const int delta = o - lane;
      const u16* uho = ex.uh + delta;
      double run = om;
      u64 pre[W];
#pragma unroll
      for (int w = 0; w < W; ++w) pre[w] = 0;
      for (int i = 0; i < j; ++i) {
        const int e = uho[(ko - 1 - i) * ex.stride];
        run = run * ex.p[e];
#pragma unroll
        for (int w = 0; w < W; ++w)
          if (w == (e >> 6)) pre[w] |= 1ull << (e & 63);
      }
      const int e = uho[(ko - 1 - j) * ex.stride];
      const int wi = e >> 6;
      const u64 bit = 1ull << (e & 63);
      const double cm = run * ex.q[e];
      const u32 off = wbase + (u32)cc;
      if (off < seg_cap) {
        u64 ck[W], cu[W];
#pragma unroll
        for (int w = 0; w < W; ++w) {
          ck[w] = ok[w] | pre[w] | (w == wi ? bit : 0ull);
          cu[w] = ou[w] | pre[w];
        }
        store_item<W>(out, qstride, out_base + off, ck, cu, cm);
        ex.store_child_of(out, qstride, out_base + off, e, delta);
      } else { /* segment full: main.cu:245-252 */
        accE.add(cm);
        ++c[C_PATHS];
        ++c[C_OVER];
      }
    }
  }
}

/* Success term of a live item: mass times the p of all its UNKNOWN hops,
 * src side first (main.cu:254-262). */
template <class EX>
__device__ inline double success_mass(const EX& ex, double mass, int k) {
  double run = mass;
  for (int j = k - 1; j >= 0; --j) run = run * ex.p[ex.uh[j * ex.stride]];
  return run;
}

/* The children of a live item, src side first: those below the truncation
 * (main.cu:537-543) and those marked in `dead` (dead_children(),
 * main.cu:579-582) are counted now, as the processed nodes they would be
 * when dequeued, instead of being queued; bit j of the result marks child j
 * as queued.  `succ` gets the success mass, the same product as
 * success_mass(). */
template <class EX>
__device__ inline u64 settle_children(const EX& ex, double mass, int k, double eps, u64 dead,
                                      Acc2& accE, Acc2& accD, u32 (&c)[C_N], u32& settled,
                                      double& succ) {
  u64 keep = 0;
  double run = mass;
  for (int j = 0; j < k; ++j) {
    const int e = ex.uh[(k - 1 - j) * ex.stride];
    const double cm = run * ex.q[e];
    if (cm < eps) {
      accE.add(cm);
      ++c[C_NODES];
      ++c[C_TRUNC];
      ++settled;
    } else if ((dead >> j) & 1ull) {
      accD.add(cm);
      ++c[C_NODES];
      ++c[C_DEAD];
      ++settled;
    } else {
      keep |= 1ull << j;
    }
    run = run * ex.p[e];
  }
  succ = run;
  return keep;
}

/* Parent context of an item's children for the feature function. */
template <int W, class EX>
__device__ inline void child_ctx(const LevelArgs& a, const EX& ex, const u64 (&known)[W],
                                 const u64 (&up)[W], double mass, int k, ChildCtx& c) {
  int pk = 0, pu = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    pk += __popcll(known[w]);
    pu += __popcll(up[w]);
  }
  c.inv_e = 1.0f / (float)(a.g.e > 0 ? a.g.e : 1);
  c.inv_n = 1.0f / (float)a.g.n;
  c.inv_k = 1.0f / (float)(k > 0 ? k : 1);
  c.u_p = a.g.e - pk;
  c.w_p = pu;
  c.f_p = pk - pu;
  c.k = k;
  c.len = ex.plen;
  c.src = a.g.src;
  c.dst = a.g.dst;
  c.deg_src = ex.out_degree(a.g.src, known, up);
  c.deg_dst = ex.in_degree(a.g.dst, known, up);
  c.log2_m = flog2(mass);
  float lp = 0.0f;
  for (int j = 0; j < k; ++j) lp += __log2f((float)ex.p[ex.uh[j * ex.stride]]);
  c.lp_total = lp;
  c.eps = a.eps;
}

/* The learned choice of the sweep direction for the wave's items (every
 * lane of the wave calls it; `pick` lanes hold a live item with at least two
 * UNKNOWN hops, whose hops the search left in the forward sweep).  MODE 1
 * (tree log) explores: the alternating rule's direction, flipped for a
 * reproducible pseudo-random quarter of the items, and the record of the
 * item gets the features of the direction taken.  Other
 * modes score both directions on the matrix cores (two 64-row tiles) and
 * sweep from the end with the smaller predicted subtree. */
template <int W, class EX, int MODE>
__device__ inline void choose_sweep(const LevelArgs& a, const EX& ex, bool live, int k,
                                    double mass, const u64 (&known)[W], const u64 (&up)[W],
                                    u32 tag_in, u16* tile) {
  const bool pick = live && k >= 2;
  const int lm = floor_log2_of(mass);
  const int head = choice_headroom(lm, a.eps);
  if (MODE == 1) {
    if (!pick) return;
    const int rev = ((ex.raw.f / ex.period) & 1) ^ explore_bit<W>(known, up);
    if (tag_in < a.log_cap) {
      int x[kChoiceFeat];
      choice_features(ex.raw, rev != 0, lm, head, a.g.e, a.g.lqq, x);
      float* dst = a.log_choice + (u64)tag_in * kChoiceFeat;
#pragma unroll
      for (int i = 0; i < kChoiceFeat; ++i) dst[i] = (float)x[i];
    }
    if (rev) ex.reverse_hops(k);
    return;
  }
  if (__ballot(pick) == 0) return; /* wave-uniform */
  const int lane = threadIdx.x & 63;
  bf16x8_t af[2];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) af[mt] = *(const bf16x8_t*)a.chooser->afrag[mt][lane];
  const u32 sign = a.chooser->sign[lane];
  float y[2];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    int x[kChoiceFeat];
    if (pick) {
      choice_features(ex.raw, d == 1, lm, head, a.g.e, a.g.lqq, x);
    } else {
#pragma unroll
      for (int i = 0; i < kChoiceFeat; ++i) x[i] = 0;
    }
    u32* trow = (u32*)(tile + lane * kFeat);
#pragma unroll
    for (int i = 0; i < kChoiceFeat; i += 2) /* small integers: exact in BF16 */
      trow[i / 2] = (u32)bf16_bits((float)x[i]) | ((u32)bf16_bits((float)x[i + 1]) << 16);
    wave_lds_sync(); /* every lane's row before the tile is read */
    y[d] = mlp_score_tile(af, sign, 0.0f, tile);
    wave_lds_sync(); /* the tile read before it is rewritten */
  }
  if (pick) {
    /* the scores are exact multiples of 2^-8 below 2^16: exact as integers */
    const int alt = (ex.raw.f / ex.period) & 1;
    const int s_alt = (int)(y[alt] * 256.0f), s_other = (int)(y[alt ^ 1] * 256.0f);
    const int rev = s_alt - s_other > a.chooser->margin ? alt ^ 1 : alt;
    if (rev) ex.reverse_hops(k);
  }
}

/* One tree level: consume buf[parity], append children to the workgroup's
 * segment of buf[parity^1].  MODE 0: level order; MODE 4: max-mass class
 * rounds; MODE 1: level order that also writes the tree log; MODE 2: class
 * rounds on the learned score (children scored 64 at a time on the matrix
 * cores); MODE 3: the forkBest rule (two-way branch) in either order. */
/* The items of one level (buf[parity], n_items of them, segment prefix sp)
 * expanded into the workgroup's segment of buf[parity ^ 1]: the body of
 * k_expand_level.  Sums and counts go to the caller's accumulators. */
template <int W, class EX, int MODE>
__device__ __forceinline__ void expand_level_items(const LevelArgs& a, const EX& ex, int parity,
                                                   int level, const u32* sp, u32 n_items,
                                                   u32* s_cnt, u32* s_hist, u16* tile, u64 t_end,
                                                   const bf16x8_t (&afrag)[2], u32 sign, float b2,
                                                   float alpha, Acc2& accR, Acc2& accE,
                                                   Acc2& accD, u32 (&c)[C_N], u32& settled) {
  const bool mass_on = MODE == 4 || ((MODE == 2 || MODE == 3) && a.mass_mode);
  /* level order: children below the truncation (any form) and dead ones
   * (row form, kSettleDead) counted at creation */
  const bool settle = MODE == 0 && a.settle != 0;
  const bool settle_dead = settle && EX::kSettle && (a.settle & kSettleDead);
  const QBuf in = a.buf[parity];
  const QBuf out = a.buf[parity ^ 1];
  const int lane = threadIdx.x & 63;
  const u64 wave_id = (u64)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const u64 n_waves = (u64)(gridDim.x * blockDim.x) >> 6;
  const u64 out_base = (u64)blockIdx.x * a.seg_cap;

  /* Small levels are spread over every wave (chunk < 64 lanes) so each
   * workgroup receives, and appends to its segment, a fair share. */
  const u64 per_wave = (n_items + n_waves - 1) / n_waves;
  const u64 chunk = per_wave >= 64 ? 64 : per_wave;
  for (u64 base = wave_id * chunk; base < n_items; base += n_waves * chunk) {
    const u32 i = (u32)(base + lane);
    u64 known[W], up[W];
    double mass = 0.0;
    int k = 0, stored = 0;
    bool live = false, carry = false;
    u64 slot_in = 0;
    u32 tag_in = 0xffffffffu;
    int fork_e = 0;
    u64 lmask = 0;     /* settle: the children that are queued */
    double succ = 0.0; /* settle: the success mass */
    /* past the time budget: the wave's items stay pending (carried) */
    const bool late = t_end && wall_clock64() >= t_end;
    const int s_base = a.n_seg <= 4096 ? wave_segment_of(sp, a.n_seg, (u32)base)
                                       : segment_of(sp, a.n_seg, (u32)base);
    if ((u64)lane < chunk && i < n_items) {
      int s = s_base;
      while (sp[s + 1] <= i) ++s;
      slot_in = (u64)s * a.seg_cap + (i - sp[s]);
      load_item<W>(in, a.stride, slot_in, known, up, mass);
      /* level order: the rows travel with the planes (one memory round) */
      if (MODE == 0 && !late) ex.load(in, a.stride, slot_in);
      if (MODE == 1 || MODE == 2) tag_in = a.tag[parity][slot_in];
      const int cl = !mass_on ? 0 : MODE == 2 ? (int)tag_in : mass_class(mass);
      if (late) {
        carry = true;
        stored = 1;
      } else if (mass_on && cl >= a.drop_class) { /* given up: queue over capacity */
        accE.add(mass);
        ++c[C_FOLD];
      } else if (mass_on && cl > a.thr_class) { /* lighter: carried */
        carry = true;
        stored = 1;
      } else {
        ++c[C_NODES];
        if (mass < a.eps) { /* main.cu:537-543 */
          accE.add(mass);
          ++c[C_TRUNC];
        } else if (MODE == 3) { /* forkBest: success, dead, or two children */
          ex.load(in, a.stride, slot_in);
          const int r = ex.fork_item(known, up);
          if (r > 0) {
            live = true;
            k = 2;
            fork_e = r - 1;
            stored = 2;
            if (mass_on)
              stored = (mass * ex.p[fork_e] >= a.eps ? 1 : 0) + (mass * ex.q[fork_e] >= a.eps ? 1 : 0);
          } else if (r == 0) { /* a WORKING path: main.cu:550-567 */
            accR.add(mass);
            ++c[C_PATHS];
          } else {
            if (r == -2) ++c[C_ERR];
            accD.add(mass); /* main.cu:579-582 */
            ++c[C_DEAD];
          }
        } else {
          if (MODE != 0) ex.load(in, a.stride, slot_in);
          if constexpr (MODE == 0 && EX::kSettle)
            k = settle_dead ? ex.find_item_marked(known, up) : ex.find_item(known, up);
          else
            k = ex.find_item(known, up);
          if (k >= 0) {
            live = true;
            stored = k;
            if constexpr (MODE == 0 && EX::kRowBytes > 0) { /* child-major emission */
              if (settle) {
                u64 dead = 0;
                if constexpr (EX::kSettle)
                  if (settle_dead) dead = ex.dead_children();
                lmask = settle_children(ex, mass, k, a.eps, dead, accE, accD, c, settled, succ);
                stored = __popcll(lmask);
              }
            }
            if ((mass_on || (settle && EX::kRowBytes == 0)) && !EX::kChoose) {
              /* children below eps are not queued */
              stored = 0;
              double run = mass;
              for (int j = k - 1; j >= 0; --j) {
                const int e = ex.uh[j * ex.stride];
                const double cm = run * ex.q[e];
                if (cm >= a.eps) ++stored;
                run = run * ex.p[e];
              }
            }
          } else {
            if (k == -2) ++c[C_ERR];
            accD.add(mass); /* main.cu:579-582 */
            ++c[C_DEAD];
            k = 0;
          }
        }
      }
    }
    if constexpr (EX::kChoose) {
      choose_sweep<W, EX, MODE>(a, ex, live, k, mass, known, up, tag_in, tile);
      if ((mass_on || settle) && live) { /* children below eps are not queued */
        stored = 0;
        double run = mass;
        for (int j = k - 1; j >= 0; --j) {
          const int e = ex.uh[j * ex.stride];
          if (run * ex.q[e] >= a.eps) ++stored;
          run = run * ex.p[e];
        }
      }
    }
    const int incl = wave_incl_scan(stored);
    const int total = __shfl(incl, 63, 64);
    u32 wbase = 0;
    u64 rbase = 0;
    if (total) {
      if (lane == 63) {
        wbase = atomicAdd(s_cnt, (u32)total);
        if (MODE == 1)
          rbase = atomicAdd((unsigned long long*)a.log_count, (unsigned long long)total);
      }
      wbase = __shfl(wbase, 63, 64);
      if (MODE == 1) rbase = __shfl(rbase, 63, 64);
    }
    u32 off = wbase + (u32)(incl - stored);
    if (carry) {
      if (off < a.seg_cap) {
        store_item<W>(out, a.stride, out_base + off, known, up, mass);
        ex.carry(in, slot_in, out, out_base + off, a.stride);
        if (MODE == 2) a.tag[parity ^ 1][out_base + off] = tag_in;
        if (mass_on) atomicAdd(&s_hist[MODE == 2 ? (int)tag_in : mass_class(mass)], 1u);
      } else {
        accE.add(mass);
        ++c[C_FOLD];
      }
    }
    ChildCtx cx;
    if ((MODE == 1 || MODE == 2) && live) child_ctx<W, EX>(a, ex, known, up, mass, k, cx);
    if (MODE == 3) {
      if (live) { /* WORKING child, then FAILED child (main.cu:355-427) */
        const int wi = fork_e >> 6;
        const u64 bit = 1ull << (fork_e & 63);
#pragma unroll
        for (int ch = 0; ch < 2; ++ch) {
          const double cm = ch == 0 ? mass * ex.p[fork_e] : mass * ex.q[fork_e];
          if (mass_on && cm < a.eps) { /* settled at creation */
            accE.add(cm);
            ++c[C_NODES];
            ++c[C_TRUNC];
            continue;
          }
          if (off < a.seg_cap) {
            u64 ck[W], cu[W];
#pragma unroll
            for (int w = 0; w < W; ++w) {
              ck[w] = known[w] | (w == wi ? bit : 0ull);
              cu[w] = up[w] | (ch == 0 && w == wi ? bit : 0ull);
            }
            store_item<W>(out, a.stride, out_base + off, ck, cu, cm);
            if (ch == 0)
              ex.store_rows(out, a.stride, out_base + off);
            else
              ex.store_child(out, a.stride, out_base + off, fork_e);
            if (mass_on) atomicAdd(&s_hist[mass_class(cm)], 1u);
          } else { /* segment full: main.cu:245-252 */
            accE.add(cm);
            ++c[C_PATHS];
            ++c[C_OVER];
          }
          ++off;
        }
      }
    } else if (MODE == 2) {
      /* children in rounds, one per lane per round, src side first; every
       * lane of the wave takes part in each round's tile score */
      int kmax = live ? k : 0;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) kmax = max(kmax, __shfl_xor(kmax, o, 64));
      double run = mass;
      float lp_prefix = 0.0f;
      u64 pre[W];
#pragma unroll
      for (int w = 0; w < W; ++w) pre[w] = 0;
      for (int t = 0; t < kmax; ++t) {
        const bool has = live && t < k;
        int e = 0;
        double cm = 0.0;
        float f[kFeat];
#pragma unroll
        for (int x = 0; x < kFeat; ++x) f[x] = 0.0f;
        if (has) {
          e = ex.uh[(k - 1 - t) * ex.stride];
          cm = run * ex.q[e];
          const u32 xy = ex.ep[e];
          child_features(cx, t, (int)(xy & 0xffffu), (int)(xy >> 16), ex.p[e], ex.q[e], cm,
                         lp_prefix, f);
        }
        u32* trow = (u32*)(tile + lane * kFeat);
#pragma unroll
        for (int x = 0; x < kFeat; x += 2)
          trow[x / 2] = (u32)bf16_bits(f[x]) | ((u32)bf16_bits(f[x + 1]) << 16);
        wave_lds_sync(); /* every lane's row before the tile is read */
        const float y = mlp_score_tile(afrag, sign, b2, tile);
        wave_lds_sync(); /* the tile read before the next round rewrites it */
        if (has) {
          const int wi = e >> 6;
          const u64 bit = 1ull << (e & 63);
          if (cm < a.eps) { /* settled at creation: counted as a truncated node */
            accE.add(cm);
            ++c[C_NODES];
            ++c[C_TRUNC];
          } else {
            if (off < a.seg_cap) {
              u64 ck[W], cu[W];
#pragma unroll
              for (int w = 0; w < W; ++w) {
                ck[w] = known[w] | pre[w] | (w == wi ? bit : 0ull);
                cu[w] = up[w] | pre[w];
              }
              store_item<W>(out, a.stride, out_base + off, ck, cu, cm);
              ex.store_child(out, a.stride, out_base + off, e);
              const int cl = score_class(flog2(cm), y, alpha, kClasses);
              a.tag[parity ^ 1][out_base + off] = (u32)cl;
              atomicAdd(&s_hist[cl], 1u);
            } else { /* segment full: main.cu:245-252 */
              accE.add(cm);
              ++c[C_PATHS];
              ++c[C_OVER];
            }
            ++off;
          }
          run = run * ex.p[e];
          lp_prefix += __log2f((float)ex.p[e]);
#pragma unroll
          for (int w = 0; w < W; ++w)
            if (w == wi) pre[w] |= bit;
        }
      }
      if (live) {
        accR.add(run); /* success term, main.cu:254-262 */
        ++c[C_PATHS];
      }
    } else if (MODE == 0 && EX::kRowBytes > 0) { /* children with rows: child-major */
      /* `late` is wave-uniform: a late wave only carries, and has no children */
      if (!late)
        emit_children_wave<W, EX>(ex, out, a.stride, out_base, a.seg_cap, wbase, incl, stored, k,
                                  lmask, settle, known, up, mass, accE, c);
      if (live) {
        accR.add(settle ? succ : success_mass(ex, mass, k));
        ++c[C_PATHS];
      }
    } else if (live) {
      /* main.cu:214-266: child j has the earlier UNKNOWN hops WORKING and
       * hop j FAILED; masses multiply left to right along the path */
      double run = mass;
      float lp_prefix = 0.0f;
      u64 pre[W];
#pragma unroll
      for (int w = 0; w < W; ++w) pre[w] = 0;
      for (int j = k - 1; j >= 0; --j) {
        const int e = ex.uh[j * ex.stride];
        const int wi = e >> 6;
        const u64 bit = 1ull << (e & 63);
        const double cm = run * ex.q[e];
        bool queue = true;
        if (mass_on || settle) {
          if (cm < a.eps) { /* settled at creation: counted as a truncated node */
            accE.add(cm);
            ++c[C_NODES];
            ++c[C_TRUNC];
            ++settled;
            queue = false;
          }
        }
        if (queue) {
          if (off < a.seg_cap) {
            u64 ck[W], cu[W];
#pragma unroll
            for (int w = 0; w < W; ++w) {
              ck[w] = known[w] | pre[w] | (w == wi ? bit : 0ull);
              cu[w] = up[w] | pre[w];
            }
            store_item<W>(out, a.stride, out_base + off, ck, cu, cm);
            ex.store_child(out, a.stride, out_base + off, e);
            if (mass_on) atomicAdd(&s_hist[mass_class(cm)], 1u);
            if (MODE == 1) {
              const u64 rid = rbase + (off - wbase);
              a.tag[parity ^ 1][out_base + off] = (u32)rid;
              if (rid < a.log_cap) {
                float f[kFeat];
                const u32 xy = ex.ep[e];
                child_features(cx, k - 1 - j, (int)(xy & 0xffffu), (int)(xy >> 16), ex.p[e],
                               ex.q[e], cm, lp_prefix, f);
                float* dstf = a.log_feat + rid * kFeat;
#pragma unroll
                for (int x = 0; x < kFeat; ++x) dstf[x] = bf16_value(bf16_bits(f[x]));
                a.log_parent[rid] = tag_in;
                a.log_level[rid] = (u32)level + 1u;
                if (a.log_mass) a.log_mass[rid] = cm;
              }
            }
          } else { /* segment full: main.cu:245-252 */
            accE.add(cm);
            ++c[C_PATHS];
            ++c[C_OVER];
          }
          ++off;
        }
        run = run * ex.p[e];
        if (MODE == 1) lp_prefix += __log2f((float)ex.p[e]);
#pragma unroll
        for (int w = 0; w < W; ++w)
          if (w == wi) pre[w] |= bit;
      }
      accR.add(run); /* success term, main.cu:254-262 */
      ++c[C_PATHS];
      if (MODE == 1 && a.log_succ && tag_in < a.log_cap) a.log_succ[tag_in] = run;
    }
  }

}

template <int W, class EX, int MODE>
__global__ void __launch_bounds__(256) k_expand_level(LevelArgs a, int parity, int level) {
  extern __shared__ __align__(16) unsigned char smem[];
  /* class rounds: always in MODE 4 (max-mass order of the path rule), never
   * in MODE 0 (level order) and MODE 1, by the argument in MODE 2 and 3 */
  const bool mass_on = MODE == 4 || ((MODE == 2 || MODE == 3) && a.mass_mode);
  /* a launch queued past the one that found the queue empty for good and
   * computed the totals has nothing to do */
  if (a.fin_on_empty && __hip_atomic_load(&a.ctl->fin_done, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT))
    return;
  u32* sp = (u32*)(smem + a.lay.off_extra);
  u32* s_cnt = sp + a.n_seg + 1;
  u32* part = s_cnt + 1;
  const u32 pre = prefetch_slot(a);
  /* workgroup 0 adds up the children settled at this level's creation
   * (level width, below): 16 loads per lane in flight together */
  const bool count_settled = a.settle && blockIdx.x == 0 && threadIdx.x < 64;
  constexpr int kPend = 16;
  u32 pv[kPend];
  if (count_settled) {
#pragma unroll
    for (int j = 0; j < kPend; ++j) {
      const int s = (int)threadIdx.x + 64 * j;
      pv[j] = s < a.n_seg ? a.seg_settle[parity][s] : 0u;
    }
  }
  /* the graph tables are staged while the segment counts are in flight */
  block_prefix(a.seg_count[parity], a.n_seg, sp, part, [&] {
    if (EX::kStage) stage_tables(a.g, a.lay, smem);
  });
  const u32 n_items = sp[a.n_seg];
  u64 pend = 0;
  if (count_settled) { /* wave 0 of workgroup 0, every lane */
    u32 t = 0;
#pragma unroll
    for (int j = 0; j < kPend; ++j) t += pv[j];
    for (int s = (int)threadIdx.x + 64 * kPend; s < a.n_seg; s += 64) t += a.seg_settle[parity][s];
    pend = wave_total_u32(t);
  }
  const u64 t_end = a.budget_ticks ? a.ctl->start_tick + a.budget_ticks : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) { /* count the level (main.cu:970-973) */
    a.ctl->last_items = n_items;
    u64 it = a.ctl->iterations, mw = a.ctl->max_width;
    /* the level's items: those queued and those settled when created */
    const u64 width = (u64)n_items + pend;
    if (width && !(t_end && wall_clock64() >= t_end)) {
      it += 1;
      if (width > mw) mw = width;
      a.ctl->iterations = it;
      a.ctl->max_width = mw;
    }
    /* a launch queued past the one that computed the totals leaves the
     * mirror alone: the host may already be on its next solve */
    if (a.hmirror && !a.ctl->fin_done) {
      a.hmirror[0] = n_items;
      a.hmirror[1] = it;
      a.hmirror[2] = mw;
    }
  }
  if (n_items == 0) {
    if (threadIdx.x == 0 && !a.append) {
      a.seg_count[parity ^ 1][blockIdx.x] = 0;
      if (a.settle) a.seg_settle[parity ^ 1][blockIdx.x] = 0;
    }
    if (mass_on && !a.append)
      for (int i = threadIdx.x; i < kClasses; i += blockDim.x)
        a.hist[(u64)blockIdx.x * kClasses + i] = 0;
    if (a.fin_on_empty && !mass_on && blockIdx.x < kFinMax) {
      /* the queue is empty for good: every slot is final */
      u32* s_flag = s_cnt;
      if (threadIdx.x == 0) *s_flag = a.ctl->fin_done;
      __syncthreads();
      const bool done = *s_flag != 0;
      __syncthreads();
      if (!done) finalize_slots(a, (u64)a.n_seg, s_flag);
    }
    return;
  }
  /* [kClasses] class counts: class rounds and the learned order only (the
   * host sizes the LDS area without them otherwise) */
  u32* s_hist = part + blockDim.x;
  u32* s_set = s_hist + (mass_on || MODE == 2 ? kClasses : 0); /* children settled at creation */
  if (threadIdx.x == 0) {
    *s_cnt = a.append ? a.seg_count[parity ^ 1][blockIdx.x] : 0u;
    *s_set = 0;
  }
  /* per-wave feature tile [64][kFeat] bf16 (learned order), 16-byte aligned */
  u16* tile = (u16*)(((size_t)(s_set + 1) + 15) & ~(size_t)15) +
              (threadIdx.x >> 6) * 64 * kFeat;
  if (mass_on)
    for (int i = threadIdx.x; i < kClasses; i += blockDim.x)
      s_hist[i] = a.append ? a.hist[(u64)blockIdx.x * kClasses + i] : 0u;
  __syncthreads();

  EX ex;
  ex.bind_graph(smem, a.lay, a.g, threadIdx.x);

  const int lane = threadIdx.x & 63;
  bf16x8_t afrag[2];
  u32 sign = 0;
  float b2 = 0.0f, alpha = 0.0f;
  if (MODE == 2) {
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) afrag[mt] = *(const bf16x8_t*)a.mlp->afrag[mt][lane];
    sign = a.mlp->sign[lane];
    b2 = a.mlp->b2;
    alpha = a.mlp->alpha;
  }

  Acc2 accR, accE, accD;
  accR.init();
  accE.init();
  accD.init();
  u32 c[C_N] = {0, 0, 0, 0, 0, 0, 0, 0};
  u32 settled = 0;

  expand_level_items<W, EX, MODE>(a, ex, parity, level, sp, n_items, s_cnt, s_hist, tile, t_end,
                                  afrag, sign, b2, alpha, accR, accE, accD, c, settled);
  if (MODE == 0 && a.settle) {
    const u32 t = wave_total_u32(settled);
    if ((threadIdx.x & 63) == 0 && t) atomicAdd(s_set, t);
  }

  /* per-workgroup flush; then publish this workgroup's segment count */
  __syncthreads();
  if (mass_on)
    for (int i = threadIdx.x; i < kClasses; i += blockDim.x)
      a.hist[(u64)blockIdx.x * kClasses + i] = s_hist[i];
  block_flush(a, accR, accE, accD, c, (unsigned char*)part, &pre);
  if (threadIdx.x == 0) {
    const u32 n = *s_cnt;
    a.seg_count[parity ^ 1][blockIdx.x] = n < a.seg_cap ? n : a.seg_cap;
    if (a.settle) a.seg_settle[parity ^ 1][blockIdx.x] = *s_set;
  }
}

/* Class rounds (max-mass order, or learned order with kTag: the class is
 * the item's tag): the pool in buffer `parity` split before a round in
 * which many items are not expanded.  Workgroup w takes segment w:
 * an item of class >= drop_class is given up to ESP (counted folded, as the
 * round itself would), an item of a class in (thr_class, drop_class) is
 * carried to segment w of the output buffer, and an item to expand is moved
 * to the front of its own segment, in order.  The round's expand launch
 * (append) then finds only items to expand, so every lane of a wave expands
 * one, and appends the children after the carried items.  Items, children
 * and counts are those of the round without the split; only which wave
 * expands which item changes.  Rows (RowNS: the row form's node-set type,
 * void for items without rows) move with their items through the lane's
 * LDS row slice. */
template <int W, typename RowNS, bool kTag>
__global__ void __launch_bounds__(256) k_class_split(LevelArgs a, int parity) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr bool kRows = !std::is_void<RowNS>::value;
  using NS = typename std::conditional<kRows, RowNS, u32>::type;
  u32* s_w = (u32*)(smem + a.lay.off_extra); /* [4] expand, [4] carry counts per wave */
  u32* s_hist = s_w + 8;                     /* [kClasses] carried items per class */
  u32* part = s_hist + kClasses;             /* block_flush scratch */
  const u32 pre = prefetch_slot(a);
  for (int i = threadIdx.x; i < kClasses; i += blockDim.x) s_hist[i] = 0;
  const QBuf in = a.buf[parity];
  const QBuf out = a.buf[parity ^ 1];
  const u32 cnt = a.seg_count[parity][blockIdx.x];
  const u32 n = cnt < a.seg_cap ? cnt : a.seg_cap;
  const u64 base = (u64)blockIdx.x * a.seg_cap;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int n_w = (blockDim.x + 63) >> 6;
  /* queue words of an item's rows: two 16-bit rows per 32-bit word when
   * n <= 16 (Expander::packed), else one per node */
  const int nn = sizeof(NS) == 4 && a.g.n <= 16 ? (a.g.n + 1) >> 1 : a.g.n;
  NS* my = (NS*)(smem + a.lay.off_rows) + threadIdx.x; /* this lane's row slice */
  const u64 below = (1ull << lane) - 1ull;
  Acc2 accR, accE, accD;
  accR.init();
  accE.init();
  accD.init();
  u32 c[C_N] = {0, 0, 0, 0, 0, 0, 0, 0};
  u32 keep = 0, carried = 0; /* workgroup-uniform */
  __syncthreads();
  for (u32 c0 = 0; c0 < n; c0 += blockDim.x) {
    const u32 i = c0 + threadIdx.x;
    u64 known[W], up[W];
    double mass = 0.0;
    int role = 0, cl = 0; /* 1 expand, 2 carry, 3 given up */
    u32 tag = 0;
    if (i < n) {
      load_item<W>(in, a.stride, base + i, known, up, mass);
      if (kTag) tag = a.tag[parity][base + i];
      /* the planes in registers before the barrier below: a use makes the
       * loads complete here, whatever the barrier waits for */
#pragma unroll
      for (int w = 0; w < W; ++w) asm volatile("" ::"v"(known[w]), "v"(up[w]));
      cl = kTag ? (int)tag : mass_class(mass);
      role = cl >= a.drop_class ? 3 : cl > a.thr_class ? 2 : 1;
      if (role == 3) { /* queue over capacity: given up to ESP */
        accE.add(mass);
        ++c[C_FOLD];
      } else if (kRows) {
        const NS* r = (const NS*)in.rows + base + i;
        for (int u = 0; u < nn; ++u) my[u * blockDim.x] = r[(u64)u * a.stride];
      }
    }
    const u64 be = __ballot(role == 1), bc = __ballot(role == 2);
    if (lane == 0) {
      s_w[wave] = (u32)__popcll(be);
      s_w[4 + wave] = (u32)__popcll(bc);
    }
    /* every item of the chunk is read (the barrier waits for the loads)
     * before any is written: an item moves to a slot at or below its own,
     * possibly one another lane of the chunk reads */
    __syncthreads();
    u32 pe = keep, pc = carried, te = 0, tc = 0;
    for (int w = 0; w < n_w; ++w) {
      const u32 xe = s_w[w], xc = s_w[4 + w];
      if (w < wave) {
        pe += xe;
        pc += xc;
      }
      te += xe;
      tc += xc;
    }
    if (role == 1) {
      const u64 d = base + pe + (u32)__popcll(be & below);
      store_item<W>(in, a.stride, d, known, up, mass);
      if (kTag) a.tag[parity][d] = tag;
      if (kRows) {
        NS* r = (NS*)in.rows + d;
        for (int u = 0; u < nn; ++u) r[(u64)u * a.stride] = my[u * blockDim.x];
      }
    } else if (role == 2) {
      const u64 d = base + pc + (u32)__popcll(bc & below);
      store_item<W>(out, a.stride, d, known, up, mass);
      if (kTag) a.tag[parity ^ 1][d] = tag;
      if (kRows) {
        NS* r = (NS*)out.rows + d;
        for (int u = 0; u < nn; ++u) r[(u64)u * a.stride] = my[u * blockDim.x];
      }
      atomicAdd(&s_hist[cl], 1u);
    }
    keep += te;
    carried += tc;
    __syncthreads(); /* s_w and the row slices are reused by the next chunk */
  }
  block_flush(a, accR, accE, accD, c, (unsigned char*)part, &pre);
  for (int i = threadIdx.x; i < kClasses; i += blockDim.x)
    a.hist[(u64)blockIdx.x * kClasses + i] = s_hist[i];
  if (threadIdx.x == 0) {
    a.seg_count[parity][blockIdx.x] = keep;
    a.seg_count[parity ^ 1][blockIdx.x] = carried;
  }
}

/* One level of a workgroup's own items: items in_base .. in_base + n - 1 of
 * `in` are expanded, one lane each, and their children appended to `out` at
 * out_base + 0 .. out_cap - 1 in the order the waves reserve room (s_cnt[0],
 * zero on entry, counts every queued child, those past out_cap included:
 * they go to ESP as overflowed, main.cu:245-252); children settled at
 * creation (a.settle) are added to s_cnt[1].  The whole workgroup calls
 * it. */
template <int W, class EX>
__device__ inline void group_pass(const LevelArgs& a, const EX& ex, const QBuf& in, u64 in_base,
                                  u32 n, const QBuf& out, u64 out_base, u32 out_cap, u32* s_cnt,
                                  Acc2& accR, Acc2& accE, Acc2& accD, u32 (&c)[C_N]) {
  const int lane = threadIdx.x & 63;
  const bool settle = a.settle != 0;
  const bool settle_dead = settle && EX::kSettle && (a.settle & kSettleDead);
  u32 settled = 0;
  for (u32 base = 0; base < n; base += blockDim.x) {
    const u32 i = base + threadIdx.x;
    u64 known[W], up[W];
    double mass = 0.0;
    int k = 0, hops = 0; /* children queued; UNKNOWN hops */
    bool live = false;
    u64 lmask = 0;
    double succ = 0.0;
    if (i < n) {
      load_item<W>(in, a.stride, in_base + i, known, up, mass);
      ex.load(in, a.stride, in_base + i); /* with the planes: one memory round */
      ++c[C_NODES];
      if (mass < a.eps) { /* main.cu:537-543 */
        accE.add(mass);
        ++c[C_TRUNC];
      } else {
        if constexpr (EX::kSettle)
          k = settle_dead ? ex.find_item_marked(known, up) : ex.find_item(known, up);
        else
          k = ex.find_item(known, up);
        if (k >= 0) {
          live = true;
          hops = k;
          if (settle) {
            u64 dead = 0;
            if constexpr (EX::kSettle)
              if (settle_dead) dead = ex.dead_children();
            lmask = settle_children(ex, mass, hops, a.eps, dead, accE, accD, c, settled, succ);
            k = __popcll(lmask);
          }
        } else {
          if (k == -2) ++c[C_ERR];
          accD.add(mass); /* main.cu:579-582 */
          ++c[C_DEAD];
          k = 0;
        }
      }
    }
    const int incl = wave_incl_scan(k);
    const int total = __shfl(incl, 63, 64);
    u32 wbase = 0;
    if (total) {
      if (lane == 63) wbase = atomicAdd(s_cnt, (u32)total);
      wbase = __shfl(wbase, 63, 64);
    }
    if (EX::kRowBytes > 0) {
      emit_children_wave<W, EX>(ex, out, a.stride, out_base, out_cap, wbase, incl, k, hops, lmask,
                                settle, known, up, mass, accE, c);
      if (live) {
        accR.add(settle ? succ : success_mass(ex, mass, hops));
        ++c[C_PATHS];
      }
    } else if (live) { /* main.cu:214-266 */
      u32 off = wbase + (u32)(incl - k);
      double run = mass;
      u64 pre[W];
#pragma unroll
      for (int w = 0; w < W; ++w) pre[w] = 0;
      for (int j = hops - 1; j >= 0; --j) {
        const int e = ex.uh[j * ex.stride];
        const int wi = e >> 6;
        const u64 bit = 1ull << (e & 63);
        const double cm = run * ex.q[e];
        /* children settled at creation were counted by settle_children() */
        if (!settle || ((lmask >> (hops - 1 - j)) & 1ull)) {
          if (off < out_cap) {
            u64 ck[W], cu[W];
#pragma unroll
            for (int w = 0; w < W; ++w) {
              ck[w] = known[w] | pre[w] | (w == wi ? bit : 0ull);
              cu[w] = up[w] | pre[w];
            }
            store_item<W>(out, a.stride, out_base + off, ck, cu, cm);
          } else { /* region full: main.cu:245-252 */
            accE.add(cm);
            ++c[C_PATHS];
            ++c[C_OVER];
          }
          ++off;
        }
        run = run * ex.p[e];
#pragma unroll
        for (int w = 0; w < W; ++w)
          if (w == wi) pre[w] |= bit;
      }
      accR.add(run); /* success term, main.cu:254-262 */
      ++c[C_PATHS];
    }
  }
  if (settle) {
    const u32 t = wave_total_u32(settled);
    if (lane == 0 && t) atomicAdd(s_cnt + 1, t);
  }
}

/* The first levels of a level-order solve (path rule) in one workgroup:
 * while a level holds at most max_in items (one lane each), the workgroup
 * expands it from segment 0 of one buffer into segment 0 of the other and
 * goes straight on to the next, with no launch, prefix scan or flush in
 * between.  It stops before the first larger level (or after max_levels),
 * leaves that level in buffer 0, segment 0, and flushes its sums to slot 0;
 * the level launches take over from there.  Same counts, levels and widest
 * level as one launch per level. */
template <int W, class EX>
__global__ void __launch_bounds__(256) k_small_levels(LevelArgs a, u32 max_in, int max_levels) {
  extern __shared__ __align__(16) unsigned char smem[];
  u32* s_cnt = (u32*)(smem + a.lay.off_extra);
  u32* part = s_cnt + 4;
  if (EX::kStage) stage_tables(a.g, a.lay, smem);
  EX ex;
  ex.bind_graph(smem, a.lay, a.g, threadIdx.x);
  Acc2 accR, accE, accD;
  accR.init();
  accE.init();
  accD.init();
  u32 c[C_N] = {0, 0, 0, 0, 0, 0, 0, 0};
  int p = 0;
  u32 n = a.seg_count[0][0];
  /* the level's children settled at creation: part of its width */
  u32 pend = a.settle ? a.seg_settle[0][0] : 0u;
  /* levels counted in registers (main.cu:970-973) and written once at the
   * end: a global read-modify-write per level would stall every barrier */
  u64 levels = 0, widest = 0;
  for (int lv = 0; n > 0 && n <= max_in && lv < max_levels; ++lv) {
    if (threadIdx.x == 0) {
      s_cnt[0] = 0;
      s_cnt[1] = 0;
    }
    levels += 1;
    const u64 width = (u64)n + pend;
    widest = width > widest ? width : widest;
    __syncthreads();
    group_pass<W, EX>(a, ex, a.buf[p], 0, n, a.buf[p ^ 1], 0, a.seg_cap, s_cnt, accR, accE, accD,
                      c);
    __syncthreads();
    const u32 made = s_cnt[0];
    pend = s_cnt[1];
    n = made < a.seg_cap ? made : a.seg_cap;
    p ^= 1;
    __syncthreads();
  }
  if (n == 0 && pend) { /* a last level of settled children only */
    levels += 1;
    widest = pend > widest ? pend : widest;
    pend = 0;
  }
  if (p == 1) { /* hand the level over in buffer 0 */
    const QBuf from = a.buf[1];
    const QBuf to = a.buf[0];
    for (u32 i = threadIdx.x; i < n; i += blockDim.x) {
#pragma unroll
      for (int w = 0; w < W; ++w) {
        to.known[(u64)w * a.stride + i] = from.known[(u64)w * a.stride + i];
        to.up[(u64)w * a.stride + i] = from.up[(u64)w * a.stride + i];
      }
      to.mass[i] = from.mass[i];
      ex.carry(from, i, to, i, a.stride);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a.seg_count[0][0] = n;
    if (a.settle) a.seg_settle[0][0] = pend;
    a.ctl->last_items = n;
    a.ctl->iterations += levels;
    if (widest > a.ctl->max_width) a.ctl->max_width = widest;
  }
  block_flush(a, accR, accE, accD, c, (unsigned char*)part);
}

/* Levels of the ramp after k_small_levels: levels that are still too narrow
 * to fill the chip, for which a launch per level would cost mostly its fixed
 * costs.  The level k_small_levels left in buffer 0, segment 0 (n0 items,
 * after `iterations` levels) is dealt out in contiguous slices, one per
 * workgroup; each workgroup then expands its slice's subtrees level after
 * level on its own (group_pass), in a region of `span` segments starting at
 * segment 1 + span * blockIdx.x, alternating between the buffers, with no
 * launch or grid-wide step in between.  Every workgroup runs the same number
 * of levels L: the largest even number for which n0 g^L stays near
 * `target` items, g the mean growth per level so far (n0^(1/iterations)),
 * so the last level ends in buffer 0.  The items entering each level are
 * added up over the workgroups (ctl->ramp_width); the last workgroup to
 * finish counts those levels and the widest of them into the control block
 * as the level launches would (main.cu:970-973), and sets the segment counts
 * that it alone may write.  A workgroup whose region fills sets
 * ctl->ramp_over (and the mirror's ramp word): its later children went to
 * ESP, so the caller solves again without the ramp.  Workgroups never wait
 * for each other. */
template <int W, class EX>
__global__ void __launch_bounds__(256) k_ramp_levels(LevelArgs a, int span, u32 target) {
  extern __shared__ __align__(16) unsigned char smem[];
  const u32 n0 = a.seg_count[0][0];
  const u64 done_levels = a.ctl->iterations;
  if (n0 < 2 || done_levels == 0) return;
  /* the same L in every workgroup: computed from the same words */
  const double g = exp2(log2((double)n0) / (double)done_levels);
  const double want = g > 1.0 ? log2((double)target / (double)n0) / log2(g) : (double)kRampMax;
  int L = want > 0.0 ? 2 * (int)(0.5 * want + 0.5) : 0;
  if (L > kRampMax) L = kRampMax;
  if (L < 2) return;
  u32* s_cnt = (u32*)(smem + a.lay.off_extra);
  u32* part = s_cnt + 4;          /* block_flush scratch: 80 bytes per wave */
  u32* s_width = s_cnt + 4 + 84;  /* [kRampMax + 1] items entering each level */
  const u32 lo = (u32)((u64)blockIdx.x * n0 / gridDim.x);
  const u32 hi = (u32)((u64)(blockIdx.x + 1) * n0 / gridDim.x);
  const u64 region = (u64)(1 + (u64)span * blockIdx.x) * a.seg_cap;
  const u32 region_cap = (u32)span * a.seg_cap;
  u32 n = hi - lo;
  /* children settled at creation of the level being expanded (those of the
   * handed level are added by the last workgroup) */
  u32 pend = 0;
  bool over = false;
  if (n > 0) {
    const u32 pre = prefetch_slot(a);
    if (EX::kStage) stage_tables(a.g, a.lay, smem);
    EX ex;
    ex.bind_graph(smem, a.lay, a.g, threadIdx.x);
    Acc2 accR, accE, accD;
    accR.init();
    accE.init();
    accD.init();
    u32 c[C_N] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int lv = 0; lv < L; ++lv) {
      if (threadIdx.x == 0) {
        s_cnt[0] = 0;
        s_cnt[1] = 0;
        s_width[lv] = n + pend;
      }
      __syncthreads();
      if (n) {
        const QBuf& in = a.buf[lv & 1];
        group_pass<W, EX>(a, ex, in, lv == 0 ? (u64)lo : region, n, a.buf[(lv + 1) & 1], region,
                          region_cap, s_cnt, accR, accE, accD, c);
      }
      __syncthreads();
      const u32 made = s_cnt[0];
      pend = s_cnt[1];
      over = over || made > region_cap;
      n = made < region_cap ? made : region_cap;
      __syncthreads();
    }
    if (threadIdx.x == 0) s_width[L] = n + pend;
    block_flush(a, accR, accE, accD, c, (unsigned char*)part, &pre);
    /* the widths go out once, after the levels (block_flush's barrier
     * orders thread 0's last LDS word before these reads) */
    if ((int)threadIdx.x <= L) {
      const u32 w = s_width[threadIdx.x];
      if (w) atomicAdd(&a.ctl->ramp_width[threadIdx.x], (u64)w);
      if (threadIdx.x == 0 && over) atomicOr(&a.ctl->ramp_over, 1u);
      __threadfence();
    }
    __syncthreads();
  }
  /* segment counts of the region (its items are contiguous from its start) */
  for (int i = threadIdx.x; i < span; i += blockDim.x) {
    const u64 before = (u64)i * a.seg_cap;
    const u64 left = n > before ? n - before : 0;
    a.seg_count[0][1 + (u64)span * blockIdx.x + i] = (u32)(left < a.seg_cap ? left : a.seg_cap);
    if (a.settle) a.seg_settle[0][1 + (u64)span * blockIdx.x + i] = i == 0 ? pend : 0u;
  }
  if (threadIdx.x != 0) return;
  __threadfence();
  if (atomicAdd(&a.ctl->ramp_done, 1u) != gridDim.x - 1) return;
  /* the last workgroup: every other one has read n0 and added its widths */
  __threadfence();
  u64 it = a.ctl->iterations, mw = a.ctl->max_width;
  for (int lv = 0; lv < L; ++lv) {
    u64 w = __hip_atomic_load(&a.ctl->ramp_width[lv], __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
    if (lv == 0 && a.settle) w += a.seg_settle[0][0];
    if (w) {
      it += 1;
      if (w > mw) mw = w;
    }
  }
  const u32 ov = __hip_atomic_load(&a.ctl->ramp_over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  a.ctl->iterations = it;
  a.ctl->max_width = mw;
  a.ctl->last_items =
      __hip_atomic_load(&a.ctl->ramp_width[L], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  a.seg_count[0][0] = 0;
  for (u64 s = 1 + (u64)span * gridDim.x; s < (u64)a.n_seg; ++s) a.seg_count[0][s] = 0;
  if (a.settle) {
    a.seg_settle[0][0] = 0;
    for (u64 s = 1 + (u64)span * gridDim.x; s < (u64)a.n_seg; ++s) a.seg_settle[0][s] = 0;
  }
  if (a.hmirror) a.hmirror[kMirRamp] = ov;
}

/* One expansion of each caller-supplied item into fixed output slots (the
 * parity entry point nr_expand_items). */
struct ItemsArgs {
  DevGraph g;
  Layout lay;
  const u64* known; /* [count][w] */
  const u64* up;
  const double* mass;
  int count, max_children;
  int* nchild;
  double* succ;
  u64* ck; /* [count][max_children][w] */
  u64* cu;
  double* cm;
};

template <int W, class EX>
__global__ void __launch_bounds__(256) k_expand_items(ItemsArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  if (EX::kStage) stage_tables(a.g, a.lay, smem);
  __syncthreads();
  EX ex;
  ex.bind_graph(smem, a.lay, a.g, threadIdx.x);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.count) return;
  u64 known[W], up[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    known[w] = a.known[(u64)i * W + w];
    up[w] = a.up[(u64)i * W + w];
  }
  const double mass = a.mass[i];
  ex.derive(known, up);
  int k = ex.find_item(known, up);
  if (k < 0) {
    a.nchild[i] = k;
    a.succ[i] = 0.0;
    return;
  }
  a.nchild[i] = k;
  double run = mass;
  u64 pre[W];
#pragma unroll
  for (int w = 0; w < W; ++w) pre[w] = 0;
  int c = 0;
  for (int j = k - 1; j >= 0; --j, ++c) {
    const int e = ex.uh[j * ex.stride];
    const int wi = e >> 6;
    const u64 bit = 1ull << (e & 63);
    if (c < a.max_children) {
      u64 base = ((u64)i * a.max_children + c) * W;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        a.ck[base + w] = known[w] | pre[w] | (w == wi ? bit : 0ull);
        a.cu[base + w] = up[w] | pre[w];
      }
      a.cm[(u64)i * a.max_children + c] = run * ex.q[e];
    }
    run = run * ex.p[e];
#pragma unroll
    for (int w = 0; w < W; ++w)
      if (w == wi) pre[w] |= bit;
  }
  a.succ[i] = run;
}


typedef void* (*KernelGetter)();

/* Kernels of one width class; batch / refill exist for the row form only. */
struct Variant {
  KernelGetter level, items, batch, refill;
  int ns_bytes; /* bytes per adjacency row carried by an item (0: none) */
  int nw;       /* CsrExpander node-set words (0: Expander) */
  KernelGetter level_log, level_learn; /* MODE 1 / MODE 2 of k_expand_level */
  KernelGetter level_fork;             /* MODE 3 (forkBest rule) */
  KernelGetter seed_rows;              /* row form: rows of resumed items */
  int lean;                            /* items carry no rows (LeanExpander) */
  KernelGetter level_mass;             /* MODE 4 (max-mass class rounds) */
  KernelGetter small;                  /* k_small_levels of this form */
  int choose;                          /* the learned chooser picks the sweep */
  KernelGetter ramp;                   /* k_ramp_levels of this form */
  int settle; /* the level kernel can settle dead children at creation */
  KernelGetter split; /* k_class_split for this form's items (class rounds) */
};

/* The alternating rule with the learned sweep choice (ChooseAltExpander, or
 * CsrChooseAltExpander when nw > 0): MODE 0, 1, 2 and 4 (nr_choose.hip,
 * nr_choose_csr.hip). */
Variant variant_choose(int W, int ns_bytes);
Variant variant_choose_csr(int W, int NW);

/* MODE 4 kernel of a variant (nr_mass.hip), filled into v. */
void add_mass_kernels(Variant& v, int W);

/* k_class_split for items of W plane words carrying rows of row_bytes per
 * node (0: no rows), classed by mass or (tag_class) by their tag,
 * nr_mass.hip. */
KernelGetter split_kernel(int W, int row_bytes, bool tag_class);

/* Lean row form (plain path rule, two terminals) for W words, node sets of
 * NS bytes (nr_lean.hip). */
Variant variant_lean(int W, int ns_bytes);

/* Alternating-sweep rule (AltExpander) for W words and node sets of
 * NS bytes: level, max-mass, tree-log, learned, small-level and
 * single-expansion kernels (nr_alt.hip). */
Variant variant_alt(int W, int ns_bytes);

/* The alternating-sweep rule on the CSR arcs (CsrAltExpander) for W words and
 * NW node-set words (nr_alt_csr.hip); undirected graphs only. */
Variant variant_alt_csr(int W, int NW);

/* MODE 3 kernels of a variant (nr_fork.hip), filled into v. */
void add_fork_kernels(Variant& v, int W);

/* MODE 1 / MODE 2 kernels of a variant (nr_learn.hip), filled into v. */
void add_learn_kernels(Variant& v, int W);

/* CSR form for W (edge-plane words) and node-set words NW (nr_csr.hip). */
Variant variant_csr(int W, int NW);

}  // namespace nr

#endif
