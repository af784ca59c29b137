/*
 * nr_big.hip - the path rule for graphs past the per-lane forms (more than
 * 256 nodes or 512 edges; up to 65535 of each, e.g. the 6301-node SNAP
 * graph p2p-Gn08C): one wave per subproblem.
 *
 * A lane-per-item search needs the item's node sets in registers, which a
 * 6301-node graph cannot afford; here the 64 lanes of a wave share one item:
 * its planes sit in the wave's LDS slice, the frontier is a bitset there
 * and each lane expands the frontier nodes of its bit position in every
 * 64-node word, node levels are bytes, and the walk back tests 64 in-arcs
 * of a node per step with a ballot (the in-arc rows are sorted by (u, e),
 * so the first hit is the lowest-index parent, lowest edge id on ties:
 * device_code.cuh:80-194 made deterministic, as in the other forms).
 * Children (main.cu:214-266) are written by all lanes, one plane word per
 * lane, into the workgroup's segment.  Items are stored item-major:
 * known[W] then up[W] (u64), so a wave's loads and stores are contiguous.
 */
#include "nr_big.h"

namespace nr {

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(256) k_init_big(LevelArgs a, const u64* root, int W,
                                                 u64 n_slots) {
  const u64 tid = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u64 stride = (u64)gridDim.x * blockDim.x;
  if (blockIdx.x == 0) {
    for (u32 i = threadIdx.x; i < sizeof(LevelCtl) / 8; i += blockDim.x) ((u64*)a.ctl)[i] = 0;
    __syncthreads();
  }
  for (u64 i = tid; i < n_slots; i += stride) {
    a.partR[i].hi = a.partR[i].lo = 0.0;
    a.partE[i].hi = a.partE[i].lo = 0.0;
    a.partD[i].hi = a.partD[i].lo = 0.0;
    for (int j = 0; j < C_N; ++j) a.cnt[i * C_N + j] = 0;
  }
  for (u64 s = tid; s < (u64)a.n_seg; s += stride) {
    a.seg_count[0][s] = s == 0 ? 1u : 0u;
    a.seg_count[1][s] = 0u;
    a.seg_settle[0][s] = 0u;
    a.seg_settle[1][s] = 0u;
  }
  for (u64 w = tid; w < 2 * (u64)W; w += stride) a.buf[0].known[w] = root[w];
  if (tid == 0) {
    a.ctl->start_tick = wall_clock64();
    a.ctl->last_items = 1;
    a.buf[0].mass[0] = 1.0;
  }
}

__global__ void __launch_bounds__(256) k_expand_big(LevelArgs a, int parity, BigShape bs) {
  extern __shared__ __align__(16) unsigned char smem[];
  u32* sp = (u32*)(smem + a.lay.off_extra);
  u32* s_cnt = sp + a.n_seg + 1;
  u32* part = s_cnt + 1;
  u32* s_set = part + blockDim.x; /* children settled at creation */
  block_prefix(a.seg_count[parity], a.n_seg, sp, part);
  const u32 n_items = sp[a.n_seg];
  /* the level's children settled when created (below the truncation) */
  u64 pend = 0;
  if (a.settle && blockIdx.x == 0 && threadIdx.x < 64) { /* wave 0, 16 loads in flight */
    u32 t = 0;
    for (int s0 = 0; s0 < a.n_seg; s0 += 64 * 16) {
      u32 v[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const int s = s0 + (int)threadIdx.x + 64 * j;
        v[j] = s < a.n_seg ? a.seg_settle[parity][s] : 0u;
      }
#pragma unroll
      for (int j = 0; j < 16; ++j) t += v[j];
    }
    pend = wave_total_u32(t);
  }
  const u64 t_end = a.budget_ticks ? a.ctl->start_tick + a.budget_ticks : 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) { /* count the level (main.cu:970-973) */
    a.ctl->last_items = n_items;
    const u64 width = (u64)n_items + pend;
    if (width && !(t_end && wall_clock64() >= t_end)) {
      a.ctl->iterations += 1;
      if (width > a.ctl->max_width) a.ctl->max_width = width;
    }
  }
  if (n_items == 0) {
    if (threadIdx.x == 0) {
      a.seg_count[parity ^ 1][blockIdx.x] = 0;
      if (a.settle) a.seg_settle[parity ^ 1][blockIdx.x] = 0;
    }
    return;
  }
  const u32 pre = prefetch_slot(a);
  if (threadIdx.x == 0) {
    *s_cnt = 0;
    *s_set = 0;
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int W = bs.w, n = a.g.n, NW = (n + 63) / 64;
  const int src = a.g.src, dst = a.g.dst;
  unsigned char* wm = smem + bs.off_waves + wave * bs.wave_bytes;
  u64* planes = (u64*)(wm + bs.off_planes); /* known[W], up[W] */
  u64* cur = (u64*)(wm + bs.off_cur);
  u64* nxt = (u64*)(wm + bs.off_nxt);
  unsigned char* lev = wm + bs.off_lev;
  u16* hops = (u16*)(wm + bs.off_hops);
  const u64 n_waves = (u64)gridDim.x * (blockDim.x >> 6);
  const u64 wave_id = (u64)blockIdx.x * (blockDim.x >> 6) + wave;
  const u64 out_base = (u64)blockIdx.x * a.seg_cap;
  const u64 row = 2 * (u64)W;
  const QBuf in = a.buf[parity];
  const QBuf out = a.buf[parity ^ 1];

  Acc2 accR, accE, accD;
  accR.init();
  accE.init();
  accD.init();
  u32 c[C_N] = {0, 0, 0, 0, 0, 0, 0, 0};
  u32 settled = 0; /* lane 0: children settled at creation */

  for (u64 i = wave_id; i < n_items; i += n_waves) { /* one item per wave */
    int s = a.n_seg <= 4096 ? wave_segment_of(sp, a.n_seg, (u32)i) : segment_of(sp, a.n_seg, (u32)i);
    while (sp[s + 1] <= (u32)i) ++s;
    const u64 slot_in = (u64)s * a.seg_cap + (i - sp[s]);
    const u64* ip = in.known + slot_in * row;
    const double mass = in.mass[slot_in];
    if (t_end && wall_clock64() >= t_end) { /* past the budget: carried unexpanded */
      u32 o = 0;
      if (lane == 0) o = atomicAdd(s_cnt, 1u);
      o = __shfl(o, 0, 64);
      if (o < a.seg_cap) {
        u64* op = out.known + (out_base + o) * row;
        for (u64 w = lane; w < row; w += 64) op[w] = ip[w];
        if (lane == 0) out.mass[out_base + o] = mass;
      } else if (lane == 0) {
        accE.add(mass);
        ++c[C_FOLD];
      }
      continue;
    }
    if (lane == 0) ++c[C_NODES];
    if (mass < a.eps) { /* main.cu:537-543 */
      if (lane == 0) {
        accE.add(mass);
        ++c[C_TRUNC];
      }
      continue;
    }
    for (u64 w = lane; w < row; w += 64) planes[w] = ip[w];
    for (int w = lane; w < (n + 7) / 8; w += 64) ((u64*)lev)[w] = ~0ull;
    for (int w = lane; w < NW; w += 64) {
      cur[w] = 0;
      nxt[w] = 0;
    }
    wave_sync();
    if (lane == 0) {
      lev[src] = 0;
      cur[src >> 6] = 1ull << (src & 63);
    }
    wave_sync();
    auto allowed = [&](int e) {
      const u64 b = 1ull << (e & 63);
      return !((planes[e >> 6] & b) && !(planes[W + (e >> 6)] & b));
    };
    /* optimistic BFS from src (device_code.cuh:80-194), whole levels */
    int L = 0;
    bool dead = false, err = false;
    if (src != dst) {
      for (;;) {
        /* The frontier's nodes dealt to all 64 lanes: lane l holds word
         * w0 + l of a block of 64 frontier words, an exclusive prefix of
         * their popcounts numbers the block's nodes, and in round r lane i
         * expands node r*64 + i (owner word by a search over the prefix,
         * then that word's j-th set bit).  Levels are sets, so the order in
         * which nodes are expanded changes nothing. */
        for (int w0 = 0; w0 < NW; w0 += 64) {
          const int wl = w0 + lane;
          const u64 word = wl < NW ? cur[wl] : 0ull;
          const int cnt = __popcll(word);
          const int incl = wave_incl_scan(cnt);
          const int total = __shfl(incl, 63, 64);
          for (int r0 = 0; r0 < total; r0 += 64) {
            const int g = r0 + lane;
            int lo = 0, hi = 63; /* first lane whose inclusive prefix exceeds g */
#pragma unroll
            for (int step = 0; step < 6; ++step) {
              const int mid = (lo + hi) >> 1;
              if (__shfl(incl, mid, 64) > g)
                hi = mid;
              else
                lo = mid + 1;
            }
            const u64 ow = __shfl(word, lo, 64);
            const int oi = __shfl(incl, lo, 64) - __shfl(cnt, lo, 64);
            if (g < total) {
              /* the (g - oi)-th set bit of ow, by halving on popcounts */
              int j = g - oi, base = 0;
              u64 x = ow;
#pragma unroll
              for (int half = 32; half >= 1; half >>= 1) {
                const u64 low = x & ((1ull << half) - 1ull);
                const int c = __popcll(low);
                if (j >= c) {
                  j -= c;
                  x >>= half;
                  base += half;
                } else {
                  x = low;
                }
              }
              const int u = (w0 + lo) * 64 + base;
              const int j1 = a.g.rp[u + 1];
              for (int jj = a.g.rp[u]; jj < j1; jj += 8) { /* eight arc loads in flight */
                u32 xs[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) xs[t] = jj + t < j1 ? a.g.arc[jj + t] : 0u;
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                  if (jj + t >= j1) break;
                  const int v = (int)(xs[t] & 0xffffu), e = (int)(xs[t] >> 16);
                  if (lev[v] != 0xff || !allowed(e)) continue;
                  lev[v] = (unsigned char)(L + 1);
                  atomicOr((unsigned long long*)&nxt[v >> 6], 1ull << (v & 63));
                }
              }
            }
          }
        }
        wave_sync();
        ++L;
        if (lev[dst] != 0xff) break;
        u64 any = 0;
        for (int w = lane; w < NW; w += 64) {
          const u64 x = nxt[w];
          any |= x;
          cur[w] = x;
          nxt[w] = 0;
        }
        wave_sync();
        if (__ballot(any != 0) == 0) {
          dead = true;
          break;
        }
        if (L >= 254) { /* levels are bytes */
          err = true;
          break;
        }
      }
    }
    int k = 0;
    if (!dead && !err) { /* walk back: lowest-index parent, lowest edge id */
      int v = dst;
      for (int l = L; l >= 1; --l) {
        const int j0 = a.g.irp[v], j1 = a.g.irp[v + 1];
        int pu = -1, pe = -1;
        for (int b0 = j0; b0 < j1; b0 += 64) {
          const int j = b0 + lane;
          u32 x = 0;
          bool ok = false;
          if (j < j1) {
            x = a.g.iarc[j];
            ok = lev[x & 0xffffu] == (unsigned char)(l - 1) && allowed((int)(x >> 16));
          }
          const u64 m = __ballot(ok);
          if (m) {
            const u32 xf = __shfl(x, (int)__builtin_ctzll(m), 64);
            pu = (int)(xf & 0xffffu);
            pe = (int)(xf >> 16);
            break;
          }
        }
        if (pu < 0 || k >= bs.hop_cap) {
          err = true;
          break;
        }
        if (!((planes[pe >> 6] >> (pe & 63)) & 1ull)) { /* an UNKNOWN hop */
          if (lane == 0) hops[k] = (u16)pe;
          ++k;
        }
        v = pu;
      }
    }
    wave_sync();
    if (err) {
      if (lane == 0) ++c[C_ERR];
      continue;
    }
    if (dead) { /* main.cu:579-582 */
      if (lane == 0) {
        accD.add(mass);
        ++c[C_DEAD];
      }
      continue;
    }
    /* children src side first (hops[] holds them dst side first): child t
     * has the earlier hops WORKING and hop t FAILED; each hop is made
     * WORKING in the wave's planes once its child is written.  With
     * a.settle a child below the truncation is counted when created
     * (main.cu:537-543) and takes no slot. */
    int kq = k;
    if (a.settle) {
      kq = 0;
      double r0 = mass;
      for (int t = 0; t < k; ++t) {
        const int e = hops[k - 1 - t];
        if (r0 * a.g.q[e] >= a.eps) ++kq;
        r0 = r0 * a.g.p[e];
      }
    }
    u32 obase = 0;
    if (lane == 0 && kq) obase = atomicAdd(s_cnt, (u32)kq);
    obase = __shfl(obase, 0, 64);
    double run = mass;
    u32 o = obase;
    for (int t = 0; t < k; ++t) {
      const int e = hops[k - 1 - t];
      const int we = e >> 6;
      const u64 b = 1ull << (e & 63);
      const double cm = run * a.g.q[e];
      if (a.settle && cm < a.eps) { /* settled at creation */
        if (lane == 0) {
          accE.add(cm);
          ++c[C_NODES];
          ++c[C_TRUNC];
          ++settled;
        }
      } else if (o < a.seg_cap) {
        u64* op = out.known + (out_base + o) * row;
        for (u64 w = lane; w < row; w += 64) op[w] = planes[w] | ((int)w == we ? b : 0ull);
        if (lane == 0) out.mass[out_base + o] = cm;
      } else if (lane == 0) { /* segment full: main.cu:245-252 */
        accE.add(cm);
        ++c[C_PATHS];
        ++c[C_OVER];
      }
      if (!(a.settle && cm < a.eps)) ++o;
      wave_sync();
      if (lane == 0) {
        planes[we] |= b;
        planes[W + we] |= b;
      }
      wave_sync();
      run = run * a.g.p[e];
    }
    if (lane == 0) { /* success term, main.cu:254-262 */
      accR.add(run);
      ++c[C_PATHS];
    }
  }

  if (lane == 0 && settled) atomicAdd(s_set, settled);
  __syncthreads();
  block_flush(a, accR, accE, accD, c, (unsigned char*)part, &pre);
  if (threadIdx.x == 0) {
    const u32 nn = *s_cnt;
    a.seg_count[parity ^ 1][blockIdx.x] = nn < a.seg_cap ? nn : a.seg_cap;
    if (a.settle) a.seg_settle[parity ^ 1][blockIdx.x] = *s_set;
  }
}

void* big_expand_kernel() { return (void*)&k_expand_big; }
void* big_init_kernel() { return (void*)&k_init_big; }

}  // namespace nr
