/*
 * nr_engine.hip - HIP kernels (gfx950, wave64) and host runtime of the
 * two-terminal reliability engine; implements nr_solve / nr_expand_items of
 * include/nr_api.h.
 *
 * Level order (the reference's order, workingPad/main.cu:963-1008): the queue
 * is a pair of HBM buffers in structure-of-arrays form (known planes, up
 * planes, FP64 mass).  One launch of k_expand_level consumes one tree level:
 * each lane takes one item (coalesced SoA loads), runs the BFS/path search
 * of nr_device.h, and the wave reserves room for all of its children with a
 * single atomic (prefix scan over the lanes' child counts), then every lane
 * writes its children at its offset.  The last workgroup to finish clamps
 * the next level to the capacity, counts the level and resets the counters,
 * so consecutive levels are launched back to back with no host round trip;
 * the host reads the control block once per chunk of levels.
 *
 * Masses are FP64 products taken left to right along the path; R, D and ESP
 * are per-lane Neumaier sums, merged per wave into double-double slots that
 * the host reduces at the end (replaces the single-address FP64 CAS of
 * main.cu:68-79).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "nr_device.h"
#include "nr_internal.h"

namespace nr {

/* Control block of a level-order solve (device memory). */
struct LevelCtl {
  u64 count[2];   /* items in buffer 0/1; the write side doubles as the
                     append counter and may exceed the capacity */
  u64 iterations; /* non-empty levels processed */
  u64 max_width;  /* widest level */
  u64 nodes, paths, dead, truncated, overflowed, folded, errors;
  u32 blocks_done;
  u32 pad;
};

struct QBuf {
  u64* known; /* [w][cap] */
  u64* up;    /* [w][cap] */
  double* mass;
};

struct LevelArgs {
  DevGraph g;
  Layout lay;
  QBuf buf[2];
  LevelCtl* ctl;
  DD* partR;
  DD* partE;
  DD* partD;
  u64 cap;
  double eps;
};

template <int W>
__device__ inline void load_item(const QBuf& b, u64 cap, u64 i, u64 (&known)[W],
                                 u64 (&up)[W], double& mass) {
#pragma unroll
  for (int w = 0; w < W; ++w) {
    known[w] = b.known[w * cap + i];
    up[w] = b.up[w * cap + i];
  }
  mass = b.mass[i];
}

template <int W>
__device__ inline void store_item(const QBuf& b, u64 cap, u64 i, const u64 (&known)[W],
                                  const u64 (&up)[W], double mass) {
#pragma unroll
  for (int w = 0; w < W; ++w) {
    b.known[w * cap + i] = known[w];
    b.up[w * cap + i] = up[w];
  }
  b.mass[i] = mass;
}

/* One tree level: consume buf[parity], append children to buf[parity^1]. */
template <int W, typename NS>
__global__ void __launch_bounds__(256) k_expand_level(LevelArgs a, int parity) {
  extern __shared__ __align__(16) unsigned char smem[];
  stage_tables(a.g, a.lay, smem);
  __syncthreads();

  Expander<W, NS> ex;
  ex.bind(smem, a.lay, threadIdx.x, a.g.src, a.g.dst);

  const QBuf in = a.buf[parity];
  const QBuf out = a.buf[parity ^ 1];
  const u64 n_items = __atomic_load_n(&a.ctl->count[parity], __ATOMIC_RELAXED);
  const int lane = threadIdx.x & 63;
  const u64 wave_id = (u64)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const u64 n_waves = (u64)(gridDim.x * blockDim.x) >> 6;

  Acc2 accR, accE, accD;
  accR.init();
  accE.init();
  accD.init();
  u32 c_nodes = 0, c_paths = 0, c_dead = 0, c_trunc = 0, c_over = 0, c_err = 0;

  for (u64 base = wave_id * 64; base < n_items; base += n_waves * 64) {
    const u64 i = base + lane;
    u64 known[W], up[W], Dn[W];
    double mass = 0.0;
    int k = 0;
    bool live = false;
    if (i < n_items) {
      load_item<W>(in, a.cap, i, known, up, mass);
      ++c_nodes;
      if (mass < a.eps) { /* main.cu:537-543 */
        accE.add(mass);
        ++c_trunc;
      } else {
#pragma unroll
        for (int w = 0; w < W; ++w) Dn[w] = known[w] & ~up[w];
        k = ex.find(known, Dn);
        if (k >= 0) {
          live = true;
        } else {
          if (k == -2) ++c_err;
          accD.add(mass); /* main.cu:579-582 */
          ++c_dead;
          k = 0;
        }
      }
    }
    const int incl = wave_incl_scan(k);
    const int total = __shfl(incl, 63, 64);
    u64 wbase = 0;
    if (total) {
      unsigned lo = 0, hi = 0;
      if (lane == 63) {
        u64 b = atomicAdd((unsigned long long*)&a.ctl->count[parity ^ 1], (u64)total);
        lo = (unsigned)b;
        hi = (unsigned)(b >> 32);
      }
      lo = __shfl(lo, 63, 64);
      hi = __shfl(hi, 63, 64);
      wbase = ((u64)hi << 32) | lo;
    }
    if (live) {
      /* main.cu:214-266: child j has the earlier UNKNOWN hops WORKING and
       * hop j FAILED; masses multiply left to right along the path */
      u64 slot = wbase + (u64)(incl - k);
      double run = mass;
      u64 pre[W];
#pragma unroll
      for (int w = 0; w < W; ++w) pre[w] = 0;
      for (int j = k - 1; j >= 0; --j, ++slot) {
        const int e = ex.uh[j * ex.stride];
        const int wi = e >> 6;
        const u64 bit = 1ull << (e & 63);
        const double cm = run * ex.q[e];
        if (slot < a.cap) {
          u64 ck[W], cu[W];
#pragma unroll
          for (int w = 0; w < W; ++w) {
            ck[w] = known[w] | pre[w] | (w == wi ? bit : 0ull);
            cu[w] = up[w] | pre[w];
          }
          store_item<W>(out, a.cap, slot, ck, cu, cm);
        } else { /* queue full: main.cu:245-252 */
          accE.add(cm);
          ++c_paths;
          ++c_over;
        }
        run = run * ex.p[e];
#pragma unroll
        for (int w = 0; w < W; ++w)
          if (w == wi) pre[w] |= bit;
      }
      accR.add(run); /* success term, main.cu:254-262 */
      ++c_paths;
    }
  }

  /* per-wave flush */
  wave_sum_dd(accR.s, accR.c);
  wave_sum_dd(accE.s, accE.c);
  wave_sum_dd(accD.s, accD.c);
  u64 t_nodes = wave_sum_u64(c_nodes), t_paths = wave_sum_u64(c_paths),
      t_dead = wave_sum_u64(c_dead), t_trunc = wave_sum_u64(c_trunc),
      t_over = wave_sum_u64(c_over), t_err = wave_sum_u64(c_err);
  if (lane == 0) {
    dd_merge_slot(&a.partR[wave_id], accR.s, accR.c);
    dd_merge_slot(&a.partE[wave_id], accE.s, accE.c);
    dd_merge_slot(&a.partD[wave_id], accD.s, accD.c);
    if (t_nodes) atomicAdd((unsigned long long*)&a.ctl->nodes, t_nodes);
    if (t_paths) atomicAdd((unsigned long long*)&a.ctl->paths, t_paths);
    if (t_dead) atomicAdd((unsigned long long*)&a.ctl->dead, t_dead);
    if (t_trunc) atomicAdd((unsigned long long*)&a.ctl->truncated, t_trunc);
    if (t_over) atomicAdd((unsigned long long*)&a.ctl->overflowed, t_over);
    if (t_err) atomicAdd((unsigned long long*)&a.ctl->errors, t_err);
  }

  /* last workgroup: close the level (main.cu:673-691 done on the device) */
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 prev = atomicAdd(&a.ctl->blocks_done, 1u);
    if (prev == gridDim.x - 1) {
      __threadfence();
      u64 produced = atomicAdd((unsigned long long*)&a.ctl->count[parity ^ 1], 0ull);
      a.ctl->count[parity ^ 1] = produced < a.cap ? produced : a.cap;
      a.ctl->count[parity] = 0;
      if (n_items) {
        a.ctl->iterations += 1;
        if (n_items > a.ctl->max_width) a.ctl->max_width = n_items;
      }
      a.ctl->blocks_done = 0;
      __threadfence();
    }
  }
}

/* Fold every queued item of buf[parity] into ESP (time budget expiry;
 * original/Seq_16.cpp:44,227 PURGEQ restated on the device). */
__global__ void __launch_bounds__(256) k_fold(LevelArgs a, int parity) {
  const u64 n_items = __atomic_load_n(&a.ctl->count[parity], __ATOMIC_RELAXED);
  const int lane = threadIdx.x & 63;
  const u64 wave_id = (u64)(blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const u64 n_waves = (u64)(gridDim.x * blockDim.x) >> 6;
  Acc2 e;
  e.init();
  u32 c = 0;
  for (u64 i = wave_id * 64 + lane; i < n_items; i += n_waves * 64) {
    e.add(a.buf[parity].mass[i]);
    ++c;
  }
  wave_sum_dd(e.s, e.c);
  u64 t = wave_sum_u64(c);
  if (lane == 0) {
    dd_merge_slot(&a.partE[wave_id], e.s, e.c);
    if (t) atomicAdd((unsigned long long*)&a.ctl->folded, t);
  }
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

template <int W, typename NS>
__global__ void __launch_bounds__(256) k_expand_items(ItemsArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  stage_tables(a.g, a.lay, smem);
  __syncthreads();
  Expander<W, NS> ex;
  ex.bind(smem, a.lay, threadIdx.x, a.g.src, a.g.dst);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.count) return;
  u64 known[W], up[W], Dn[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    known[w] = a.known[(u64)i * W + w];
    up[w] = a.up[(u64)i * W + w];
    Dn[w] = known[w] & ~up[w];
  }
  const double mass = a.mass[i];
  int k = ex.find(known, Dn);
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

/* ------------------------------------------------------------------ */
/* host runtime                                                        */
/* ------------------------------------------------------------------ */

#define NR_HIP(call)                                   \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) {                            \
      last_hip_error() = _e;                           \
      return NR_ERR_DEVICE;                            \
    }                                                  \
  } while (0)

hipError_t& last_hip_error() {
  static thread_local hipError_t e = hipSuccess;
  return e;
}

/* RAII bag of device allocations. */
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

/* Host-built graph tables. */
struct HostTables {
  std::vector<u64> sadj, inc;
  std::vector<u32> ep, deg;
  std::vector<u16> eidm;
};

int build_tables(const nr_graph* g, int W, HostTables& t) {
  const int n = g->n;
  t.sadj.assign(n, 0);
  t.inc.assign((size_t)n * W, 0);
  t.ep.assign(std::max(g->e, 1), 0);
  t.deg.assign(n, 0);
  t.eidm.assign((size_t)n * n, 0xffff);
  for (int e = 0; e < g->e; ++e) {
    int a = g->ea[e], b = g->eb[e];
    t.ep[e] = (u32)a | ((u32)b << 16);
    t.inc[(size_t)a * W + (e >> 6)] |= 1ull << (e & 63);
    t.inc[(size_t)b * W + (e >> 6)] |= 1ull << (e & 63);
    if (a != b) {
      t.sadj[a] |= 1ull << b;
      t.sadj[b] |= 1ull << a;
    }
    t.eidm[(size_t)a * n + b] = (u16)e;
    t.eidm[(size_t)b * n + a] = (u16)e;
  }
  for (int u = 0; u < n; ++u) {
    int c = 0;
    for (int i = 0; i < W; ++i) c += __builtin_popcountll(t.inc[(size_t)u * W + i]);
    t.deg[u] = (u32)c;
  }
  return NR_OK;
}

struct DeviceGraph {
  DevMem mem;
  DevGraph g;
};

int upload_graph(const nr_graph* hg, int W, DeviceGraph& d) {
  HostTables t;
  build_tables(hg, W, t);
  const int n = hg->n, e = std::max(hg->e, 1);
  double *p, *q;
  u64 *sadj, *inc;
  u32 *ep, *deg;
  u16* eidm;
  int rc;
  if ((rc = d.mem.alloc(&p, e)) || (rc = d.mem.alloc(&q, e)) ||
      (rc = d.mem.alloc(&sadj, n)) || (rc = d.mem.alloc(&inc, (size_t)n * W)) ||
      (rc = d.mem.alloc(&ep, e)) || (rc = d.mem.alloc(&deg, n)) ||
      (rc = d.mem.alloc(&eidm, (size_t)n * n + 2)))
    return rc;
  std::vector<double> hp(e, 0.0), hq(e, 0.0);
  for (int i = 0; i < hg->e; ++i) {
    hp[i] = hg->p[i];
    hq[i] = hg->q[i];
  }
  NR_HIP(hipMemcpy(p, hp.data(), sizeof(double) * e, hipMemcpyHostToDevice));
  NR_HIP(hipMemcpy(q, hq.data(), sizeof(double) * e, hipMemcpyHostToDevice));
  NR_HIP(hipMemcpy(sadj, t.sadj.data(), sizeof(u64) * n, hipMemcpyHostToDevice));
  NR_HIP(hipMemcpy(inc, t.inc.data(), sizeof(u64) * n * W, hipMemcpyHostToDevice));
  NR_HIP(hipMemcpy(ep, t.ep.data(), sizeof(u32) * e, hipMemcpyHostToDevice));
  NR_HIP(hipMemcpy(deg, t.deg.data(), sizeof(u32) * n, hipMemcpyHostToDevice));
  NR_HIP(hipMemcpy(eidm, t.eidm.data(), sizeof(u16) * n * n, hipMemcpyHostToDevice));
  d.g = DevGraph{p, q, sadj, inc, ep, deg, eidm, hg->n, hg->e, W, hg->src, hg->dst};
  return NR_OK;
}

/* Width classes: W 64-bit words per edge-state plane, node sets of 32 or 64
 * bits.  Each class is a separate instantiation of the kernels. */

template <int W, typename NS>
struct Kern {
  static void* level() { return (void*)&k_expand_level<W, NS>; }
  static void* items() { return (void*)&k_expand_items<W, NS>; }
};

typedef void* (*KernelGetter)();

struct Variant {
  KernelGetter level, items;
  int ns_bytes;
};

template <typename NS>
Variant variant_for_w(int W) {
  switch (W) {
#define NR_W(X) \
  case X: return Variant{&Kern<X, NS>::level, &Kern<X, NS>::items, (int)sizeof(NS)};
    NR_W(1) NR_W(2) NR_W(3) NR_W(4) NR_W(5) NR_W(6) NR_W(7) NR_W(8)
#undef NR_W
    default: return Variant{nullptr, nullptr, 0};
  }
}

/* Largest block (<= 256, multiple of 64) whose LDS footprint fits in 64 KiB;
 * a requested size is honoured when it fits. */
int choose_block(const nr_graph* g, int W, int ns_bytes, int requested) {
  const int cands[4] = {requested, 256, 128, 64};
  for (int b : cands) {
    if (b <= 0 || b > 256 || b % 64) continue;
    Layout L = make_layout(g->n, std::max(g->e, 1), W, b, ns_bytes);
    if (L.total_bytes <= 64 * 1024) return b;
  }
  return 0;
}

int pick_variant(const nr_graph* g, Variant& v) {
  const int W = mask_words(g);
  if (g->n > 64 || W > 8) return NR_ERR_UNSUPPORTED;
  if (!g->symmetric) return NR_ERR_UNSUPPORTED;
  v = g->n <= 32 ? variant_for_w<u32>(W) : variant_for_w<u64>(W);
  return v.level ? NR_OK : NR_ERR_UNSUPPORTED;
}

void initial_planes(const nr_graph* g, int W, std::vector<u64>& known,
                    std::vector<u64>& up) {
  /* main.cu:717-748: p <= 0 FAILED, p >= 1 WORKING, otherwise UNKNOWN */
  known.assign(W, 0);
  up.assign(W, 0);
  for (int e = 0; e < g->e; ++e) {
    u64 bit = 1ull << (e & 63);
    if (g->p[e] <= 0.0) {
      known[e >> 6] |= bit;
    } else if (g->p[e] >= 1.0) {
      known[e >> 6] |= bit;
      up[e >> 6] |= bit;
    }
  }
}

double dd_reduce(const std::vector<DD>& v) {
  double s = 0.0, c = 0.0;
  auto add = [&](double x) {
    double t = s + x;
    if (std::fabs(s) >= std::fabs(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
  };
  for (const DD& d : v) {
    add(d.hi);
    add(d.lo);
  }
  return s + c;
}

int solve_level(const nr_graph* g, const nr_config* cfg, nr_result* out) {
  Variant var;
  int rc = pick_variant(g, var);
  if (rc) return rc;
  const int W = mask_words(g);
  NR_HIP(hipSetDevice(cfg->device));
  hipDeviceProp_t prop;
  NR_HIP(hipGetDeviceProperties(&prop, cfg->device));

  DeviceGraph dg;
  if ((rc = upload_graph(g, W, dg))) return rc;
  const int block = choose_block(g, W, var.ns_bytes, cfg->block_size);
  if (block <= 0) return NR_ERR_UNSUPPORTED;
  Layout lay = make_layout(g->n, std::max(g->e, 1), W, block, var.ns_bytes);
  const void* kfun = (const void*)var.level();
  if (lay.total_bytes > 48 * 1024)
    NR_HIP(hipFuncSetAttribute(kfun, hipFuncAttributeMaxDynamicSharedMemorySize,
                               lay.total_bytes));
  int per_cu = 0;
  NR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfun, block,
                                                      lay.total_bytes));
  if (per_cu < 1) return NR_ERR_UNSUPPORTED;
  const int grid = cfg->grid_blocks > 0 ? cfg->grid_blocks : per_cu * prop.multiProcessorCount;
  const u64 n_waves = (u64)grid * block / 64;

  const u64 cap = cfg->queue_capacity > 0 ? (u64)cfg->queue_capacity : (1ull << 20);
  DevMem mem;
  LevelArgs a;
  std::memset(&a, 0, sizeof a);
  a.g = dg.g;
  a.lay = lay;
  a.cap = cap;
  a.eps = cfg->truncation_eps > 0 ? cfg->truncation_eps : 0.0;
  for (int b = 0; b < 2; ++b) {
    if ((rc = mem.alloc(&a.buf[b].known, cap * W)) || (rc = mem.alloc(&a.buf[b].up, cap * W)) ||
        (rc = mem.alloc(&a.buf[b].mass, cap)))
      return rc;
  }
  if ((rc = mem.alloc(&a.ctl, 1)) || (rc = mem.alloc(&a.partR, n_waves)) ||
      (rc = mem.alloc(&a.partE, n_waves)) || (rc = mem.alloc(&a.partD, n_waves)))
    return rc;
  NR_HIP(hipMemset(a.ctl, 0, sizeof(LevelCtl)));
  NR_HIP(hipMemset(a.partR, 0, sizeof(DD) * n_waves));
  NR_HIP(hipMemset(a.partE, 0, sizeof(DD) * n_waves));
  NR_HIP(hipMemset(a.partD, 0, sizeof(DD) * n_waves));
  /* seed the root item (main.cu:853-855) */
  {
    std::vector<u64> k0, u0;
    initial_planes(g, W, k0, u0);
    for (int w = 0; w < W; ++w) {
      NR_HIP(hipMemcpy(a.buf[0].known + (size_t)w * cap, &k0[w], 8, hipMemcpyHostToDevice));
      NR_HIP(hipMemcpy(a.buf[0].up + (size_t)w * cap, &u0[w], 8, hipMemcpyHostToDevice));
    }
    double one = 1.0;
    NR_HIP(hipMemcpy(a.buf[0].mass, &one, 8, hipMemcpyHostToDevice));
    u64 c1 = 1;
    NR_HIP(hipMemcpy(&a.ctl->count[0], &c1, 8, hipMemcpyHostToDevice));
  }
  NR_HIP(hipDeviceSynchronize());

  hipStream_t stream;
  NR_HIP(hipStreamCreate(&stream));
  struct StreamGuard {
    hipStream_t s;
    std::vector<hipEvent_t> ev;
    ~StreamGuard() {
      for (auto e : ev) (void)hipEventDestroy(e);
      (void)hipStreamDestroy(s);
    }
  } sg{stream, {}};
  auto new_event = [&](hipEvent_t* e) -> int {
    NR_HIP(hipEventCreate(e));
    sg.ev.push_back(*e);
    return NR_OK;
  };
  hipEvent_t ev_start, ev_stop;
  if ((rc = new_event(&ev_start)) || (rc = new_event(&ev_stop))) return rc;
  const size_t max_timed = 4096;
  std::vector<hipEvent_t> k_beg, k_end;

  auto t_begin = std::chrono::steady_clock::now();
  const bool has_budget = cfg->time_budget_s > 0;
  NR_HIP(hipEventRecord(ev_start, stream));
  int level = 0;
  int chunk = 4;
  bool finished = false, deadline = false;
  LevelCtl hctl;
  long long launches = 0;
  while (!finished) {
    for (int c = 0; c < chunk; ++c, ++level) {
      const int parity = level & 1;
      if (k_beg.size() < max_timed) {
        hipEvent_t b, e;
        if ((rc = new_event(&b)) || (rc = new_event(&e))) return rc;
        k_beg.push_back(b);
        k_end.push_back(e);
        NR_HIP(hipEventRecord(b, stream));
      }
      void* args[] = {&a, (void*)&parity};
      NR_HIP(hipLaunchKernel(kfun, dim3(grid), dim3(block), args, lay.total_bytes, stream));
      ++launches;
      if (k_end.size() == (size_t)launches) NR_HIP(hipEventRecord(k_end.back(), stream));
    }
    NR_HIP(hipMemcpyAsync(&hctl, a.ctl, sizeof hctl, hipMemcpyDeviceToHost, stream));
    NR_HIP(hipStreamSynchronize(stream));
    if (hctl.count[0] == 0 && hctl.count[1] == 0) finished = true;
    if (!finished && has_budget) {
      double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count();
      if (el >= cfg->time_budget_s) {
        const int parity = level & 1;
        void* args[] = {&a, (void*)&parity};
        NR_HIP(hipLaunchKernel((const void*)&k_fold, dim3(grid), dim3(block), args, 0, stream));
        deadline = true;
        finished = true;
      }
    }
    chunk = std::min(chunk * 2, 16);
  }
  NR_HIP(hipEventRecord(ev_stop, stream));
  NR_HIP(hipEventSynchronize(ev_stop));
  float ms = 0.f;
  NR_HIP(hipEventElapsedTime(&ms, ev_start, ev_stop));
  double kms = 0.0;
  const long long useful = std::min<long long>((long long)k_end.size(), launches);
  for (long long i = 0; i < useful; ++i) {
    float x = 0.f;
    NR_HIP(hipEventElapsedTime(&x, k_beg[i], k_end[i]));
    kms += x;
  }
  NR_HIP(hipMemcpy(&hctl, a.ctl, sizeof hctl, hipMemcpyDeviceToHost));
  std::vector<DD> pr(n_waves), pe(n_waves), pd(n_waves);
  NR_HIP(hipMemcpy(pr.data(), a.partR, sizeof(DD) * n_waves, hipMemcpyDeviceToHost));
  NR_HIP(hipMemcpy(pe.data(), a.partE, sizeof(DD) * n_waves, hipMemcpyDeviceToHost));
  NR_HIP(hipMemcpy(pd.data(), a.partD, sizeof(DD) * n_waves, hipMemcpyDeviceToHost));
  if (hctl.errors) return NR_ERR_DEVICE;

  std::memset(out, 0, sizeof(*out));
  out->R = dd_reduce(pr);
  out->ESP = dd_reduce(pe);
  out->D = dd_reduce(pd);
  out->paths = hctl.paths;
  out->nodes = hctl.nodes;
  out->dead = hctl.dead;
  out->truncated = hctl.truncated;
  out->overflowed = hctl.overflowed;
  out->folded = hctl.folded;
  out->iterations = (long long)hctl.iterations;
  out->max_queue_items = (long long)hctl.max_width;
  out->item_bytes = 16LL * W + 8;
  out->elapsed_ms = ms;
  out->kernel_ms = kms;
  out->kernel_launches = launches;
  out->deadline_hit = deadline ? 1 : 0;
  out->overflowed_flag = hctl.overflowed ? 1 : 0;
  out->grid_blocks = grid;
  out->block_size = block;
  out->lds_bytes = lay.total_bytes;
  out->blocks_per_cu = per_cu;
  out->compute_units = prop.multiProcessorCount;
  return NR_OK;
}

int expand_items(const nr_graph* g, int device, int count, const u64* known,
                 const u64* up, const double* mass, int max_children, int* nchild,
                 double* succ, u64* ck, u64* cu, double* cm) {
  Variant var;
  int rc = pick_variant(g, var);
  if (rc) return rc;
  if (count <= 0) return NR_OK;
  const int W = mask_words(g);
  NR_HIP(hipSetDevice(device));
  DeviceGraph dg;
  if ((rc = upload_graph(g, W, dg))) return rc;
  const int block = choose_block(g, W, var.ns_bytes, 0);
  if (block <= 0) return NR_ERR_UNSUPPORTED;
  Layout lay = make_layout(g->n, std::max(g->e, 1), W, block, var.ns_bytes);
  const void* kfun = (const void*)var.items();
  if (lay.total_bytes > 48 * 1024)
    NR_HIP(hipFuncSetAttribute(kfun, hipFuncAttributeMaxDynamicSharedMemorySize,
                               lay.total_bytes));
  DevMem mem;
  ItemsArgs a;
  std::memset(&a, 0, sizeof a);
  a.g = dg.g;
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
  NR_HIP(hipMemcpy(dk, known, sizeof(u64) * count * W, hipMemcpyHostToDevice));
  NR_HIP(hipMemcpy(du, up, sizeof(u64) * count * W, hipMemcpyHostToDevice));
  NR_HIP(hipMemcpy(dm, mass, sizeof(double) * count, hipMemcpyHostToDevice));
  NR_HIP(hipMemset(a.ck, 0, sizeof(u64) * nc * W));
  NR_HIP(hipMemset(a.cu, 0, sizeof(u64) * nc * W));
  NR_HIP(hipMemset(a.cm, 0, sizeof(double) * nc));
  a.known = dk;
  a.up = du;
  a.mass = dm;
  const int grid = (count + block - 1) / block;
  void* args[] = {&a};
  NR_HIP(hipLaunchKernel(kfun, dim3(grid), dim3(block), args, lay.total_bytes, 0));
  NR_HIP(hipDeviceSynchronize());
  NR_HIP(hipMemcpy(nchild, a.nchild, sizeof(int) * count, hipMemcpyDeviceToHost));
  NR_HIP(hipMemcpy(succ, a.succ, sizeof(double) * count, hipMemcpyDeviceToHost));
  if (ck) NR_HIP(hipMemcpy(ck, a.ck, sizeof(u64) * nc * W, hipMemcpyDeviceToHost));
  if (cu) NR_HIP(hipMemcpy(cu, a.cu, sizeof(u64) * nc * W, hipMemcpyDeviceToHost));
  if (cm) NR_HIP(hipMemcpy(cm, a.cm, sizeof(double) * nc, hipMemcpyDeviceToHost));
  return NR_OK;
}

}  // namespace nr

extern "C" {

int nr_solve(const nr_graph* g, const nr_config* cfg, nr_result* out) {
  if (!g || !out) return NR_ERR_ARG;
  nr_config def;
  if (!cfg) {
    nr_config_init(&def);
    cfg = &def;
  }
  if (cfg->queue_capacity < 0 || cfg->truncation_eps < 0 || cfg->time_budget_s < 0)
    return NR_ERR_ARG;
  switch (cfg->order) {
    case NR_ORDER_LEVEL: return nr::solve_level(g, cfg, out);
    default: return NR_ERR_UNSUPPORTED;
  }
}

int nr_solve_batch(const nr_graph* const* graphs, int count, const nr_config* cfg,
                   nr_result* results) {
  if (!graphs || count < 0 || !results) return NR_ERR_ARG;
  for (int i = 0; i < count; ++i) {
    int rc = nr_solve(graphs[i], cfg, &results[i]);
    if (rc) return rc;
  }
  return NR_OK;
}

int nr_expand_items(const nr_graph* g, int device, int count,
                    const unsigned long long* known, const unsigned long long* up,
                    const double* mass, int max_children, int* nchild, double* succ,
                    unsigned long long* child_known, unsigned long long* child_up,
                    double* child_mass) {
  if (!g || count < 0 || (count > 0 && (!known || !up || !mass || !nchild || !succ)))
    return NR_ERR_ARG;
  return nr::expand_items(g, device, count, known, up, mass, max_children, nchild, succ,
                          child_known, child_up, child_mass);
}

int nr_device_query(int device, char* name, int name_len, int* compute_units,
                    int* wave_size, long long* total_mem) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return NR_ERR_DEVICE;
  if (name && name_len > 0) {
    std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (compute_units) *compute_units = prop.multiProcessorCount;
  if (wave_size) *wave_size = prop.warpSize;
  if (total_mem) *total_mem = (long long)prop.totalGlobalMem;
  return NR_OK;
}

}  // extern "C"
