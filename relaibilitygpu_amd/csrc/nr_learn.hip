/*
 * nr_learn.hip - the learned subproblem scorer: kernels of the level engine
 * that write the tree log (MODE 1) or order the queue by the learned score
 * (MODE 2), the training kernels of the MLP (FP32 forward / backward with
 * deterministic reductions, Adam), and the model entry points of
 * include/nr_api.h.  No counterpart in the reference: this is the learned
 * work-order part added on top of its exact decomposition.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/nr_api.h"
#include "nr_internal.h"
#include "nr_learn.h"
#include "nr_level.h"

struct nr_model {
  float params[nr::kMlpParams];
  float alpha;
  bool has_order = false;
  nr::MlpDev dev; /* packed inference form of the order scorer */
  /* learned path choice: integer weights a[k][h] (the layer-1 weights times
   * |w2|, one scale for the layer, rounded to |a| <= 255), signs of w2, and
   * the device form with weights a / 256 */
  bool has_chooser = false;
  int cw[nr::kFeat * nr::kHidden];
  int cs[nr::kHidden];
  int cmargin = 0;
  nr::MlpDev cdev;
};

namespace nr {

/* ---- MODE 1 / MODE 2 instantiations ---- */
template <int W, typename NS>
struct KernLearn {
  static void* log() { return (void*)&k_expand_level<W, Expander<W, NS>, 1>; }
  static void* learn() { return (void*)&k_expand_level<W, Expander<W, NS>, 2>; }
};

template <typename NS>
static void learn_rows(Variant& v, int W) {
  switch (W) {
#define NR_W(X)                                 \
  case X:                                       \
    v.level_log = &KernLearn<X, NS>::log;       \
    v.level_learn = &KernLearn<X, NS>::learn;   \
    break;
    NR_W(1) NR_W(2) NR_W(3) NR_W(4) NR_W(5) NR_W(6) NR_W(7) NR_W(8)
#undef NR_W
    default: break;
  }
}

void add_learn_kernels_csr(Variant& v, int W); /* nr_learn_csr.hip */

void add_learn_kernels(Variant& v, int W) {
  if (v.nw) {
    add_learn_kernels_csr(v, W);
  } else if (v.ns_bytes == 4) {
    learn_rows<u32>(v, W);
  } else {
    learn_rows<u64>(v, W);
  }
}

/* ---- training ---- */
constexpr int kTB = 64; /* samples per training workgroup (one wave) */

/* Forward + backward of the MLP for up to kTB samples per workgroup; writes
 * the workgroup's gradient sums (unscaled) and loss sum to its partial slot.
 * Loss per sample: (y_hat - y)^2 / 2. */
__global__ void __launch_bounds__(kTB) k_mlp_grad(const float* X, const float* Y,
                                                  const int* idx, long long n,
                                                  const float* P, float* partial,
                                                  double* loss_part) {
  __shared__ float sW1[kFeat * kHidden];
  __shared__ float sw2[kHidden];
  __shared__ float sX[kTB][kFeat];
  __shared__ float sDZ[kTB][kHidden + 1];
  __shared__ float sRH[kTB][kHidden + 1];
  __shared__ float sR[kTB];
  const int t = threadIdx.x;
  for (int i = t; i < kFeat * kHidden; i += kTB) sW1[i] = P[i];
  sw2[t] = P[kFeat * kHidden + t];
  const float b2 = P[kFeat * kHidden + kHidden];
  const long long s0 = (long long)blockIdx.x * kTB;
  const long long sidx = s0 + t;
  const bool valid = sidx < n;
  const long long row = valid ? (idx ? (long long)idx[sidx] : sidx) : 0;
  for (int k = 0; k < kFeat; ++k) sX[t][k] = valid ? X[row * kFeat + k] : 0.0f;
  __syncthreads();
  float z[kHidden];
  float yh = b2;
#pragma unroll
  for (int h = 0; h < kHidden; ++h) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < kFeat; ++k) acc += sX[t][k] * sW1[k * kHidden + h];
    z[h] = acc;
    yh += fmaxf(acc, 0.0f) * sw2[h];
  }
  const float r = valid ? yh - Y[row] : 0.0f;
  sR[t] = r;
#pragma unroll
  for (int h = 0; h < kHidden; ++h) {
    sDZ[t][h] = z[h] > 0.0f ? r * sw2[h] : 0.0f;
    sRH[t][h] = r * fmaxf(z[h], 0.0f);
  }
  __syncthreads();
  float* out = partial + (size_t)blockIdx.x * kMlpParams;
  for (int pi = t; pi < kMlpParams; pi += kTB) {
    float acc = 0.0f;
    if (pi < kFeat * kHidden) {
      const int k = pi / kHidden, h = pi % kHidden;
      for (int s = 0; s < kTB; ++s) acc += sX[s][k] * sDZ[s][h];
    } else if (pi < kFeat * kHidden + kHidden) {
      const int h = pi - kFeat * kHidden;
      for (int s = 0; s < kTB; ++s) acc += sRH[s][h];
    } else {
      for (int s = 0; s < kTB; ++s) acc += sR[s];
    }
    out[pi] = acc;
  }
  if (t == 0) {
    double l = 0.0;
    for (int s = 0; s < kTB; ++s) l += 0.5 * (double)sR[s] * (double)sR[s];
    loss_part[blockIdx.x] = l;
  }
}

/* Sum of the partial slots in slot order (deterministic), scaled. */
__global__ void __launch_bounds__(256) k_mlp_reduce(const float* partial, int n_part,
                                                    float scale, float* G,
                                                    const double* loss_part, double* loss) {
  const int pi = blockIdx.x * blockDim.x + threadIdx.x;
  if (pi < kMlpParams) {
    float acc = 0.0f;
    for (int b = 0; b < n_part; ++b) acc += partial[(size_t)b * kMlpParams + pi];
    G[pi] = acc * scale;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && loss) {
    double l = 0.0;
    for (int b = 0; b < n_part; ++b) l += loss_part[b];
    *loss = l;
  }
}

/* Adam (Kingma & Ba 2015) with decoupled weight decay on W1 and w2. */
__global__ void __launch_bounds__(256) k_adam(float* P, const float* G, float* M, float* V,
                                              float lr, float b1, float b2, float eps,
                                              float wd, int step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kMlpParams) return;
  const float g = G[i];
  const float m = b1 * M[i] + (1.0f - b1) * g;
  const float v = b2 * V[i] + (1.0f - b2) * g * g;
  M[i] = m;
  V[i] = v;
  const float mh = m / (1.0f - powf(b1, (float)step));
  const float vh = v / (1.0f - powf(b2, (float)step));
  float p = P[i];
  if (i < kMlpParams - 1) p -= lr * wd * p;
  P[i] = p - lr * mh / (sqrtf(vh) + eps);
}

/* Scores of n feature rows with the inference form (the matrix-core tile
 * function of the expand kernel), one wave per 64 rows. */
__global__ void __launch_bounds__(64) k_mlp_score(const float* X, long long n,
                                                  const MlpDev* m, float* Y) {
  __shared__ __align__(16) u16 tile[64 * kFeat];
  const int lane = threadIdx.x;
  const long long row = (long long)blockIdx.x * 64 + lane;
  u32* trow = (u32*)(tile + lane * kFeat);
#pragma unroll
  for (int x = 0; x < kFeat; x += 2) {
    const float f0 = row < n ? X[row * kFeat + x] : 0.0f;
    const float f1 = row < n ? X[row * kFeat + x + 1] : 0.0f;
    trow[x / 2] = (u32)bf16_bits(f0) | ((u32)bf16_bits(f1) << 16);
  }
  bf16x8_t a[2];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) a[mt] = *(const bf16x8_t*)m->afrag[mt][lane];
  __syncthreads();
  const float y = mlp_score_tile(a, m->sign[lane], m->b2, tile);
  if (row < n) Y[row] = y;
}

/* Exact inference form of a chooser (see nr_learn.h): W1'[k][h] =
 * W1[k][h] |w2[h]| scaled so that the largest magnitude is 255/256 and
 * rounded to the 1/256 grid; the output bias plays no part in a choice. */
void pack_chooser(const float* P, float margin, int* a, int* sg, int* mi, MlpDev& d) {
  std::memset(&d, 0, sizeof d);
  const float* W1 = P;
  const float* w2 = P + kFeat * kHidden;
  double mx = 0.0;
  for (int k = 0; k < kFeat; ++k)
    for (int h = 0; h < kHidden; ++h)
      mx = std::max(mx, std::fabs((double)W1[k * kHidden + h] * std::fabs((double)w2[h])));
  const double c = mx > 0.0 ? 255.0 / mx : 0.0;
  for (int k = 0; k < kFeat; ++k)
    for (int h = 0; h < kHidden; ++h) {
      const double v = (double)W1[k * kHidden + h] * std::fabs((double)w2[h]) * c;
      long r = std::lrint(v);
      a[k * kHidden + h] = (int)std::max(-255L, std::min(255L, r));
    }
  for (int h = 0; h < kHidden; ++h) sg[h] = w2[h] < 0.0f ? 1 : 0;
  /* a prediction difference dy (log2 subtree nodes) is about c * dy in the
   * integer units of the scores (2^-8 of the device form's output) */
  const double m = (double)margin * c;
  *mi = m >= 2e9 ? 2000000000 : (int)std::lrint(std::max(0.0, m));
  for (int mt = 0; mt < 2; ++mt)
    for (int l = 0; l < 64; ++l)
      for (int j = 0; j < 8; ++j) {
        const int k = 8 * (l / 32) + j, h = 32 * mt + (l % 32);
        d.afrag[mt][l][j] = bf16_bits((float)a[k * kHidden + h] / 256.0f); /* exact */
      }
  for (int l = 0; l < 64; ++l) {
    u32 m = 0;
    for (int mt = 0; mt < 2; ++mt)
      for (int r = 0; r < 16; ++r)
        if (sg[32 * mt + mfma32_row(r, l)]) m |= 1u << (16 * mt + r);
    d.sign[l] = m;
  }
  d.margin = *mi;
}

void pack_model(const float* P, float alpha, MlpDev& d) {
  std::memset(&d, 0, sizeof d);
  const float* W1 = P;
  const float* w2 = P + kFeat * kHidden;
  for (int mt = 0; mt < 2; ++mt)
    for (int l = 0; l < 64; ++l)
      for (int j = 0; j < 8; ++j) {
        const int k = 8 * (l / 32) + j, h = 32 * mt + (l % 32);
        d.afrag[mt][l][j] = bf16_bits(W1[k * kHidden + h] * std::fabs(w2[h]));
      }
  for (int l = 0; l < 64; ++l) {
    u32 m = 0;
    for (int mt = 0; mt < 2; ++mt)
      for (int r = 0; r < 16; ++r) {
        const int h = 32 * mt + mfma32_row(r, l);
        if (w2[h] < 0.0f) m |= 1u << (16 * mt + r);
      }
    d.sign[l] = m;
  }
  d.b2 = P[kMlpParams - 1];
  d.alpha = alpha;
}

/* Device buffers of one training / gradient call. */
struct TrainBufs {
  std::vector<void*> ptrs;
  ~TrainBufs() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  int alloc(T** out, size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return NR_ERR_NOMEM;
    ptrs.push_back(p);
    *out = (T*)p;
    return NR_OK;
  }
};

#define NR_CHK(x)                                \
  do {                                           \
    if ((x) != hipSuccess) return NR_ERR_DEVICE; \
  } while (0)

int mlp_grad_dev(hipStream_t st, const float* dX, const float* dY, const int* didx, long long n,
                 const float* dP, float* dG, double* dloss, TrainBufs& tb, float*& part,
                 double*& lpart, long long& part_cap) {
  const long long blocks = (n + kTB - 1) / kTB;
  if (blocks > part_cap) {
    int rc;
    if ((rc = tb.alloc(&part, (size_t)blocks * kMlpParams)) || (rc = tb.alloc(&lpart, blocks)))
      return rc;
    part_cap = blocks;
  }
  hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)blocks), dim3(kTB), 0, st, dX, dY, didx, n, dP,
                     part, lpart);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, st,
                     (const float*)part, (int)blocks, 1.0f / (float)n, dG, (const double*)lpart,
                     dloss);
  NR_CHK(hipGetLastError());
  return NR_OK;
}

}  // namespace nr

using namespace nr;

extern "C" {

int nr_mlp_init(unsigned long long seed, float* params) {
  return nr::guarded([&]() -> int {
    if (!params) return NR_ERR_ARG;
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    const float s1 = std::sqrt(2.0f / kFeat), s2 = std::sqrt(1.0f / kHidden);
    for (int i = 0; i < kFeat * kHidden; ++i) params[i] = nd(rng) * s1;
    for (int h = 0; h < kHidden; ++h) params[kFeat * kHidden + h] = nd(rng) * s2;
    params[kMlpParams - 1] = 0.0f;
    return NR_OK;
  });
}

int nr_mlp_grad(const float* X, const float* Y, long long n, const float* params, float* grad,
                double* loss, int device) {
  return nr::guarded([&]() -> int {
    if (!X || !Y || n <= 0 || !params || !grad) return NR_ERR_ARG;
    NR_CHK(hipSetDevice(device));
    TrainBufs tb;
    float *dX, *dY, *dP, *dG, *part = nullptr;
    double *dl, *lpart = nullptr;
    long long cap = 0;
    int rc;
    if ((rc = tb.alloc(&dX, (size_t)n * kFeat)) || (rc = tb.alloc(&dY, n)) ||
        (rc = tb.alloc(&dP, kMlpParams)) || (rc = tb.alloc(&dG, kMlpParams)) ||
        (rc = tb.alloc(&dl, 1)))
      return rc;
    hipStream_t st;
    NR_CHK(hipStreamCreate(&st));
    struct S {
      hipStream_t s;
      ~S() { (void)hipStreamDestroy(s); }
    } guard{st};
    NR_CHK(hipMemcpyAsync(dX, X, sizeof(float) * n * kFeat, hipMemcpyHostToDevice, st));
    NR_CHK(hipMemcpyAsync(dY, Y, sizeof(float) * n, hipMemcpyHostToDevice, st));
    NR_CHK(hipMemcpyAsync(dP, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st));
    if ((rc = mlp_grad_dev(st, dX, dY, nullptr, n, dP, dG, dl, tb, part, lpart, cap))) return rc;
    double hl = 0.0;
    NR_CHK(hipMemcpyAsync(grad, dG, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, st));
    NR_CHK(hipMemcpyAsync(&hl, dl, sizeof(double), hipMemcpyDeviceToHost, st));
    NR_CHK(hipStreamSynchronize(st));
    if (loss) *loss = hl / (double)n;
    return NR_OK;
  });
}

int nr_train_config_init(nr_train_config* c) {
  return nr::guarded([&]() -> int {
    if (!c) return NR_ERR_ARG;
    std::memset(c, 0, sizeof *c);
    c->epochs = 8;
    c->batch = 4096;
    c->lr = 3e-3f;
    c->weight_decay = 1e-5f;
    c->seed = 1;
    c->device = 0;
    c->holdout = 0.1f;
    return NR_OK;
  });
}

int nr_mlp_train(const float* X, const float* Y, long long n, const nr_train_config* cfg,
                 float* params, float* train_loss, float* holdout_loss) {
  return nr::guarded([&]() -> int {
    if (!X || !Y || n < 2 || !cfg || !params || cfg->epochs < 1 || cfg->batch < 1 ||
        cfg->holdout < 0.0f || cfg->holdout >= 1.0f)
      return NR_ERR_ARG;
    NR_CHK(hipSetDevice(cfg->device));
    const long long n_hold = (long long)((double)n * cfg->holdout);
    const long long n_train = n - n_hold;
    std::mt19937_64 rng(cfg->seed);
    std::vector<int> order((size_t)n);
    for (long long i = 0; i < n; ++i) order[(size_t)i] = (int)i;
    std::shuffle(order.begin(), order.end(), rng);
    std::vector<int> hold(order.begin() + n_train, order.end());
    std::vector<int> train(order.begin(), order.begin() + n_train);
    TrainBufs tb;
    float *dX, *dY, *dP, *dG, *dM, *dV, *part = nullptr;
    double *dl, *lpart = nullptr;
    int *didx, *dhold;
    long long cap = 0;
    int rc;
    if ((rc = tb.alloc(&dX, (size_t)n * kFeat)) || (rc = tb.alloc(&dY, n)) ||
        (rc = tb.alloc(&dP, kMlpParams)) || (rc = tb.alloc(&dG, kMlpParams)) ||
        (rc = tb.alloc(&dM, kMlpParams)) || (rc = tb.alloc(&dV, kMlpParams)) ||
        (rc = tb.alloc(&dl, 1)) || (rc = tb.alloc(&didx, n_train)) ||
        (rc = tb.alloc(&dhold, std::max<long long>(n_hold, 1))))
      return rc;
    hipStream_t st;
    NR_CHK(hipStreamCreate(&st));
    struct S {
      hipStream_t s;
      ~S() { (void)hipStreamDestroy(s); }
    } guard{st};
    NR_CHK(hipMemcpyAsync(dX, X, sizeof(float) * n * kFeat, hipMemcpyHostToDevice, st));
    NR_CHK(hipMemcpyAsync(dY, Y, sizeof(float) * n, hipMemcpyHostToDevice, st));
    NR_CHK(hipMemcpyAsync(dP, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st));
    if (n_hold)
      NR_CHK(hipMemcpyAsync(dhold, hold.data(), sizeof(int) * n_hold, hipMemcpyHostToDevice, st));
    std::vector<float> zeros(kMlpParams, 0.0f);
    NR_CHK(hipMemcpyAsync(dM, zeros.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st));
    NR_CHK(hipMemcpyAsync(dV, zeros.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st));
    int step = 0;
    for (int ep = 0; ep < cfg->epochs; ++ep) {
      std::shuffle(train.begin(), train.end(), rng);
      NR_CHK(hipMemcpyAsync(didx, train.data(), sizeof(int) * n_train, hipMemcpyHostToDevice, st));
      double sum = 0.0;
      for (long long i0 = 0; i0 < n_train; i0 += cfg->batch) {
        const long long nb = std::min<long long>(cfg->batch, n_train - i0);
        if ((rc = mlp_grad_dev(st, dX, dY, didx + i0, nb, dP, dG, dl, tb, part, lpart, cap)))
          return rc;
        hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, st, dP,
                           (const float*)dG, dM, dV, cfg->lr, 0.9f, 0.999f, 1e-8f,
                           cfg->weight_decay, ++step);
        double hl = 0.0;
        NR_CHK(hipMemcpyAsync(&hl, dl, sizeof(double), hipMemcpyDeviceToHost, st));
        NR_CHK(hipStreamSynchronize(st));
        sum += hl;
      }
      if (train_loss) train_loss[ep] = (float)(sum / (double)n_train);
      if (holdout_loss) {
        double hl = 0.0;
        if (n_hold) {
          if ((rc = mlp_grad_dev(st, dX, dY, dhold, n_hold, dP, dG, dl, tb, part, lpart, cap)))
            return rc;
          NR_CHK(hipMemcpyAsync(&hl, dl, sizeof(double), hipMemcpyDeviceToHost, st));
          NR_CHK(hipStreamSynchronize(st));
          hl /= (double)n_hold;
        }
        holdout_loss[ep] = (float)hl;
      }
    }
    NR_CHK(hipMemcpyAsync(params, dP, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, st));
    NR_CHK(hipStreamSynchronize(st));
    return NR_OK;
  });
}

int nr_model_create(const float* params, float alpha, nr_model** out) {
  return nr::guarded([&]() -> int {
    if (!out) return NR_ERR_ARG;
    nr_model* m = new (std::nothrow) nr_model;
    if (!m) return NR_ERR_NOMEM;
    m->alpha = alpha;
    if (params) {
      std::memcpy(m->params, params, sizeof(float) * kMlpParams);
      pack_model(params, alpha, m->dev);
      m->has_order = true;
    }
    *out = m;
    return NR_OK;
  });
}

int nr_model_set_chooser(nr_model* m, const float* params, float margin) {
  return nr::guarded([&]() -> int {
    if (!m || !params || !(margin >= 0.0f)) return NR_ERR_ARG;
    for (int i = 0; i < kMlpParams; ++i)
      if (!std::isfinite(params[i])) return NR_ERR_ARG;
    pack_chooser(params, margin, m->cw, m->cs, &m->cmargin, m->cdev);
    m->has_chooser = true;
    return NR_OK;
  });
}

int nr_model_get_chooser(const nr_model* m, int* weights, int* signs, int* margin) {
  return nr::guarded([&]() -> int {
    if (!m || !m->has_chooser) return NR_ERR_ARG;
    if (weights) std::memcpy(weights, m->cw, sizeof m->cw);
    if (signs) std::memcpy(signs, m->cs, sizeof m->cs);
    if (margin) *margin = m->cmargin;
    return NR_OK;
  });
}

void nr_model_destroy(nr_model* m) { delete m; }

int nr_model_score(const nr_model* m, const float* X, long long n, float* Y, int device) {
  return nr::guarded([&]() -> int {
    if (!m || !m->has_order || !X || !Y || n < 0) return NR_ERR_ARG;
    if (n == 0) return NR_OK;
    NR_CHK(hipSetDevice(device));
    TrainBufs tb;
    float *dX, *dY;
    MlpDev* dm;
    int rc;
    if ((rc = tb.alloc(&dX, (size_t)n * kFeat)) || (rc = tb.alloc(&dY, n)) ||
        (rc = tb.alloc(&dm, 1)))
      return rc;
    NR_CHK(hipMemcpy(dX, X, sizeof(float) * n * kFeat, hipMemcpyHostToDevice));
    NR_CHK(hipMemcpy(dm, &m->dev, sizeof(MlpDev), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mlp_score, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, 0,
                       (const float*)dX, n, (const MlpDev*)dm, dY);
    NR_CHK(hipGetLastError());
    NR_CHK(hipMemcpy(Y, dY, sizeof(float) * n, hipMemcpyDeviceToHost));
    return NR_OK;
  });
}

}  // extern "C"

namespace nr {
const MlpDev* model_dev(const nr_model* m) { return m && m->has_order ? &m->dev : nullptr; }
const MlpDev* model_chooser(const nr_model* m) {
  return m && m->has_chooser ? &m->cdev : nullptr;
}
}  // namespace nr
