/*
 * nr_learn.h - the learned subproblem scorer on the device: child features
 * and the two-layer MLP (16 features -> 64 hidden, ReLU -> 1) evaluated for
 * 64 children at a time with gfx950 BF16 matrix-core instructions
 * (v_mfma_f32_32x32x16_bf16), FP32 accumulation.
 *
 * The score predicts log2 of the size of a child's subtree under the plain
 * method (its nodes, dead and truncated included).  It is a work-order hint
 * only: every order expands the same exact decomposition.
 *
 * Inference form of a trained model (fp32 W1[16][64], the hidden bias being
 * row 15 against the constant feature 15, w2[64], b2):
 *     W1'[k][h] = bf16(W1[k][h] * |w2[h]|),  s_h = sign(w2[h]),
 *     y = b2 + sum_h s_h * relu(sum_k bf16(x_k) * W1'[k][h])      (FP32 sums)
 * which equals b2 + sum_h w2[h] relu(x.W1[:,h]) up to the BF16 rounding of
 * the weights and features, since relu(z) |w| = relu(z |w|).
 *
 * Matrix layout: H^T (64 hidden x 64 children) = W1'^T (64 x 16) X^T (16 x 64)
 * as 2 x 2 tiles of 32x32x16.  Operand A (W1'^T rows) is fixed per model and
 * kept in VGPRs, pre-packed per lane by the host; operand B (children's
 * features) is read from a per-wave LDS tile [64 children][16] bf16.  In the
 * accumulator of a 32x32 tile, lane l holds column l % 32 (a child) and
 * rows 8*(r/4) + 4*(l/32) + r%4 for r = 0..15 (hidden units), so each lane
 * sums its 32 hidden units of a child and one exchange with lane l ^ 32
 * completes the child's output: lane l ends with child l's score.
 */
#ifndef NR_LEARN_H
#define NR_LEARN_H

#include <hip/hip_runtime.h>

#include "nr_device.h"

namespace nr {

constexpr int kFeat = 16;
constexpr int kHidden = 64;
/* W1[16][64] (row 15 multiplies the constant feature: the hidden bias),
 * w2[64], b2 */
constexpr int kMlpParams = kFeat * kHidden + kHidden + 1;

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

/* Device form of a model: per-lane operand A fragments and sign masks. */
struct MlpDev {
  u16 afrag[2][64][8]; /* [row tile][lane][k]: W1'[8*(l/32)+k][32*mt + l%32] */
  u32 sign[64];        /* bit 16*mt + r of lane l: s_h < 0 for h = 32*mt + row(r, l) */
  float b2;
  float alpha; /* score = log2(mass) - alpha * y */
  int margin;  /* path chooser: advantage (2^-8 units) needed to leave the
                  alternating direction */
  float pad;
};

/* Round-to-nearest-even fp32 -> bf16 bits (finite inputs). */
__host__ __device__ inline u16 bf16_bits(float f) {
  u32 u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__host__ __device__ inline float bf16_value(u16 b) {
  u32 u = (u32)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

/* Accumulator row of register r in lane l of a 32x32 tile. */
__host__ __device__ inline int mfma32_row(int r, int lane) {
  return 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
}

/* log2 of a positive double without underflow (exponent + mantissa). */
__device__ inline float flog2(double m) {
  if (!(m > 0.0)) return -1100.0f;
  u64 bits;
  memcpy(&bits, &m, 8);
  int e = (int)((bits >> 52) & 0x7ff);
  if (e == 0) return -1075.0f; /* subnormal: below every threshold in use */
  const double mant = __longlong_as_double((long long)((bits & 0x000fffffffffffffull) |
                                                        0x3ff0000000000000ull));
  return (float)(e - 1023) + __log2f((float)mant);
}

/* Features of child t of an item (t counted from the src side), all from
 * the parent's expansion.  Scaled to roughly [-1, 1]; f15 is the constant
 * that carries the hidden bias. */
struct ChildCtx {
  float inv_e, inv_n, inv_k;
  int u_p, w_p, f_p; /* parent's UNKNOWN / WORKING / FAILED edge counts */
  int k, len;        /* parent's unknown hops and path length */
  int deg_src, deg_dst;
  int src, dst;
  float log2_m;      /* parent mass */
  float lp_total;    /* sum of log2 p over the parent's unknown hops */
  double eps;
};

__device__ inline void child_features(const ChildCtx& c, int t, int e_a, int e_b, double p_e,
                                      double q_e, double cm, float lp_prefix, float* f) {
  const float lcm = flog2(cm);
  const float lpe = __log2f((float)p_e);
  const bool at_src = e_a == c.src || e_b == c.src;
  const bool at_dst = e_a == c.dst || e_b == c.dst;
  f[0] = lcm * (1.0f / 64.0f);
  f[1] = (float)(c.u_p - t - 1) * c.inv_e;
  f[2] = (float)(c.w_p + t) * c.inv_e;
  f[3] = (float)(c.f_p + 1) * c.inv_e;
  f[4] = (float)(t + 1) * c.inv_k;
  f[5] = (float)c.k * c.inv_n;
  f[6] = (float)c.len * c.inv_n;
  f[7] = (float)q_e;
  f[8] = (c.lp_total - lp_prefix - lpe) * (1.0f / 16.0f);
  float head = 1.0f;
  if (c.eps > 0.0) head = fminf(fmaxf((lcm - flog2(c.eps)) * (1.0f / 64.0f), 0.0f), 1.0f);
  f[9] = head;
  f[10] = fminf((float)(c.deg_src - (at_src ? 1 : 0)), 16.0f) * (1.0f / 16.0f);
  f[11] = fminf((float)(c.deg_dst - (at_dst ? 1 : 0)), 16.0f) * (1.0f / 16.0f);
  f[12] = (at_src || at_dst) ? 1.0f : 0.0f;
  f[13] = c.log2_m * (1.0f / 64.0f);
  f[14] = (float)(c.k - t - 1) * c.inv_n;
  f[15] = 1.0f;
}

/* Scores of the 64 children in a wave's tile (rows of zeros for lanes
 * without a child are harmless).  Every lane of the wave must call it;
 * returns this lane's child's y. */
__device__ inline float mlp_score_tile(const bf16x8_t (&a)[2], u32 sign, float b2,
                                       const u16* tile) {
  const int lane = threadIdx.x & 63;
  bf16x8_t b[2];
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    const u16* src = tile + (32 * nt + (lane & 31)) * kFeat + 8 * (lane >> 5);
    b[nt] = *(const bf16x8_t*)src;
  }
  float part[2] = {0.0f, 0.0f};
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      f32x16_t acc = {};
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[mt], b[nt], acc, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const float v = fmaxf(acc[r], 0.0f);
        part[nt] += ((sign >> (16 * mt + r)) & 1u) ? -v : v;
      }
    }
  }
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) part[nt] += __shfl_xor(part[nt], 32, 64);
  return (lane < 32 ? part[0] : part[1]) + b2;
}

/* ------------------------------------------------------------------ */
/* Learned path choice: the sweep direction of the fewest-UNKNOWN path  */
/* ------------------------------------------------------------------ */
/*
 * The chooser is a second network of the same shape whose input is an
 * item's path and the two terminals' components, seen from the end a sweep
 * would start at, and whose output predicts log2 of the item's subtree size
 * when swept from that end.  Both directions are scored; the sweep leaves
 * the alternating rule's direction only when the other one's prediction is
 * smaller by more than the model's margin (a conservative step of policy
 * improvement over that rule: a large margin is the rule itself).  Either
 * direction is an exact decomposition, so exactness is untouched.
 *
 * Exact arithmetic: every feature is an integer with |x| <= 63 and every
 * weight of the inference form is a/256 with |a| <= 255 (one scale for the
 * whole layer, |w2| folded in).  Then every BF16 operand is exact, every
 * product is a multiple of 2^-8, the hidden sums stay below 16*63 and the
 * output below 64*16*63 < 2^16: all partial sums are exact in FP32 whatever
 * order the matrix cores add them in, and the choice is the same integer
 * comparison a CPU makes.
 */
constexpr int kChoiceFeat = 16;

__host__ __device__ inline int floor_log2_of(double m) {
  unsigned long long bits;
  memcpy(&bits, &m, 8);
  const int e = (int)((bits >> 52) & 0x7ff);
  return e == 0 ? -1075 : e - 1023; /* zero and subnormals: below every use */
}
__host__ __device__ inline int floor_half(int v) { return v >= 0 ? v / 2 : -((1 - v) / 2); }
__host__ __device__ inline int clamp63(int v) { return v < -63 ? -63 : (v > 63 ? 63 : v); }

/* The headroom feature measures an item's mass against the truncation eps,
 * or against 2^-64 when eps is smaller or zero, in half-binades, saturating
 * at 32 (64 binades): only mass relative to the truncation enters the
 * choice, so a solve truncated deeper than the training logs, or not at all,
 * presents the same kind of states to the chooser. */
constexpr int kChoiceDepth = 32;

__host__ __device__ inline int choice_headroom(int lm, double eps) {
  const int le = eps > 0x1p-64 ? floor_log2_of(eps) : -64;
  const int h = floor_half(lm - le);
  return h < 0 ? 0 : (h > kChoiceDepth ? kChoiceDepth : h);
}

/* The 16 features of sweeping an item's path from end A (src, or dst when
 * from_dst) towards end B: lm = floor(log2 mass), ne = edges. */
__host__ __device__ inline void choice_features(const ChoiceRaw& r, bool from_dst, int lm,
                                                int head, int ne, const unsigned char* lqq,
                                                int* x) {
  const int nA = from_dst ? r.nT : r.nS, cA = from_dst ? r.cutT : r.cutS;
  const int oA = from_dst ? r.outT : r.outS;
  const int nB = from_dst ? r.nS : r.nT, cB = from_dst ? r.cutS : r.cutT;
  const int oB = from_dst ? r.outS : r.outT;
  const int e = ne > 0 ? ne : 1;
  x[0] = clamp63(r.k);          /* UNKNOWN hops = children */
  x[1] = clamp63(nA);           /* start end's component: nodes, */
  x[2] = clamp63(cA);           /* UNKNOWN edges leaving it, */
  x[3] = clamp63(oA - cA);      /* FAILED edges leaving it */
  x[4] = clamp63(nB);           /* the same for the far end */
  x[5] = clamp63(cB);
  x[6] = clamp63(oB - cB);
  x[7] = (32 * r.f) / e;        /* FAILED and UNKNOWN fractions, 1/32 steps */
  x[8] = (32 * r.u) / e;
  x[9] = 0;                     /* (unused) */
  x[10] = head;                 /* headroom over eps */
  x[11] = -clamp63(r.lp);       /* path success probability, -4 log2 steps */
  x[12] = -(int)lqq[from_dst ? r.e_dst : r.e_src]; /* q of the first failed hop */
  x[13] = -(int)lqq[from_dst ? r.e_src : r.e_dst]; /* q of the last */
  x[14] = clamp63(cA - cB);     /* cut asymmetry */
  x[15] = 1;                    /* carries the hidden bias */
}

/* Alternating-rule expanders whose sweep direction the chooser picks (the
 * search records a ChoiceRaw and leaves the hops in the forward sweep). */
template <int W, typename NS>
struct ChooseAltExpander : AltExpander<W, NS> {
  static constexpr bool kChoose = true;
  mutable ChoiceRaw raw;
  __device__ int find_item(const u64 (&known)[W], const u64 (&up)[W]) const {
    return this->template find_alt<true>(known, up, &raw);
  }
};
template <int W, int NW>
struct CsrChooseAltExpander : CsrAltExpander<W, NW> {
  static constexpr bool kChoose = true;
  mutable ChoiceRaw raw;
  __device__ int find_item(const u64 (&known)[W], const u64 (&up)[W]) const {
    return this->template find_alt<true>(known, up, &raw);
  }
};

/* Class of a scored child: two classes per binade of mass / predicted
 * work, class 0 first. */
__device__ inline int score_class(float log2_mass, float y, float alpha, int n_classes) {
  const float s = log2_mass - alpha * y;
  int c = (int)floorf(-2.0f * s);
  return c < 0 ? 0 : (c >= n_classes ? n_classes - 1 : c);
}

}  // namespace nr

#endif
