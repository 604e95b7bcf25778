// Learned lifting (an extension with no counterpart in the reference): an
// integer lifting transform whose predict step at levels >= 3 adds to the
// interpolet (hat) predict the output of a small MLP
//     8 -> 32 -> 32 -> 1, GELU, bf16 weights and activations, fp32 accumulate,
// evaluated with bf16 MFMA (v_mfma_f32_16x16x16_bf16 for the first layer,
// v_mfma_f32_16x16x32_bf16 for the other two), whose inputs are the 4
// coarse neighbours on each side of the predicted point along the lifted
// axis, divided by their max-abs.  The interpolet predict is the fallback.
//
// Why integers: a prediction that passes through bf16 roundings is a
// discontinuous function of its inputs, so a floating-point lifting would
// reconstruct coarse values that differ from the forward ones in the last bit
// and the inverse prediction could then differ by a bf16 step.  On integers
// (fixed point: q = round(f / scale)) every lifting step is exact
// (Calderbank-Daubechies-Sweldens-Yeo integer lifting): the inverse sees
// bit-identical coarse values, recomputes the bit-identical prediction (same
// kernel, same instruction sequence, deterministic MFMA) and subtracts it back.
// The only round-trip error is the quantization, |f - scale*q| <= scale/2.
//
// One axis sweep (levels l = L .. lmin+1 forward, reverse order inverse; the
// storage stride of level l is s = 2^(L-l), period n = 2^L):
//   predict (odd points x = s (mod 2s)), with coarse neighbours
//     c[-4..-1, 1..4] = g[x -/+ s, 3s, 5s, 7s] (periodic):
//       P = floor((c[-1] + c[1] + 1) / 2)                  interpolet
//         + round(sigma * mlp(c / sigma))                  levels >= 3 when on
//       (sigma = max |c|; the MLP term is clamped to +-4 sigma)
//       forward g[x] -= P,  inverse g[x] += P
//   update (even points, when enabled; the integer CDF 5/3 update):
//       forward g[e] += floor((g[e-s] + g[e+s] + 2) / 4), inverse g[e] -= ...
// Forward: predict then update per level, finest level first.  Inverse: undo
// update then predict per level, coarsest first.  Axes: forward 0..D-1,
// inverse D-1..0 (the learned steps of different axes do not commute).
#include "common.hpp"
#include "lifting.hpp"

#include <cmath>
#include <cstdint>

namespace paliwa_hip {
namespace learned {

constexpr int kThreads = 256;
constexpr int kTile = 16;    // lines per tile
constexpr int kLogTile = 4;
constexpr int kHidden = 32;
constexpr int kInputs = 8;
// network groups of 16 points interleaved per wavefront in the transform
constexpr int kGroupsInFlight = 2;
// Packed fp32 parameter layout.
constexpr int kW1 = 0;                         // [32][8]
constexpr int kB1 = kW1 + kHidden * kInputs;   // [32]
constexpr int kW2 = kB1 + kHidden;             // [32][32]
constexpr int kB2 = kW2 + kHidden * kHidden;   // [32]
constexpr int kW3 = kB2 + kHidden;             // [32]
constexpr int kB3 = kW3 + kHidden;             // [1]
constexpr int kParams = kB3 + 1;               // 1377

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// fp32 -> bf16 bits, round to nearest even (gfx950's v_cvt_pk_bf16_f32).
__device__ __forceinline__ short to_bf16(float x) {
  return __builtin_bit_cast(short, static_cast<__bf16>(x));
}
__device__ __forceinline__ float from_bf16(short b) {
  return __uint_as_float(static_cast<uint32_t>(static_cast<uint16_t>(b)) << 16);
}
__device__ __forceinline__ float round_bf16(float x) {
  return from_bf16(to_bf16(x));
}
// bf16 pairs: one v_cvt_pk_bf16_f32 (the same rounding as to_bf16)
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));

// GELU without transcendentals, evaluated two values per instruction in
// packed fp16:
//   gelu(z) = z * g(z),  g = clamp(0.5 + s Q(s^2 - 3/8), 0, 1),  s = fp16(z / 4),
// Q of degree 4 with fp16 coefficients, every step an fp16 multiply-add
// (one rounding each), z * g in fp32, then the bf16 rounding every activation
// gets.  In exact arithmetic g is a monotone minimax fit to the gate of the
// tanh form
//   gelu(z) = 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))
// within 1.07e-3 of it in z * g; with the fp16 roundings, within 2.04e-3 for
// every fp16 s (below the bf16 rounding of activations above 0.5 in
// magnitude).  g reaches 1 at z = 3.378 and stays there (0 below -3.378):
// the polynomial's leading coefficient is positive and, in fp16, s^2 - 3/8
// and Q overflow to +inf for large |s|, which the clamp maps to 1 or 0, so no
// clamp of s is needed.  Per pair of activations: a conversion to fp16, five
// packed multiply-adds for the polynomial and one (with the [0, 1] clamp) for
// g, two fp32 products (v_fma_mix_f32 reads g's fp16 halves), one bf16 pack.
// The transform's both directions and the training kernels all evaluate g
// with these fp16 operations, so the network the inverse evaluates is bit for
// bit the one the forward evaluated.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
constexpr float kGateScale = 0.25f; // s = z * kGateScale
constexpr float kGateShift = 0.375f;
constexpr float kGateQ0 = 0.8046875f, kGateQ1 = -0.9580078125f, kGateQ2 = 1.541015625f,
                kGateQ3 = -2.6484375f, kGateQ4 = 2.87109375f; // exact fp16 values

template <typename V>
__device__ __forceinline__ V splat(float x) {
  return static_cast<V>(static_cast<_Float16>(x));
}
template <>
__device__ __forceinline__ f16x2 splat<f16x2>(float x) {
  _Float16 const h = static_cast<_Float16>(x);
  return f16x2{h, h};
}
template <>
__device__ __forceinline__ f16x4 splat<f16x4>(float x) {
  _Float16 const h = static_cast<_Float16>(x);
  return f16x4{h, h, h, h};
}
// V = _Float16 (the training kernels) or f16x4 (the transform): the same
// fp16 operations element by element, so both give bit-identical gates.
template <typename V>
__device__ __forceinline__ V gate(V s) {
  V const w = __builtin_elementwise_fma(s, s, splat<V>(-kGateShift));
  V p = __builtin_elementwise_fma(splat<V>(kGateQ4), w, splat<V>(kGateQ3));
  p = __builtin_elementwise_fma(p, w, splat<V>(kGateQ2));
  p = __builtin_elementwise_fma(p, w, splat<V>(kGateQ1));
  p = __builtin_elementwise_fma(p, w, splat<V>(kGateQ0));
  V const g = __builtin_elementwise_fma(s, p, splat<V>(0.5f));
  return __builtin_elementwise_min(__builtin_elementwise_max(g, splat<V>(0.0f)),
                                   splat<V>(1.0f));
}
// gelu of one fp32 pre-activation (training kernels)
__device__ __forceinline__ float gelu(float z) {
  _Float16 const g = gate(static_cast<_Float16>(z * kGateScale));
  return z * static_cast<float>(g);
}
// d/dz of z * g(z) for the exact-arithmetic g (straight through the fp16
// roundings): g + z g'(z) where 0 < raw g < 1, g'(z) = (Q(w) + 2 s^2 Q'(w)) / 4.
__device__ __forceinline__ float gelu_grad(float z) {
  float const s = z * kGateScale;
  float const w = s * s - kGateShift;
  float p = __builtin_fmaf(kGateQ4, w, kGateQ3);
  p = __builtin_fmaf(p, w, kGateQ2);
  p = __builtin_fmaf(p, w, kGateQ1);
  p = __builtin_fmaf(p, w, kGateQ0);
  float const raw = __builtin_fmaf(s, p, 0.5f);
  float dp = __builtin_fmaf(4.0f * kGateQ4, w, 3.0f * kGateQ3);
  dp = __builtin_fmaf(dp, w, 2.0f * kGateQ2);
  dp = __builtin_fmaf(dp, w, kGateQ1);
  float const dg = kGateScale * __builtin_fmaf(2.0f * s * s, dp, p);
  float const g = __builtin_amdgcn_fmed3f(raw, 0.0f, 1.0f);
  return raw > 0.0f && raw < 1.0f ? __builtin_fmaf(z, dg, g) : g;
}

// Per-lane register fragments of the network for the MFMA evaluation.
// Transposed formulation: each layer computes H^T = W * X^T with A = weight
// rows (units) x K and B = X^T (K x 16 points), so the D layout of one layer
// (lane l: units 4*(l/16)+j of a 16-unit row block, point l%16) is the B
// layout of the next.  Layer 1 runs v_mfma_f32_16x16x16_bf16 (K = the 8
// inputs, padded with zero weights), layers 2 and 3 gfx950's
// v_mfma_f32_16x16x32_bf16, whose B fragment (lane l: 8 consecutive k of
// point l%16) is the two row blocks' D fragments side by side: lane group
// q = l/16 holds units 4q..4q+3 and 16+4q..16+4q+3, the k order kunit()
// gives the weights.  Biases are the accumulators' initial values (fp32).
// Layers 1 and 2 run on weights and biases scaled by 1/4 and the output layer
// on weights scaled by 4 (exact in bf16 and fp32), so the accumulators hold
// z/4 = s before its fp16 rounding, the products z/4 * g are the
// activations / 4, and the output is r; every rounding scales with them, so
// each value is exactly the unscaled network's.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

struct Net {
  bf16x4 w1[2]; // A fragments of W1 / 4 per row block (k >= 8: 0)
  f32x4 b1[2];  // b1 / 4 of the lane's 4 D units per row block
  bf16x8 w2[2]; // A fragments of W2 per row block, k order kunit()
  f32x4 b2[2];  // b2 / 4
  bf16x8 w3;    // 4 * output weights in every row, k order kunit()
  float b3;
};

// hidden unit of slot j (0..7) of lane group q in a 16x16x32 fragment
__device__ __forceinline__ int kunit(int q, int j) {
  return j < 4 ? 4 * q + j : 16 + 4 * q + (j - 4);
}

__device__ __forceinline__ void load_net(float const *prm, Net &net) {
  int const lane = threadIdx.x & 63;
  int const row = lane & 15;
  int const q = lane >> 4;
#pragma unroll
  for (int blk = 0; blk < 2; ++blk) {
    int const unit = 16 * blk + row;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int const k = 4 * q + j;
      net.w1[blk][j] = to_bf16(k < kInputs ? prm[kW1 + unit * kInputs + k] * kGateScale : 0.0f);
      net.b1[blk][j] = prm[kB1 + 16 * blk + 4 * q + j] * kGateScale;
      net.b2[blk][j] = prm[kB2 + 16 * blk + 4 * q + j] * kGateScale;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j)
      net.w2[blk][j] = static_cast<__bf16>(prm[kW2 + unit * kHidden + kunit(q, j)]);
  }
#pragma unroll
  for (int j = 0; j < 8; ++j)
    net.w3[j] = static_cast<__bf16>(prm[kW3 + kunit(q, j)] * 4.0f);
  net.b3 = prm[kB3];
}

// One D fragment of pre-activations s = z/4 (fp32) -> the 4 activations / 4,
// bf16, packed as one half of the next layer's B fragment.  The gate runs on
// an fp16 4-vector: two packed instructions per step, independent of each
// other, so neither waits on the other's result.
__device__ __forceinline__ u32x2_t activate(f32x4 d) {
  f16x4 const g = gate(__builtin_convertvector(d, f16x4));
  // z/4 * g in fp32, one rounding (v_fma_mix_f32 with a zero addend)
  float const h0 = __builtin_fmaf(d[0], static_cast<float>(g[0]), 0.0f);
  float const h1 = __builtin_fmaf(d[1], static_cast<float>(g[1]), 0.0f);
  float const h2 = __builtin_fmaf(d[2], static_cast<float>(g[2]), 0.0f);
  float const h3 = __builtin_fmaf(d[3], static_cast<float>(g[3]), 0.0f);
  return u32x2_t{__builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2_t{h0, h1}, bf16x2_t)),
                 __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2_t{h2, h3}, bf16x2_t))};
}

__device__ __forceinline__ bf16x8 side_by_side(u32x2_t const (&h)[2]) {
  return __builtin_bit_cast(bf16x8, u32x4{h[0][0], h[0][1], h[1][0], h[1][1]});
}

// The network on NB groups of 16 points of one wavefront, layer by layer
// across the groups (one group's MFMAs run while another's activations
// issue): lane l supplies inputs x[4*(l/16) .. +3] of point l%16 of each
// group (lane groups 2 and 3: any finite values, their weights are 0).
// out[b] = r of point l%16 of group b, in every lane.  The arithmetic
// of a group does not depend on NB.
template <int NB>
__device__ __forceinline__ void mlp16(Net const &net, bf16x4 const (&x)[NB], float (&out)[NB]) {
  u32x2_t h1[NB][2], h2[NB][2];
#pragma unroll
  for (int blk = 0; blk < 2; ++blk)
#pragma unroll
    for (int b = 0; b < NB; ++b)
      h1[b][blk] = activate(__builtin_amdgcn_mfma_f32_16x16x16bf16_1k(net.w1[blk], x[b],
                                                                      net.b1[blk], 0, 0, 0));
#pragma unroll
  for (int blk = 0; blk < 2; ++blk)
#pragma unroll
    for (int b = 0; b < NB; ++b)
      h2[b][blk] = activate(__builtin_amdgcn_mfma_f32_16x16x32_bf16(
          net.w2[blk], side_by_side(h1[b]), net.b2[blk], 0, 0, 0));
  f32x4 const b3 = {net.b3, net.b3, net.b3, net.b3};
#pragma unroll
  for (int b = 0; b < NB; ++b)
    out[b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(net.w3, side_by_side(h2[b]), b3, 0, 0,
                                                     0)[0];
}

// Network inputs of one predicted point from its 8 coarse neighbours
// c[0..7] (positions -7s, -5s, ..., +7s).
//   raw:       u_k = c_k / max|c|,                 correction = max|c| * r
//   detrended: D_k = 2 c_k - (c_3 + c_4) - (c_4 - c_3)(2k - 7), twice the
//              deviation of c_k from the interpolet line through c_3, c_4
//              (exact in integers; D_3 = D_4 = 0), u_k = D_k / max|D|,
//              correction = max|D| / 2 * r
// Raw neighbours normalized by their max-abs carry only the ~2^-9 relative
// resolution of bf16, below the curvature the interpolet misses on resolved
// data; detrending removes the part the interpolet already predicts first.
enum Features : int32_t { kFeaturesRaw = 0, kFeaturesDetrended = 1 };

// Largest correction, in units of the feature kind's correction unit.  For
// detrended inputs one unit (max|D| / 2) already exceeds the residual of any
// point between two monotone runs of neighbours (a step between c_3 and c_4
// gives |x - hat| <= max|D| / 12), and a tight bound limits the growth of
// the details, hence of the coarse values, where the network is wrong.
template <int32_t KIND>
constexpr float correction_bound() {
  return KIND == kFeaturesDetrended ? 1.0f : 4.0f;
}

// Features in fp32: a deterministic function of the coarse values (the
// forward and the inverse run these same instructions on the same inputs),
// which is all the integer lifting step needs to stay exactly invertible.
// Returns the correction unit; |u| <= 1 + 2^-22 (the approximate reciprocal
// can round max|f| * inv above 1), which the bf16 rounding of the network
// inputs takes back to 1.
template <int32_t KIND>
__device__ __forceinline__ float features(int const (&c)[8], float (&u)[8], float &norm) {
  float f[8];
  if constexpr (KIND == kFeaturesDetrended) {
    float const mid = static_cast<float>(c[3]) + static_cast<float>(c[4]);
    float const slope = static_cast<float>(c[4]) - static_cast<float>(c[3]);
#pragma unroll
    for (int k = 0; k < 8; ++k) // (2 c_k - mid) - slope (2k - 7), two roundings
      f[k] = k == 3 || k == 4
                 ? 0.0f
                 : __builtin_fmaf(-slope, static_cast<float>(2 * k - 7),
                                  __builtin_fmaf(2.0f, static_cast<float>(c[k]), -mid));
  } else {
#pragma unroll
    for (int k = 0; k < 8; ++k)
      f[k] = static_cast<float>(c[k]);
  }
  norm = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    norm = fmaxf(norm, fabsf(f[k]));
  float const inv = norm > 0.0f ? __builtin_amdgcn_rcpf(norm) : 0.0f;
#pragma unroll
  for (int k = 0; k < 8; ++k) // detrended D_3 = D_4 = 0 whatever inv is
    u[k] = KIND == kFeaturesDetrended && (k == 3 || k == 4) ? 0.0f : f[k] * inv;
  return KIND == kFeaturesDetrended ? 0.5f * norm : norm;
}

struct LearnedArgs {
  int32_t *data;
  int64_t A, B;
  int32_t m, L;
  int32_t lmin;
  int32_t inverse;
  int32_t update;
  uint32_t mlp_mask;     // bit l: MLP predict at level l
  int32_t features;      // kFeatures*
  float const *weights;  // kParams fp32
  double *stats;         // [L+1][2] sum |d| with MLP / interpolet, or null
  // pair extraction (training data), or null
  float *pairs;          // [capacity][10]: 8 inputs, target, key
  uint32_t *pair_count;
  uint32_t pair_capacity;
  uint32_t sample_threshold; // keep a candidate when hash < threshold
  uint32_t salt;
  int32_t *overflow;     // set to 1 if an intermediate left the int32 range
};

__device__ __forceinline__ uint32_t mix(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return h;
}

// Workgroups are dispatched round-robin over the 8 XCDs (bid % 8); giving each
// XCD one contiguous range of tiles puts neighbouring tiles, which share the
// cache lines at their edges (a strided tile row is 64 bytes), in one L2.
__device__ __forceinline__ int64_t xcd_tile(uint32_t bid, uint32_t nb) {
  constexpr uint32_t kXcds = 8;
  uint32_t const xcd = bid % kXcds, local = bid / kXcds;
  uint32_t const q = nb / kXcds, r = nb % kXcds;
  return xcd < r ? int64_t(xcd) * (q + 1) + local
                 : int64_t(r) * (q + 1) + int64_t(xcd - r) * q + local;
}

// Orders one wavefront's LDS accesses around an exchange between its lanes
// (a wavefront's LDS instructions execute in program order; this keeps the
// compiler from moving loads and stores across it).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool STRIDED>
__device__ __forceinline__ int lidx(int line, int r, int m) {
  return STRIDED ? r * kTile + line : line * m + r;
}

// KIND: the network's input features (kFeatures*).
template <bool STRIDED, int32_t KIND>
__global__ __launch_bounds__(kThreads) void learned_sweep(LearnedArgs a) {
  extern __shared__ int32_t tile[];
  __shared__ int32_t overflow_flag;
  __shared__ u32x4 xfer[kThreads]; // network inputs, one 16-byte slot per lane
  int const tid = threadIdx.x;
  int const lane = tid & 63;
  int const wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  int const m = a.m;
  int const n = 1 << a.L;
  int64_t const lines = a.A * a.B;
  int64_t const line0 = xcd_tile(blockIdx.x, gridDim.x) * kTile;
  int const nl = static_cast<int>(lines - line0 < kTile ? lines - line0 : kTile);
  if (tid == 0)
    overflow_flag = 0;

  Net net;
  bool const any_mlp = (a.mlp_mask >> 3) != 0 || a.pairs != nullptr;
  if (any_mlp && a.weights != nullptr)
    load_net(a.weights, net);

  // line base offsets: lines are the (a, b) pairs of the [A][m][B] view
  __shared__ int64_t line_base[kTile];
  if (tid < kTile) {
    int64_t const ln = line0 + tid;
    int64_t const aa = ln / a.B;
    line_base[tid] = aa * m * a.B + (ln - aa * a.B);
  }
  __syncthreads();
  // load: strided, the tile's columns of one row are consecutive in memory;
  // contiguous (B == 1), the tile is nl * m consecutive elements
  if constexpr (STRIDED) {
    for (int e = tid; e < kTile * m; e += kThreads) {
      int const l = e & (kTile - 1), r = e / kTile;
      if (l < nl)
        tile[r * kTile + l] = a.data[line_base[l] + int64_t(r) * a.B];
    }
  } else {
    int32_t const *src = a.data + line_base[0];
    for (int e = tid; e < nl * m; e += kThreads)
      tile[e] = src[e];
  }
  __syncthreads();
  int const nmask = n - 1;

  int const nlev = a.L - a.lmin;
  for (int step = 0; step < nlev; ++step) {
    int const lev = a.inverse ? a.lmin + 1 + step : a.L - step;
    int const s = 1 << (a.L - lev);
    int const cnt = 1 << (lev - 1); // odd points per line
    // inverse: undo the update first
    if (a.update && a.inverse) {
      for (int p = tid; p < kTile * cnt; p += kThreads) {
        int const l = p & (kTile - 1), k = p >> kLogTile;
        if (l >= nl)
          continue;
        int const e = 2 * s * k;
        int const lo = (e - s) & nmask;
        int64_t const u = (int64_t(tile[lidx<STRIDED>(l, lo, m)]) +
                           tile[lidx<STRIDED>(l, e + s, m)] + 2) >> 2;
        tile[lidx<STRIDED>(l, e, m)] -= static_cast<int32_t>(u);
      }
      __syncthreads();
    }
    bool const use_mlp = lev >= 3 && ((a.mlp_mask >> lev) & 1u);
    bool const want_pairs = a.pairs != nullptr && lev >= 3;
    bool const want_stats = a.stats != nullptr && lev >= 3 && use_mlp;
    if (use_mlp || want_pairs || want_stats) {
      // batches of 16 points per wavefront: one odd point index k across the
      // tile's 16 lines (lane & 15 = line), so the neighbour reads of a batch
      // are 16 consecutive LDS words (strided layout) or words m apart (m odd)
      double l1_mlp = 0.0, l1_hat = 0.0;
      constexpr float bound = correction_bound<KIND>();
      int const l = lane & (kTile - 1);
      int const q = lane >> 4;
      // A wavefront takes 4 odd point indices k at a time across the tile's
      // 16 lines: lane (q, l) owns point (line l, k = k0 + q), gathers its
      // neighbours and computes its features (64 distinct points per
      // instruction; the neighbour reads of 16 lanes are 16 consecutive LDS
      // words in the strided layout, words m apart (m odd) otherwise).  The
      // network then runs as 4 MFMA groups of 16 points: lane (q, l) stores
      // its point's 8 bf16 features in its 16-byte slot of the wave's LDS
      // row, group g reads the B fragment of its 16 points back (lane
      // (q', l): 8 bytes, features 4(q' & 1)..+3 of point (g, l); lane groups
      // 2 and 3 read features too, against zero weights), and lane (q, l)
      // keeps the output of group q.
      // Neighbours x -/+ 7s .. s wrap around the period only for k < 3 or
      // k > cnt - 5; a step whose 4 indices k0..k0+3 avoid both (a test on
      // wave-uniform values) addresses them at fixed offsets from x.
      int const row_stride = STRIDED ? kTile : 1;
      for (int k0 = 4 * wave; k0 < cnt; k0 += 4 * (kThreads / 64)) {
        int const k = k0 + q;
        bool const valid = l < nl && k < cnt;
        int const kk = k < cnt ? k : k0;
        int const x = s + 2 * s * kk;
        int const xi = lidx<STRIDED>(l, x, m);
        int c[8];
        if (k0 >= 3 && k0 + 3 <= cnt - 5) {
          int32_t const *px = tile + xi;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            int const off = (2 * (3 - j) + 1) * s * row_stride; // 7s, 5s, 3s, s rows
            c[j] = px[-off];
            c[7 - j] = px[off];
          }
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            int const off = (2 * (3 - j) + 1) * s;
            c[j] = tile[lidx<STRIDED>(l, (x - off) & nmask, m)];
            c[7 - j] = tile[lidx<STRIDED>(l, (x + off) & nmask, m)];
          }
        }
        float u[8], norm;
        float const unit = features<KIND>(c, u, norm);
        float r = 0.0f; // network output
        if (a.weights != nullptr) {
          u32x4 word;
#pragma unroll
          for (int j = 0; j < 4; ++j) // bf16 pairs, one v_cvt_pk_bf16_f32 each
            word[j] = __builtin_bit_cast(
                uint32_t, __builtin_convertvector(f32x2_t{u[2 * j], u[2 * j + 1]}, bf16x2_t));
          xfer[tid] = word;
          wave_sync();
          u32x2_t const *row = reinterpret_cast<u32x2_t const *>(xfer + (tid & ~63));
#pragma unroll
          for (int g0 = 0; g0 < 4; g0 += kGroupsInFlight) {
            bf16x4 xs[kGroupsInFlight];
            float out[kGroupsInFlight];
#pragma unroll
            for (int b = 0; b < kGroupsInFlight; ++b)
              xs[b] = __builtin_bit_cast(bf16x4, row[2 * (16 * (g0 + b) + l) + (q & 1)]);
            mlp16<kGroupsInFlight>(net, xs, out);
#pragma unroll
            for (int b = 0; b < kGroupsInFlight; ++b)
              r = g0 + b == q ? out[b] : r;
          }
          wave_sync(); // every lane's reads before the next iteration's stores
        }
        int64_t const hat = (int64_t(c[3]) + c[4] + 1) >> 1;
        // clamped to +-bound units and to +-2^30: within that the correction
        // converts to int32 in one instruction; a larger one would overflow
        // the int32 detail anyway (flagged below)
        float const lim = __builtin_amdgcn_fmed3f(bound * unit, 0.0f, 1073741824.0f);
        float const corr = __builtin_amdgcn_fmed3f(unit * r, -lim, lim);
        int64_t const pm = hat + __float2int_rn(corr);
        int32_t const xv = tile[xi];
        if (!valid)
          continue;
        if (want_pairs && !a.inverse) {
          uint32_t const key = mix(static_cast<uint32_t>(line0 + l) * 2654435761u ^
                                   mix(static_cast<uint32_t>(x) ^ (a.salt << 8) ^
                                       static_cast<uint32_t>(lev)));
          if (key < a.sample_threshold && norm > 0.0f) {
            uint32_t const slot = atomicAdd(a.pair_count, 1u);
            if (slot < a.pair_capacity) {
              float *pr = a.pairs + size_t(slot) * 10;
              for (int j = 0; j < 8; ++j)
                pr[j] = fminf(fmaxf(u[j], -1.0f), 1.0f);
              // the predictor clamps its correction to +-bound units, so
              // targets beyond that cannot be fitted: clamp them too
              pr[8] = fminf(fmaxf(static_cast<float>(int64_t(xv) - hat) / unit, -bound), bound);
              pr[9] = __uint_as_float(key);
            }
          }
        }
        if (want_stats) {
          l1_mlp += fabs(double(int64_t(xv) - pm));
          l1_hat += fabs(double(int64_t(xv) - hat));
        }
        int64_t const pred = use_mlp ? pm : hat;
        int64_t const out = a.inverse ? int64_t(xv) + pred : int64_t(xv) - pred;
        if (out > INT32_MAX || out < INT32_MIN)
          overflow_flag = 1;
        tile[xi] = static_cast<int32_t>(out);
      }
      if (want_stats) {
        atomicAdd(&a.stats[2 * lev + 0], l1_mlp);
        atomicAdd(&a.stats[2 * lev + 1], l1_hat);
      }
    } else {
      for (int p = tid; p < kTile * cnt; p += kThreads) {
        int const l = p & (kTile - 1), k = p >> kLogTile;
        if (l >= nl)
          continue;
        int const x = s + 2 * s * k;
        int const lo = x - s;
        int const hi = (x + s) & nmask;
        int64_t const hat = (int64_t(tile[lidx<STRIDED>(l, lo, m)]) +
                             tile[lidx<STRIDED>(l, hi, m)] + 1) >> 1;
        int64_t const xv = tile[lidx<STRIDED>(l, x, m)];
        int64_t const out = a.inverse ? xv + hat : xv - hat;
        if (out > INT32_MAX || out < INT32_MIN)
          overflow_flag = 1;
        tile[lidx<STRIDED>(l, x, m)] = static_cast<int32_t>(out);
      }
    }
    __syncthreads();
    if (a.update && !a.inverse) {
      for (int p = tid; p < kTile * cnt; p += kThreads) {
        int const l = p & (kTile - 1), k = p >> kLogTile;
        if (l >= nl)
          continue;
        int const e = 2 * s * k;
        int const lo = (e - s) & nmask;
        int64_t const u = (int64_t(tile[lidx<STRIDED>(l, lo, m)]) +
                           tile[lidx<STRIDED>(l, e + s, m)] + 2) >> 2;
        int64_t const out = int64_t(tile[lidx<STRIDED>(l, e, m)]) + u;
        if (out > INT32_MAX || out < INT32_MIN)
          overflow_flag = 1;
        tile[lidx<STRIDED>(l, e, m)] = static_cast<int32_t>(out);
      }
      __syncthreads();
    }
  }
  // periodic image
  if (m != n)
    for (int l = tid; l < nl; l += kThreads)
      tile[lidx<STRIDED>(l, n, m)] = tile[lidx<STRIDED>(l, 0, m)];
  __syncthreads();
  if constexpr (STRIDED) {
    for (int e = tid; e < kTile * m; e += kThreads) {
      int const l = e & (kTile - 1), r = e / kTile;
      if (l < nl)
        a.data[line_base[l] + int64_t(r) * a.B] = tile[r * kTile + l];
    }
  } else {
    int32_t *dst = a.data + line_base[0];
    for (int e = tid; e < nl * m; e += kThreads)
      dst[e] = tile[e];
  }
  if (tid == 0 && overflow_flag != 0 && a.overflow != nullptr)
    atomicOr(a.overflow, 1);
}

// ---------------------------------------------------------------------------
// quantization

__global__ void quantize_kernel(float const *in, int32_t *out, int64_t count,
                                float inv_scale) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < count;
       i += int64_t(gridDim.x) * blockDim.x)
    out[i] = static_cast<int32_t>(rintf(in[i] * inv_scale));
}

__global__ void dequantize_kernel(int32_t const *in, float *out, int64_t count,
                                  float scale) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < count;
       i += int64_t(gridDim.x) * blockDim.x)
    out[i] = static_cast<float>(in[i]) * scale;
}

// ---------------------------------------------------------------------------
// MLP evaluation on given inputs (MFMA path, as inside the transform)

__global__ __launch_bounds__(kThreads) void mlp_eval_kernel(
    float const *inputs, float const *weights, float *out, int32_t count) {
  Net net;
  load_net(weights, net);
  int const lane = threadIdx.x & 63;
  int const wave_global = (blockIdx.x * kThreads + threadIdx.x) >> 6;
  int const waves = (gridDim.x * kThreads) >> 6;
  int const q = lane >> 4;
  for (int b = wave_global; b * 16 < count; b += waves) {
    int const p = b * 16 + (lane & 15);
    bool const valid = p < count;
    bf16x4 x;
#pragma unroll
    for (int j = 0; j < 4; ++j) // lane groups 2, 3: zeros against zero weights
      x[j] = to_bf16(q < 2 && valid ? inputs[size_t(p) * 8 + 4 * q + j] : 0.0f);
    bf16x4 const xs[1] = {x};
    float r[1];
    mlp16<1>(net, xs, r);
    if (valid && lane < 16)
      out[p] = r[0];
  }
}

// ---------------------------------------------------------------------------
// training: loss and gradient of the mean absolute error over a batch
//   loss = mean_i |mlp(u_i) - t_i|,  d loss / d r_i = sign(r_i - t_i) / n
// (0 where r_i == t_i).  The absolute error is the detail L1 the predict is
// meant to lower; its minimizer is the conditional median of the target,
// which the heavy tails of the residuals near fronts do not pull the way
// they pull the mean a squared error fits.
// Forward in fp32 on bf16-rounded weights, inputs and activations, fp32
// biases and the fp16 gate (the arithmetic of the MFMA path); backward in
// fp32 through the rounded values.
// Each workgroup of 64 rows writes its partial gradient and loss to a
// workspace row; reduce_partials then sums the rows in block order, so the
// result does not depend on the order in which workgroups finish.

constexpr int kTrainThreads = 64;
constexpr int kPartial = kParams + 1; // gradient, then loss

__global__ __launch_bounds__(kTrainThreads) void mlp_grad_kernel(
    float const *pairs, int32_t stride, int32_t const *index, int32_t count,
    float const *weights, float *partials, float scale) {
  __shared__ float W2[kHidden][kHidden + 1];
  __shared__ float sh_h1[kTrainThreads][kHidden + 1];
  __shared__ float sh_d2[kTrainThreads][kHidden + 1];
  __shared__ float sh_h2[kTrainThreads][kHidden + 1];
  __shared__ float sh_d1[kTrainThreads][kHidden + 1];
  __shared__ float sh_in[kTrainThreads][kInputs + 1];
  __shared__ float sh_d3[kTrainThreads];
  __shared__ float sh_loss[kTrainThreads];
  int const tid = threadIdx.x;
  for (int e = tid; e < kHidden * kHidden; e += kTrainThreads)
    W2[e / kHidden][e % kHidden] = round_bf16(weights[kW2 + e]);
  __syncthreads();
  int const i = blockIdx.x * kTrainThreads + tid;
  bool const valid = i < count;
  float u[kInputs];
  float t = 0.0f;
  if (valid) {
    int64_t const row = index != nullptr ? index[i] : i;
    for (int k = 0; k < kInputs; ++k)
      u[k] = round_bf16(pairs[row * stride + k]);
    t = pairs[row * stride + kInputs];
  } else {
    for (int k = 0; k < kInputs; ++k)
      u[k] = 0.0f;
  }
  float z1[kHidden], h1[kHidden];
  for (int h = 0; h < kHidden; ++h) {
    float acc = weights[kB1 + h];
    for (int k = 0; k < kInputs; ++k)
      acc += round_bf16(weights[kW1 + h * kInputs + k]) * u[k];
    z1[h] = acc;
    h1[h] = round_bf16(gelu(acc));
  }
  float r = weights[kB3];
  float z2[kHidden];
  for (int h = 0; h < kHidden; ++h) {
    float acc = 0.0f;
    for (int k = 0; k < kHidden; ++k)
      acc += W2[h][k] * h1[k];
    acc += weights[kB2 + h];
    z2[h] = acc;
    float const h2 = round_bf16(gelu(acc));
    sh_h2[tid][h] = valid ? h2 : 0.0f;
    r += round_bf16(weights[kW3 + h]) * h2;
  }
  float const err = valid ? r - t : 0.0f;
  // d loss / d r, scale = 1/batch
  float const d3 = (err > 0.0f ? 1.0f : (err < 0.0f ? -1.0f : 0.0f)) * scale;
  sh_loss[tid] = fabsf(err) * scale;
  sh_d3[tid] = d3;
  for (int h = 0; h < kHidden; ++h) {
    float const d2 = d3 * round_bf16(weights[kW3 + h]) * gelu_grad(z2[h]);
    sh_d2[tid][h] = d2;
  }
  __syncthreads();
  for (int k = 0; k < kHidden; ++k) {
    float acc = 0.0f;
    for (int h = 0; h < kHidden; ++h)
      acc += W2[h][k] * sh_d2[tid][h];
    sh_d1[tid][k] = acc * gelu_grad(z1[k]);
    sh_h1[tid][k] = h1[k];
  }
  for (int k = 0; k < kInputs; ++k)
    sh_in[tid][k] = u[k];
  __syncthreads();
  // block reductions over the batch rows held in LDS
  for (int e = tid; e < kPartial; e += kTrainThreads) {
    float acc = 0.0f;
    if (e == kParams) {
      for (int b = 0; b < kTrainThreads; ++b)
        acc += sh_loss[b];
    } else if (e < kB1) {
      int const h = e / kInputs, k = e % kInputs;
      for (int b = 0; b < kTrainThreads; ++b)
        acc += sh_d1[b][h] * sh_in[b][k];
    } else if (e < kW2) {
      int const h = e - kB1;
      for (int b = 0; b < kTrainThreads; ++b)
        acc += sh_d1[b][h];
    } else if (e < kB2) {
      int const h = (e - kW2) / kHidden, k = (e - kW2) % kHidden;
      for (int b = 0; b < kTrainThreads; ++b)
        acc += sh_d2[b][h] * sh_h1[b][k];
    } else if (e < kW3) {
      int const h = e - kB2;
      for (int b = 0; b < kTrainThreads; ++b)
        acc += sh_d2[b][h];
    } else if (e < kB3) {
      int const h = e - kW3;
      for (int b = 0; b < kTrainThreads; ++b)
        acc += sh_d3[b] * sh_h2[b][h];
    } else {
      for (int b = 0; b < kTrainThreads; ++b)
        acc += sh_d3[b];
    }
    partials[int64_t(blockIdx.x) * kPartial + e] = acc;
  }
}

__global__ void reduce_partials(float const *partials, int32_t blocks,
                                float *grad, float *loss_sum) {
  int const e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kPartial)
    return;
  // block order, 16 loads in flight at a time (the adds stay in that order)
  float acc = 0.0f;
  int b = 0;
  for (; b + 16 <= blocks; b += 16) {
    float v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j)
      v[j] = partials[int64_t(b + j) * kPartial + e];
#pragma unroll
    for (int j = 0; j < 16; ++j)
      acc += v[j];
  }
  for (; b < blocks; ++b)
    acc += partials[int64_t(b) * kPartial + e];
  if (e < kParams)
    grad[e] += acc;
  else
    *loss_sum += acc;
}

// Adam on the fp32 master weights; clears the gradient for the next step.
__global__ void adam_kernel(float *weights, float *grad, float *m1, float *m2,
                            float lr, float beta1, float beta2, float eps,
                            float corr1, float corr2) {
  int const e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kParams)
    return;
  float const g = grad[e];
  float const a = beta1 * m1[e] + (1.0f - beta1) * g;
  float const b = beta2 * m2[e] + (1.0f - beta2) * g * g;
  m1[e] = a;
  m2[e] = b;
  weights[e] -= lr * (a / corr1) / (sqrtf(b / corr2) + eps);
  grad[e] = 0.0f;
}

__global__ void fill_kernel(float *p, int32_t count, float v) {
  int const e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < count)
    p[e] = v;
}

uint32_t blocks_for(int64_t work, int threads) {
  int64_t b = (work + threads - 1) / threads;
  return static_cast<uint32_t>(b < 1 ? 1 : (b > (1 << 20) ? (1 << 20) : b));
}

} // namespace learned

// Partial-sum floats of one gradient evaluation over `count` rows (one row of
// kPartial per workgroup of kTrainThreads rows).
int64_t learned_gradient_floats(int32_t count) {
  using namespace learned;
  return count < 1 ? 0 : int64_t(blocks_for(count, kTrainThreads)) * kPartial;
}

// Largest row count of a gradient evaluation (one workgroup per 64 rows, at
// most 2^20 workgroups).
int32_t learned_gradient_max_rows() { return (1 << 20) * learned::kTrainThreads; }

// The two launches of a gradient evaluation: per-workgroup partial sums into
// `partials` (learned_gradient_floats(count) floats, every one written), then
// their sum in workgroup order into grad and *loss_sum.
hipError_t launch_learned_gradient(float const *pairs, int32_t stride,
                                   int32_t const *index, int32_t count,
                                   float const *weights, float *grad,
                                   float *loss_sum, float *partials,
                                   hipStream_t stream) {
  using namespace learned;
  int32_t const blocks = static_cast<int32_t>(blocks_for(count, kTrainThreads));
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(kTrainThreads), 0, stream,
                     pairs, stride, index, count, weights, partials,
                     1.0f / float(count));
  hipLaunchKernelGGL(reduce_partials, dim3((kPartial + 255) / 256), dim3(256), 0,
                     stream, partials, blocks, grad, loss_sum);
  return hipGetLastError();
}

} // namespace paliwa_hip

using namespace paliwa_hip;
using namespace paliwa_hip::learned;

namespace {

int hip_status(hipError_t err) {
  if (err != hipSuccess)
    return fail(PALIWA_ERR_HIP, std::string("HIP: ") + hipGetErrorString(err));
  return PALIWA_OK;
}

int check_int_grid(void const *data, int32_t ndim, int64_t const *extents,
                   int32_t const *level, int32_t const *lmin) {
  int64_t elements = 0;
  // the extents rules are those of the float grids
  int rc = check_grid(data, PALIWA_FLOAT32, ndim, extents, level, nullptr,
                      &elements);
  if (rc != PALIWA_OK)
    return rc;
  if (lmin == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "minimum_level is required");
  for (int d = 0; d < ndim; ++d) {
    if (lmin[d] < 0 || lmin[d] > level[d])
      return fail(PALIWA_ERR_LEVEL, "require 0 <= minimum_level <= level");
    int64_t const m = extents[d];
    if (m * kTile * int64_t(sizeof(int32_t)) > 160 * 1024)
      return fail(PALIWA_ERR_UNSUPPORTED,
                  "learned lifting holds lines of at most 2^11 points");
  }
  return PALIWA_OK;
}

int learned_run(int32_t *data, int32_t ndim, int64_t const *extents,
                int32_t const *level, int32_t const *lmin, int32_t axis,
                float const *weights, uint32_t const *mlp_mask, int32_t update,
                int32_t feature_kind, bool inverse, double *stats, float *pairs,
                uint32_t *pair_count, uint32_t capacity, uint32_t threshold,
                uint32_t salt, int32_t *overflow, hipStream_t stream) {
  int rc = check_int_grid(data, ndim, extents, level, lmin);
  if (rc != PALIWA_OK)
    return rc;
  if (axis < -1 || axis >= ndim)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "axis out of range");
  if (feature_kind != kFeaturesRaw && feature_kind != kFeaturesDetrended)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "unknown feature kind");
  for (int i = 0; i < ndim; ++i) {
    int const d = inverse ? ndim - 1 - i : i;
    if (axis >= 0 && d != axis)
      continue;
    SweepParams p = make_params(data, ndim, extents, level, d);
    if (level[d] == lmin[d])
      continue;
    LearnedArgs a{};
    a.data = data;
    a.A = p.A;
    a.B = p.B;
    a.m = p.m;
    a.L = p.L;
    a.lmin = lmin[d];
    a.inverse = inverse ? 1 : 0;
    a.update = update;
    a.mlp_mask = mlp_mask != nullptr ? mlp_mask[d] : 0u;
    a.features = feature_kind;
    a.weights = weights;
    a.stats = stats != nullptr ? stats + size_t(d) * 2 * (PALIWA_MAX_LEVEL + 1) : nullptr;
    a.pairs = pairs;
    a.pair_count = pair_count;
    a.pair_capacity = capacity;
    a.sample_threshold = threshold;
    a.salt = salt * 131u + uint32_t(d);
    a.overflow = overflow;
    if ((a.mlp_mask >> 3) != 0 && weights == nullptr)
      return fail(PALIWA_ERR_INVALID_ARGUMENT, "weights are required");
    // levels lmin+1 .. L that would run the network
    uint32_t const in_range = ((2u << p.L) - 1u) & ~((2u << lmin[d]) - 1u) & ~7u;
    if ((a.mlp_mask & in_range) == 0 && pairs == nullptr) {
      // interpolet only: the fused register ladder on int32, whose passes
      // compute x + sign * floor((lo + hi + bias) / 2^shift)
      int const nlev = p.L - lmin[d];
      p.nlevels = nlev;
      for (int i = 0; i < nlev; ++i)
        p.shift[i] = static_cast<int8_t>(inverse ? nlev - 1 - i : i);
      struct Pass {
        int8_t odd;
        double sign, bias, shift;
      };
      Pass const predict{1, inverse ? 1.0 : -1.0, 1.0, 1.0};
      Pass const upd{0, inverse ? -1.0 : 1.0, 2.0, 2.0};
      Pass passes[2] = {predict, upd};
      p.npass = update ? 2 : 1;
      if (update && inverse) {
        passes[0] = upd;
        passes[1] = predict;
      }
      for (int f = 0; f < p.npass; ++f) {
        p.offset[f] = passes[f].odd;
        p.coef[f][0] = passes[f].sign;
        p.coef[f][1] = passes[f].bias;
        p.coef[f][2] = passes[f].shift;
      }
      p.overflow = overflow;
      classify(p);
      if ((rc = hip_status(launch_fused_i32(p, stream))) != PALIWA_OK)
        return rc;
      continue;
    }
    int64_t const tiles = (p.A * p.B + kTile - 1) / kTile;
    size_t const bytes = size_t(kTile) * p.m * sizeof(int32_t);
    bool const strided = p.B > 1;
    bool const detrended = feature_kind == kFeaturesDetrended;
    void (*kern)(LearnedArgs) =
        strided ? (detrended ? learned_sweep<true, kFeaturesDetrended>
                             : learned_sweep<true, kFeaturesRaw>)
                : (detrended ? learned_sweep<false, kFeaturesDetrended>
                             : learned_sweep<false, kFeaturesRaw>);
    if (bytes > 65536) {
      rc = hip_status(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          static_cast<int>(bytes)));
      if (rc != PALIWA_OK)
        return rc;
    }
    hipLaunchKernelGGL(kern, dim3(uint32_t(tiles)), dim3(kThreads), bytes, stream, a);
    if ((rc = hip_status(hipGetLastError())) != PALIWA_OK)
      return rc;
  }
  last_error().clear();
  return PALIWA_OK;
}

} // namespace

extern "C" {

int paliwa_learned_param_count(void) { return kParams; }

int paliwa_quantize_f32(float const *in, int32_t *out, int64_t count,
                        float scale, void *stream) {
  if (in == nullptr || out == nullptr || count < 0 || !(scale > 0.0f))
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "bad quantization arguments");
  hipLaunchKernelGGL(quantize_kernel, dim3(blocks_for(count, 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), in, out, count, 1.0f / scale);
  return hip_status(hipGetLastError());
}

int paliwa_dequantize_f32(int32_t const *in, float *out, int64_t count,
                          float scale, void *stream) {
  if (in == nullptr || out == nullptr || count < 0 || !(scale > 0.0f))
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "bad quantization arguments");
  hipLaunchKernelGGL(dequantize_kernel, dim3(blocks_for(count, 256)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), in, out, count, scale);
  return hip_status(hipGetLastError());
}

int paliwa_learned_hierarchize(int32_t *data, int32_t ndim,
                               int64_t const *extents, int32_t const *level,
                               int32_t const *minimum_level, int32_t axis,
                               float const *weights, uint32_t const *mlp_mask,
                               int32_t update, int32_t features, double *stats,
                               int32_t *overflow, void *stream) {
  return learned_run(data, ndim, extents, level, minimum_level, axis, weights,
                     mlp_mask, update, features, false, stats, nullptr,
                     nullptr, 0, 0, 0, overflow,
                     static_cast<hipStream_t>(stream));
}

int paliwa_learned_dehierarchize(int32_t *data, int32_t ndim,
                                 int64_t const *extents, int32_t const *level,
                                 int32_t const *minimum_level, int32_t axis,
                                 float const *weights,
                                 uint32_t const *mlp_mask, int32_t update,
                                 int32_t features, int32_t *overflow,
                                 void *stream) {
  return learned_run(data, ndim, extents, level, minimum_level, axis, weights,
                     mlp_mask, update, features, true, nullptr, nullptr,
                     nullptr, 0, 0, 0, overflow,
                     static_cast<hipStream_t>(stream));
}

int paliwa_learned_extract_pairs(int32_t *data, int32_t ndim,
                                 int64_t const *extents, int32_t const *level,
                                 int32_t const *minimum_level, int32_t update,
                                 int32_t features, float *pairs,
                                 uint32_t *pair_count, uint32_t capacity,
                                 uint32_t sample_threshold, uint32_t salt,
                                 void *stream) {
  if (pairs == nullptr || pair_count == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "pairs and pair_count are required");
  return learned_run(data, ndim, extents, level, minimum_level, -1, nullptr,
                     nullptr, update, features, false, nullptr, pairs,
                     pair_count, capacity, sample_threshold, salt, nullptr,
                     static_cast<hipStream_t>(stream));
}

int paliwa_learned_mlp_eval(float const *inputs, float const *weights,
                            float *out, int32_t count, void *stream) {
  if (inputs == nullptr || weights == nullptr || out == nullptr || count < 0)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "bad arguments");
  hipLaunchKernelGGL(mlp_eval_kernel, dim3(blocks_for(int64_t(count) * 4, kThreads)),
                     dim3(kThreads), 0, static_cast<hipStream_t>(stream), inputs,
                     weights, out, count);
  return hip_status(hipGetLastError());
}

int64_t paliwa_learned_gradient_workspace(int32_t count) {
  return learned_gradient_floats(count);
}

int paliwa_learned_gradient(float const *pairs, int32_t stride,
                            int32_t const *index, int32_t count,
                            float const *weights, float *grad,
                            float *loss_sum, float *workspace, void *stream) {
  if (pairs == nullptr || weights == nullptr || grad == nullptr ||
      loss_sum == nullptr || workspace == nullptr || count < 1 || stride < 9)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "bad arguments");
  if (count > learned_gradient_max_rows())
    return fail(PALIWA_ERR_UNSUPPORTED, "at most 2^26 rows per gradient");
  return hip_status(launch_learned_gradient(pairs, stride, index, count, weights, grad,
                                            loss_sum, workspace,
                                            static_cast<hipStream_t>(stream)));
}

int paliwa_learned_adam(float *weights, float *grad, float *m1, float *m2,
                        float lr, float beta1, float beta2, float eps,
                        int32_t step, void *stream) {
  if (weights == nullptr || grad == nullptr || m1 == nullptr || m2 == nullptr ||
      step < 1)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "bad arguments");
  float const c1 = 1.0f - std::pow(beta1, float(step));
  float const c2 = 1.0f - std::pow(beta2, float(step));
  hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0,
                     static_cast<hipStream_t>(stream), weights, grad, m1, m2, lr,
                     beta1, beta2, eps, c1, c2);
  return hip_status(hipGetLastError());
}

int paliwa_learned_fill(float *p, int32_t count, float value, void *stream) {
  if (p == nullptr || count < 0)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "bad arguments");
  hipLaunchKernelGGL(fill_kernel, dim3((count + 255) / 256), dim3(256), 0,
                     static_cast<hipStream_t>(stream), p, count, value);
  return hip_status(hipGetLastError());
}

} // extern "C"
