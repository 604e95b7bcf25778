// The fused axis-sweep kernel (device code, included by fused_f64.hip,
// fused_f32.hip and fused_i32.hip, one translation unit per value type so they
// compile in parallel).
//
// A workgroup of 256 threads owns a tile of TB = 256/G whole lines.  Each line
// of period n = K*G is split into G chunks of K consecutive points, one chunk
// per lane, held in registers; the G lanes of a line are consecutive lanes of
// one wavefront.  Every level and every pass of the sweep runs on those
// registers:
//   * a level whose storage stride S is below K updates the points of the pass
//     inside the chunk; the single neighbour outside the chunk comes from the
//     adjacent lane by a wavefront shuffle (periodic within the line);
//   * a level whose stride is a multiple of K involves only the first point of
//     each chunk; its two neighbours are q = S/K lanes away (two shuffles).
// So the sweep costs one HBM read and one HBM write per point whatever the
// number of levels and passes.
//
// Staging.  Lines along the contiguous axis (B == 1) are copied between HBM
// and LDS as one flat run (the tile is contiguous in memory); lines along a
// strided axis are copied row by row, TB consecutive columns per row.  Along a
// strided axis with one lane per line (G == 1, short lines) each lane loads
// its own column straight into registers: the TB lanes of a row read TB
// consecutive elements, which is already coalesced.
//
// Pipelining.  The grid is persistent (as many workgroups as the device holds
// at once); each workgroup walks its tiles and issues all HBM loads of the
// next tile (K+1 per lane, into registers) before it computes and stores the
// current one, so loads stay in flight while LDS, shuffles and stores work.
//
// Level schedules.  fused_ladder covers level lists whose storage strides are
// consecutive powers of two in increasing (hierarchize) or decreasing
// (dehierarchize) order with the pass-parity patterns of the reference's
// tables ({odd}, {odd, even}, {even, odd}); every register index is then a
// compile-time constant.  fused_generic runs any level list and any filter
// list by dispatching on the stride at run time.
#pragma once

#include "lifting.hpp"

#include <type_traits>

namespace paliwa_hip {
namespace fused {

constexpr int kThreads = 256;
constexpr int kXcds = 8;
// Workgroup size (see fused_threads): along a strided axis a tile holds NT/G
// columns and every row of the tile is one run of NT/G consecutive points in
// memory; wider runs waste less of the cache lines at their ends (rows start
// at any 8-byte offset when the inner extent is odd, as with periodic images).
template <typename T> constexpr int threads_for(bool strided, int K, int G) {
  return fused_threads(strided, K, G, int(sizeof(T)), std::is_integral_v<T>);
}

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v / 2); }

template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f.template operator()<B>();
    static_for<B + 1, E>(f);
  }
}

template <int B, int E, typename F>
__device__ __forceinline__ void static_for_desc(F &&f) {
  if constexpr (B < E) {
    f.template operator()<E - 1>();
    static_for_desc<B, E - 1>(f);
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs (bid % 8).  Give each
// XCD one contiguous range of tiles: neighbouring tiles share the cache lines
// at their edges, which then meet in the same L2.
__device__ __forceinline__ int64_t xcd_tile(uint32_t bid, uint32_t nb) {
  uint32_t const xcd = bid % kXcds, local = bid / kXcds;
  uint32_t const q = nb / kXcds, r = nb % kXcds;
  return xcd < r ? int64_t(xcd) * (q + 1) + local
                 : int64_t(r) * (q + 1) + int64_t(xcd - r) * q + local;
}

template <typename T>
__device__ __forceinline__ T lift(T c0, T lo, T c1, T x, T c2, T hi) {
  return c0 * lo + c1 * x + c2 * hi;
}

// Sticky per-workgroup flag of the int32 sweeps: some result left the int32
// range.  Cleared at the start of sweep_tiles, reported at its end.
__device__ __forceinline__ int *int_overflow_flag() {
  __shared__ int flag;
  return &flag;
}

// int32 lifting step (the integer interpolet predict and CDF 5/3 update):
// the coefficient slots carry (sign, bias, shift) and the result is
// x + sign * floor((lo + hi + bias) / 2^shift), evaluated in 64 bits.
__device__ __forceinline__ int32_t lift(int32_t sign, int32_t lo, int32_t bias,
                                        int32_t x, int32_t shift, int32_t hi) {
  int64_t const q = (int64_t(lo) + hi + bias) >> shift;
  int64_t const out = sign > 0 ? int64_t(x) + q : int64_t(x) - q;
  if (out != int64_t(static_cast<int32_t>(out)))
    *int_overflow_flag() = 1;
  return static_cast<int32_t>(out);
}

// Cross-lane moves within a line's group of G consecutive lanes (aligned to
// G).  Groups of 2 and 4 lanes use DPP quad permutations, groups of 8 and 16
// DPP row rotations (row_ror:k gives lane i the value of lane (i - k) mod 16
// of its 16-lane row); groups of 32 and 64 lanes, which no single DPP move
// rotates, use ds_bpermute.  A DPP move is a VALU instruction per 32-bit word.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_word(uint32_t x) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_mov_dpp(static_cast<int>(x), CTRL, 0xF, 0xF, false));
}
template <int CTRL, typename T>
__device__ __forceinline__ T dpp_move(T v) {
  if constexpr (sizeof(T) == 8) {
    uint64_t const b = __builtin_bit_cast(uint64_t, v);
    uint32_t const lo = dpp_word<CTRL>(static_cast<uint32_t>(b));
    uint32_t const hi = dpp_word<CTRL>(static_cast<uint32_t>(b >> 32));
    return __builtin_bit_cast(T, (uint64_t(hi) << 32) | lo);
  } else {
    return __builtin_bit_cast(T, dpp_word<CTRL>(__builtin_bit_cast(uint32_t, v)));
  }
}
constexpr int dpp_row_ror(int k) { return 0x120 + k; } // k in 1..15
// quad_perm control: lane j of a quad takes lane perm[j], the lane d places
// further in j's group of G (G = 2 or 4)
template <int G, int D> constexpr int dpp_quad_rotate() {
  int c = 0;
  for (int j = 0; j < 4; ++j)
    c |= ((j & ~(G - 1)) + (((j % G) + D) % G)) << (2 * j);
  return c;
}

// The value v of lane gbase + ((gi + D) mod G), -G < D < G.
template <typename T, int G, int D>
__device__ __forceinline__ T group_fetch(T v, int gbase, int gi) {
  constexpr int d = ((D % G) + G) % G;
  if constexpr (d == 0) {
    return v;
  } else if constexpr (G <= 4) {
    return dpp_move<dpp_quad_rotate<G, d>()>(v);
  } else if constexpr (G == 16) {
    return dpp_move<dpp_row_ror(16 - d)>(v);
  } else if constexpr (G == 8) { // two rotations of the row, one per side of the group end
    T const ahead = dpp_move<dpp_row_ror(16 - d)>(v); // lane i + d
    T const wrap = dpp_move<dpp_row_ror(8 - d)>(v);   // lane i + d - 8
    return gi + d < 8 ? ahead : wrap;
  } else {
    return __shfl(v, gbase + ((gi + d) & (G - 1)), 64);
  }
}

// One pass at storage stride S < K on this lane's chunk.
template <typename T, int K, int G, int S, int ODD>
__device__ __forceinline__ void pass_in_lane(T (&v)[K], int gbase, int gi,
                                             T c0, T c1, T c2) {
  if constexpr (ODD) {
    T const nb = group_fetch<T, G, 1>(v[0], gbase, gi);
#pragma unroll
    for (int j = S; j < K; j += 2 * S) {
      T const hi = (j + S < K) ? v[j + S] : nb;
      v[j] = lift(c0, v[j - S], c1, v[j], c2, hi);
    }
  } else {
    T const pb = group_fetch<T, G, -1>(v[K - S], gbase, gi);
#pragma unroll
    for (int j = 0; j < K; j += 2 * S) {
      T const lo = (j >= S) ? v[j - S] : pb;
      v[j] = lift(c0, lo, c1, v[j], c2, v[j + S]);
    }
  }
}

// One pass at a stride that is q >= 1 chunks long (q a power of two below G).
template <typename T, int K, int G>
__device__ __forceinline__ void pass_cross_lane(T (&v)[K], int gbase, int gi,
                                                int q, int odd, T c0, T c1,
                                                T c2) {
  T lo, hi;
  if constexpr (G <= 16) {
    auto fetch = [&](auto q_tag) {
      constexpr int Q = decltype(q_tag)::value;
      lo = group_fetch<T, G, -Q>(v[0], gbase, gi);
      hi = group_fetch<T, G, Q>(v[0], gbase, gi);
    };
    if (q == 1)
      fetch(std::integral_constant<int, 1>{});
    else if (G > 4 && q == 4)
      fetch(std::integral_constant<int, (G > 4 ? 4 : 1)>{});
    else if (G > 8 && q == 8)
      fetch(std::integral_constant<int, (G > 8 ? 8 : 1)>{});
    else
      fetch(std::integral_constant<int, (G > 2 ? 2 : 1)>{});
  } else {
    lo = __shfl(v[0], gbase + ((gi - q) & (G - 1)), 64);
    hi = __shfl(v[0], gbase + ((gi + q) & (G - 1)), 64);
  }
  if ((gi & (2 * q - 1)) == odd * q)
    v[0] = lift(c0, lo, c1, v[0], c2, hi);
}

template <int PAT> constexpr int pat_len() { return PAT == kPatOdd ? 1 : 2; }
template <int PAT> constexpr int pat_odd(int f) {
  return PAT == kPatOdd ? 1 : (PAT == kPatOddEven ? (f == 0) : (f != 0));
}

template <typename T> struct Coefs {
  T c[2][3];
};

template <typename T, int K, int G, int PAT, int SH>
__device__ __forceinline__ void level_in_lane(T (&v)[K], int gbase, int gi,
                                              Coefs<T> const &c) {
  static_for<0, pat_len<PAT>()>([&]<int F>() {
    pass_in_lane<T, K, G, (1 << SH), pat_odd<PAT>(F)>(v, gbase, gi, c.c[F][0],
                                                      c.c[F][1], c.c[F][2]);
  });
}

template <typename T, int K, int G, int PAT>
__device__ __forceinline__ void level_cross_lane(T (&v)[K], int gbase, int gi,
                                                 int q, Coefs<T> const &c) {
  static_for<0, pat_len<PAT>()>([&]<int F>() {
    pass_cross_lane<T, K, G>(v, gbase, gi, q, pat_odd<PAT>(F), c.c[F][0],
                             c.c[F][1], c.c[F][2]);
  });
}

// Storage strides 2^lo .. 2^hi, increasing (ASC) or decreasing.
template <typename T, int K, int G, int PAT, bool ASC>
__device__ __forceinline__ void run_ladder(T (&v)[K], int gbase, int gi,
                                           int lo, int hi,
                                           Coefs<T> const &c) {
  constexpr int LOGK = ilog2(K);
  if constexpr (ASC) {
    static_for<0, LOGK>([&]<int SH>() {
      if (SH >= lo && SH <= hi)
        level_in_lane<T, K, G, PAT, SH>(v, gbase, gi, c);
    });
    if constexpr (G > 1)
      for (int sh = lo > LOGK ? lo : LOGK; sh <= hi; ++sh)
        level_cross_lane<T, K, G, PAT>(v, gbase, gi, 1 << (sh - LOGK), c);
  } else {
    if constexpr (G > 1)
      for (int sh = hi; sh >= LOGK && sh >= lo; --sh)
        level_cross_lane<T, K, G, PAT>(v, gbase, gi, 1 << (sh - LOGK), c);
    static_for_desc<0, LOGK>([&]<int SH>() {
      if (SH >= lo && SH <= hi)
        level_in_lane<T, K, G, PAT, SH>(v, gbase, gi, c);
    });
  }
}

// Any level list, any filter list: stride dispatch at run time.
template <typename T, int K, int G>
__device__ __forceinline__ void run_generic(T (&v)[K], int gbase, int gi,
                                            SweepParams const &p) {
  constexpr int LOGK = ilog2(K);
  for (int l = 0; l < p.nlevels; ++l) {
    int const sh = p.shift[l];
    for (int f = 0; f < p.npass; ++f) {
      int const odd = p.offset[f];
      T const c0 = static_cast<T>(p.coef[f][0]);
      T const c1 = static_cast<T>(p.coef[f][1]);
      T const c2 = static_cast<T>(p.coef[f][2]);
      if (sh < LOGK) {
        static_for<0, LOGK>([&]<int SH>() {
          if (sh == SH) {
            if (odd)
              pass_in_lane<T, K, G, (1 << SH), 1>(v, gbase, gi, c0, c1, c2);
            else
              pass_in_lane<T, K, G, (1 << SH), 0>(v, gbase, gi, c0, c1, c2);
          }
        });
      } else {
        if constexpr (G > 1)
          pass_cross_lane<T, K, G>(v, gbase, gi, 1 << (sh - LOGK), odd, c0,
                                   c1, c2);
      }
    }
  }
}

// LDS geometry of a tile of TB lines of ES-byte points (ES = sizeof(T)).
//   contiguous axis: line-major, point r of line l at l*LP + r + (r >> LOGK)*PAD
//   strided axis:    row-major,  point r of column c at r*TB + c + (r >> LOGK)*PAD
// PAD points after every chunk of K rows put the register loads of a
// wavefront's lanes (G lanes per line, chunks K rows apart) on distinct banks
// of the 32 x 4-byte LDS banks: a b64 access serves 16 lanes per cycle, a b32
// access 32, and lane gi of a line lands on bank offset gi * PAD * ES / 4, so
// PAD = 128 / (ES * G) spreads the G lanes (and the lines beside them) over
// all banks; 1 from G = 16 (float64) or 32 (float32) up.  With one lane per
// line (G = 1) the lanes are a line pitch apart, which is odd: no pad, so that
// the staging copy of lines with an image (odd m) runs without gaps.
template <int K, int G, bool STRIDED, typename T, int NTO = 0> struct TileGeom {
  static constexpr int ES = int(sizeof(T));
  static constexpr int NT = NTO ? NTO : threads_for<T>(STRIDED, K, G); // NTO: override
  static constexpr int TB = NT / G;
  static constexpr int LOGK = ilog2(K);
  static constexpr int PAD = G == 1 ? 0 : (128 / (ES * G) < 1 ? 1 : 128 / (ES * G));
  __host__ __device__ static int line_pitch(int m) { return (m + G * PAD) | 1; }
  __host__ __device__ static int lds_elems(int m, bool strided) {
    if (strided && G == 1)
      return 0; // columns go straight to registers
    return strided ? m * TB + G * PAD + 2 : TB * line_pitch(m);
  }
};

// Per-thread staging state of one tile.
//   contiguous axis: the tile is `count` consecutive points from `src`; thread
//     t moves points t, t+256, t+512, ... (the first K of them exist in every
//     full tile; the (K+1)-th only for the first TB threads, with images).
//   strided axis: thread t moves rows row0, row0+G, ..., row0+(K-1)G of its
//     column (always inside the lattice) and, with images, row n (only the
//     threads with row0 == 0).  `lane` points at row row0 of the column and
//     `step` = G*B is the uniform distance between its rows, so the per-lane
//     state is one pointer whatever K is.
template <typename T> struct Stage {
  T *src = nullptr;   // contiguous: first point of the tile
  T *lane = nullptr;  // strided: row row0 of this thread's column (or null)
  int64_t step = 0;   // strided: G * B
  int64_t image = 0;  // strided: offset of row n from `lane` (0: no image row)
  int count = 0;      // contiguous: points in the tile
  bool full = false;  // contiguous: the tile holds TB whole lines
};

// Issue every HBM load of one tile for this thread (K+1 per lane at most,
// all in flight together).
template <typename T, int K, int G, bool STRIDED, int NTO = 0>
__device__ __forceinline__ Stage<T> load_tile(SweepParams const &p,
                                              TileRef const &ref,
                                              T (&buf)[K + 1]) {
  using Geom = TileGeom<K, G, STRIDED, T, NTO>;
  constexpr int TB = Geom::TB;
  int const tid = threadIdx.x;
  int const m = p.m;
  int const n = 1 << p.L;
  T *const g = static_cast<T *>(ref.data);
  Stage<T> st;
  if constexpr (!STRIDED) {
    int64_t const line0 = ref.local * TB;
    int64_t const left = ref.A - line0;
    int const nl = static_cast<int>(left < TB ? left : TB);
    st.src = g + line0 * m;
    st.count = nl * m;
    st.full = nl == TB;
    T const *const s = st.src + tid;
    if (st.full) {
#pragma unroll
      for (int i = 0; i < K; ++i)
        buf[i] = s[i * Geom::NT];
      if (tid + K * Geom::NT < st.count)
        buf[K] = s[K * Geom::NT];
    } else {
#pragma unroll
      for (int i = 0; i < K + 1; ++i)
        if (tid + i * Geom::NT < st.count)
          buf[i] = s[i * Geom::NT];
    }
  } else {
    int const col = tid % TB;
    int const row0 = tid / TB;
    int64_t const c0 = ref.local * TB + col;
    if (c0 < ref.A * ref.B) {
      int64_t const a = c0 / ref.B;
      int64_t const b = c0 - a * ref.B;
      st.lane = g + a * m * ref.B + b + row0 * ref.B;
      st.step = int64_t(G) * ref.B;
      st.image = (m != n && row0 == 0) ? int64_t(n) * ref.B : 0;
#pragma unroll
      for (int i = 0; i < K; ++i)
        buf[i] = st.lane[i * st.step];
      if (st.image != 0)
        buf[K] = st.lane[st.image];
    }
  }
  return st;
}

template <int K, int G, bool STRIDED, typename T, int NTO = 0>
__device__ __forceinline__ int lds_index(int line, int r, int LP) {
  using Geom = TileGeom<K, G, STRIDED, T, NTO>;
  if constexpr (STRIDED)
    return r * Geom::TB + line + (r >> Geom::LOGK) * Geom::PAD;
  else
    return line * LP + r + (r >> Geom::LOGK) * Geom::PAD;
}

// Registers holding a loaded tile -> LDS.
template <typename T, int K, int G, bool STRIDED, int NTO = 0>
__device__ __forceinline__ void stash_tile(SweepParams const &p,
                                           Stage<T> const &st,
                                           T const (&buf)[K + 1], T *lds,
                                           uint32_t magic_m, int LP) {
  using Geom = TileGeom<K, G, STRIDED, T, NTO>;
  int const tid = threadIdx.x;
  int const m = p.m;
  if constexpr (!STRIDED) {
#pragma unroll
    for (int i = 0; i < K + 1; ++i) {
      int const e = tid + i * Geom::NT;
      if ((st.full && i < K) || e < st.count) {
        int const l = __umulhi(static_cast<uint32_t>(e), magic_m);
        int const r = e - l * m;
        lds[lds_index<K, G, false, T, NTO>(l, r, LP)] = buf[i];
      }
    }
  } else {
    int const col = tid % Geom::TB;
    int const row0 = tid / Geom::TB;
    if (st.lane != nullptr) {
#pragma unroll
      for (int i = 0; i < K; ++i)
        lds[lds_index<K, G, true, T, NTO>(col, row0 + i * G, LP)] = buf[i];
      if (st.image != 0)
        lds[lds_index<K, G, true, T, NTO>(col, 1 << p.L, LP)] = buf[K];
    }
  }
}

// LDS -> HBM for one tile.
template <typename T, int K, int G, bool STRIDED, int NTO = 0>
__device__ __forceinline__ void store_tile(SweepParams const &p,
                                           Stage<T> const &st, T const *lds,
                                           uint32_t magic_m, int LP) {
  using Geom = TileGeom<K, G, STRIDED, T, NTO>;
  int const tid = threadIdx.x;
  int const m = p.m;
  T out[K + 1];
  if constexpr (!STRIDED) {
#pragma unroll
    for (int i = 0; i < K + 1; ++i) {
      int const e = tid + i * Geom::NT;
      if ((st.full && i < K) || e < st.count) {
        int const l = __umulhi(static_cast<uint32_t>(e), magic_m);
        int const r = e - l * m;
        out[i] = lds[lds_index<K, G, false, T, NTO>(l, r, LP)];
      }
    }
    T *const s = st.src + tid;
    if (st.full) {
#pragma unroll
      for (int i = 0; i < K; ++i)
        s[i * Geom::NT] = out[i];
      if (tid + K * Geom::NT < st.count)
        s[K * Geom::NT] = out[K];
    } else {
#pragma unroll
      for (int i = 0; i < K + 1; ++i)
        if (tid + i * Geom::NT < st.count)
          s[i * Geom::NT] = out[i];
    }
  } else {
    int const col = tid % Geom::TB;
    int const row0 = tid / Geom::TB;
    if (st.lane != nullptr) {
#pragma unroll
      for (int i = 0; i < K; ++i)
        out[i] = lds[lds_index<K, G, true, T, NTO>(col, row0 + i * G, LP)];
      if (st.image != 0)
        out[K] = lds[lds_index<K, G, true, T, NTO>(col, 1 << p.L, LP)];
#pragma unroll
      for (int i = 0; i < K; ++i)
        st.lane[i * st.step] = out[i];
      if (st.image != 0)
        st.lane[st.image] = out[K];
    }
  }
}

// Persistent kernel body.  Workgroup w processes tiles w', w'+W, w'+2W, ...
// (W = grid size, w' = w regrouped by XCD so that consecutive tiles run on one
// XCD).  The HBM loads of the next tile are issued before the current tile is
// computed and stored, so every workgroup keeps loads in flight throughout.
// MODE 0 = ladder (PAT, ASC used), 1 = generic level/filter lists.
template <typename T, int K, int G, bool STRIDED, int MODE, int PAT, bool ASC,
          int NTO, typename Src>
__device__ __forceinline__ void sweep_tiles(SweepParams const &p,
                                            Src const &src, uint32_t magic_m,
                                            int64_t ntiles, T *lds) {
  using Geom = TileGeom<K, G, STRIDED, T, NTO>;
  constexpr int TB = Geom::TB;
  int const tid = threadIdx.x;
  int const m = p.m;
  int const n = 1 << p.L;
  int const li = tid / G;
  int const gi = tid % G;
  int const gbase = (tid & 63) - gi;
  uint32_t const nwg = gridDim.x;

  Coefs<T> c;
  if constexpr (MODE == 0) {
#pragma unroll
    for (int f = 0; f < pat_len<PAT>(); ++f)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        c.c[f][k] = static_cast<T>(p.coef[f][k]);
  }
  auto compute = [&](T(&v)[K]) {
    if constexpr (MODE == 0)
      run_ladder<T, K, G, PAT, ASC>(v, gbase, gi, p.lo, p.hi, c);
    else
      run_generic<T, K, G>(v, gbase, gi, p);
  };

  constexpr bool kInt = std::is_integral_v<T>;
  if constexpr (kInt) {
    if (tid == 0)
      *int_overflow_flag() = 0;
    __syncthreads();
  }
  int64_t t = xcd_tile(blockIdx.x, nwg);
  if constexpr (STRIDED && G == 1) {
    // Lanes load their own chunks straight between HBM and registers: lane
    // (li, gi) holds rows gi*K .. gi*K+K-1 of column li of the tile.  With
    // one lane per line (G == 1) the TB lanes of a wavefront read TB
    // consecutive columns of one row, which is already coalesced.
    struct Column {
      T *p;      // row gi*K of this lane's column
      int64_t B; // row stride
    };
    auto column = [&](int64_t tile) -> Column {
      if (tile >= ntiles)
        return {nullptr, 1};
      TileRef const ref = src(tile);
      int64_t const col = ref.local * TB + li;
      if (col >= ref.A * ref.B)
        return {nullptr, ref.B};
      int64_t const a = col / ref.B;
      return {static_cast<T *>(ref.data) + a * m * ref.B + (col - a * ref.B) +
                  int64_t(gi) * K * ref.B,
              ref.B};
    };
    T buf[K];
    Column cur = column(t);
    if (cur.p != nullptr) {
#pragma unroll
      for (int j = 0; j < K; ++j)
        buf[j] = cur.p[j * cur.B];
    }
    for (; t < ntiles; t += nwg) {
      T v[K];
#pragma unroll
      for (int j = 0; j < K; ++j)
        v[j] = buf[j];
      Column const next = column(t + nwg);
      if (next.p != nullptr) {
#pragma unroll
        for (int j = 0; j < K; ++j)
          buf[j] = next.p[j * next.B];
      }
      if (cur.p != nullptr) {
        compute(v);
#pragma unroll
        for (int j = 0; j < K; ++j)
          cur.p[j * cur.B] = v[j];
        if (m != n && gi == 0)
          cur.p[static_cast<int64_t>(n) * cur.B] = v[0];
      }
      cur = next;
    }
  } else {
    int const LP = Geom::line_pitch(m);
    T buf[K + 1];
    Stage<T> cur;
    if (t < ntiles)
      cur = load_tile<T, K, G, STRIDED, NTO>(p, src(t), buf);
    for (; t < ntiles; t += nwg) {
      __syncthreads(); // the previous tile's LDS reads are done
      stash_tile<T, K, G, STRIDED, NTO>(p, cur, buf, lds, magic_m, LP);
      __syncthreads();
      Stage<T> nxt;
      if (t + nwg < ntiles)
        nxt = load_tile<T, K, G, STRIDED, NTO>(p, src(t + nwg), buf);
      T v[K];
#pragma unroll
      for (int j = 0; j < K; ++j)
        v[j] = lds[lds_index<K, G, STRIDED, T, NTO>(li, gi * K + j, LP)];
      if constexpr (kInt) {
        // lines past the grid's end (last tile) hold stale LDS words; they
        // are never stored, but their arithmetic must not raise the overflow
        // flag: run them on zeros
        TileRef const ref = src(t);
        int64_t const lines = STRIDED ? ref.A * ref.B : ref.A;
        if (ref.local * TB + li >= lines) {
#pragma unroll
          for (int j = 0; j < K; ++j)
            v[j] = T(0);
        }
      }
      compute(v);
#pragma unroll
      for (int j = 0; j < K; ++j)
        lds[lds_index<K, G, STRIDED, T, NTO>(li, gi * K + j, LP)] = v[j];
      if (m != n && gi == 0)
        lds[lds_index<K, G, STRIDED, T, NTO>(li, n, LP)] = v[0];
      __syncthreads();
      store_tile<T, K, G, STRIDED, NTO>(p, cur, lds, magic_m, LP);
      cur = nxt;
    }
  }
  if constexpr (kInt) {
    __syncthreads();
    if (tid == 0 && *int_overflow_flag() != 0 && p.overflow != nullptr)
      atomicOr(p.overflow, 1);
  }
}

template <typename T, int K, int G, bool STRIDED, int PAT, bool ASC, int NTO = 0>
__global__ __launch_bounds__(NTO ? NTO : threads_for<T>(STRIDED, K, G)) void fused_ladder(
    SweepParams p, uint32_t magic_m, int64_t ntiles) {
  extern __shared__ unsigned char smem_raw[];
  sweep_tiles<T, K, G, STRIDED, 0, PAT, ASC, NTO>(p, OneGrid{p.data, p.A, p.B},
                                             magic_m, ntiles,
                                             reinterpret_cast<T *>(smem_raw));
}

// The same sweep over the tiles of many grids with equal line length along
// the axis (one launch of a batched transform).
template <typename T, int K, int G, bool STRIDED, int PAT, bool ASC, int NTO = 0>
__global__ __launch_bounds__(NTO ? NTO : threads_for<T>(STRIDED, K, G)) void fused_ladder_batched(
    SweepParams p, uint32_t magic_m, int64_t ntiles, ManyGrids grids) {
  extern __shared__ unsigned char smem_raw[];
  sweep_tiles<T, K, G, STRIDED, 0, PAT, ASC, NTO>(p, grids, magic_m, ntiles,
                                             reinterpret_cast<T *>(smem_raw));
}

template <typename T, int K, int G, bool STRIDED>
__global__ __launch_bounds__(threads_for<T>(STRIDED, K, G)) void fused_generic(SweepParams p,
                                                          uint32_t magic_m,
                                                          int64_t ntiles) {
  extern __shared__ unsigned char smem_raw[];
  sweep_tiles<T, K, G, STRIDED, 1, 0, true, 0>(p, OneGrid{p.data, p.A, p.B},
                                            magic_m, ntiles,
                                            reinterpret_cast<T *>(smem_raw));
}

// Number of workgroups of `kernel` the device holds at once (all CUs), cached
// per (kernel, LDS size): the persistent grid.
int resident_workgroups(const void *kernel, size_t lds_bytes, int threads);

template <typename Kernel, typename... Extra>
hipError_t launch_one(Kernel kernel, SweepParams const &p, size_t bytes,
                      int64_t tiles, int threads, hipStream_t stream,
                      Extra... extra) {
  if (bytes > 65536) {
    hipError_t err = hipFuncSetAttribute(
        reinterpret_cast<const void *>(kernel),
        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
    if (err != hipSuccess)
      return err;
  }
  int64_t const resident =
      resident_workgroups(reinterpret_cast<const void *>(kernel), bytes, threads);
  uint32_t const grid =
      static_cast<uint32_t>(tiles < resident ? tiles : resident);
  // e / m by multiply-high: exact while (points per tile) * m < 2^32.
  uint32_t const magic = static_cast<uint32_t>(0xffffffffu / p.m + 1);
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(threads), bytes, stream, p,
                     magic, tiles, extra...);
  return hipGetLastError();
}

// The width of a strided float tile whose 512-thread rows are shorter than 256
// bytes (float64 lines of 512 points and up, float32 of 512 and up) is chosen
// at launch from the line count (ladder_threads_for): the single-grid
// launcher from the grid's lines, a batched plan per group from the group's.
template <typename T, int K, int G, bool STRIDED> constexpr bool width_at_launch() {
  return STRIDED && G >= 4 && G <= 64 && !std::is_integral_v<T> &&
         threads_for<T>(STRIDED, K, G) <= 1024 && (512 / G) * int64_t(sizeof(T)) < 256;
}
template <typename T, int K, int G> constexpr bool wide_fits() {
  return 2 * (int64_t(K) * G + 1) * (512 / G) * int64_t(sizeof(T)) <= kWideTileLds;
}
template <typename T, int K, int G> constexpr int ladder_threads(int64_t lines) {
  return width_at_launch<T, K, G, true>() ? ladder_threads_for(K, G, int(sizeof(T)), lines) : 0;
}

template <typename T, int K, int G, bool STRIDED, int NTO>
hipError_t launch_ladder_nt(SweepParams const &p, int64_t lines, hipStream_t stream) {
  using Geom = TileGeom<K, G, STRIDED, T, NTO>;
  int64_t const tiles = (lines + Geom::TB - 1) / Geom::TB;
  size_t const bytes = sizeof(T) * Geom::lds_elems(p.m, STRIDED);
  bool const asc = p.schedule == kScheduleLadderAsc;
  switch (p.pattern) {
  case kPatOdd:
    return asc ? launch_one(fused_ladder<T, K, G, STRIDED, kPatOdd, true, NTO>, p, bytes, tiles, NTO, stream)
               : launch_one(fused_ladder<T, K, G, STRIDED, kPatOdd, false, NTO>, p, bytes, tiles, NTO, stream);
  case kPatOddEven:
    return asc ? launch_one(fused_ladder<T, K, G, STRIDED, kPatOddEven, true, NTO>, p, bytes, tiles, NTO, stream)
               : launch_one(fused_ladder<T, K, G, STRIDED, kPatOddEven, false, NTO>, p, bytes, tiles, NTO, stream);
  default:
    return asc ? launch_one(fused_ladder<T, K, G, STRIDED, kPatEvenOdd, true, NTO>, p, bytes, tiles, NTO, stream)
               : launch_one(fused_ladder<T, K, G, STRIDED, kPatEvenOdd, false, NTO>, p, bytes, tiles, NTO, stream);
  }
}

template <typename T, int K, int G, bool STRIDED>
hipError_t launch_kg_layout(SweepParams const &p, hipStream_t stream) {
  using Geom = TileGeom<K, G, STRIDED, T>;
  int64_t const lines = p.A * p.B;
  if constexpr (width_at_launch<T, K, G, STRIDED>()) {
    if ((p.schedule == kScheduleLadderAsc || p.schedule == kScheduleLadderDesc) &&
        p.pattern != kPatNone) {
      int const nt = ladder_threads<T, K, G>(lines);
      if constexpr (wide_fits<T, K, G>())
        if (nt == 1024)
          return launch_ladder_nt<T, K, G, STRIDED, 1024>(p, lines, stream);
      return nt == 256 ? launch_ladder_nt<T, K, G, STRIDED, 256>(p, lines, stream)
                       : launch_ladder_nt<T, K, G, STRIDED, 512>(p, lines, stream);
    }
  }
  int64_t const tiles = (lines + Geom::TB - 1) / Geom::TB;
  size_t const bytes = sizeof(T) * Geom::lds_elems(p.m, STRIDED);
  if (p.schedule == kScheduleLadderAsc || p.schedule == kScheduleLadderDesc) {
    bool const asc = p.schedule == kScheduleLadderAsc;
    switch (p.pattern) {
    case kPatOdd:
      return asc ? launch_one(fused_ladder<T, K, G, STRIDED, kPatOdd, true>, p, bytes, tiles, Geom::NT, stream)
                 : launch_one(fused_ladder<T, K, G, STRIDED, kPatOdd, false>, p, bytes, tiles, Geom::NT, stream);
    case kPatOddEven:
      return asc ? launch_one(fused_ladder<T, K, G, STRIDED, kPatOddEven, true>, p, bytes, tiles, Geom::NT, stream)
                 : launch_one(fused_ladder<T, K, G, STRIDED, kPatOddEven, false>, p, bytes, tiles, Geom::NT, stream);
    case kPatEvenOdd:
      return asc ? launch_one(fused_ladder<T, K, G, STRIDED, kPatEvenOdd, true>, p, bytes, tiles, Geom::NT, stream)
                 : launch_one(fused_ladder<T, K, G, STRIDED, kPatEvenOdd, false>, p, bytes, tiles, Geom::NT, stream);
    default:
      break;
    }
  }
  return launch_one(fused_generic<T, K, G, STRIDED>, p, bytes, tiles, Geom::NT, stream);
}

template <typename T, int K, int G>
hipError_t launch_kg(SweepParams const &p, hipStream_t stream) {
  return p.B > 1 ? launch_kg_layout<T, K, G, true>(p, stream)
                 : launch_kg_layout<T, K, G, false>(p, stream);
}

// Batched launches exist for the ladder schedules only (the named
// hierarchize / dehierarchize transforms).
// `tiles` tiles of p.threads threads (the compiled default when 0; the plan
// built its tile table with the same width).
template <typename T, int K, int G, bool STRIDED, int NTO>
hipError_t launch_batched_nt(SweepParams const &p, ManyGrids grids, int64_t tiles,
                             hipStream_t stream) {
  using Geom = TileGeom<K, G, STRIDED, T, NTO>;
  size_t const bytes = sizeof(T) * Geom::lds_elems(p.m, STRIDED);
  bool const asc = p.schedule == kScheduleLadderAsc;
  switch (p.pattern) {
  case kPatOdd:
    return asc ? launch_one(fused_ladder_batched<T, K, G, STRIDED, kPatOdd, true, NTO>, p, bytes, tiles, Geom::NT, stream, grids)
               : launch_one(fused_ladder_batched<T, K, G, STRIDED, kPatOdd, false, NTO>, p, bytes, tiles, Geom::NT, stream, grids);
  case kPatOddEven:
    return asc ? launch_one(fused_ladder_batched<T, K, G, STRIDED, kPatOddEven, true, NTO>, p, bytes, tiles, Geom::NT, stream, grids)
               : launch_one(fused_ladder_batched<T, K, G, STRIDED, kPatOddEven, false, NTO>, p, bytes, tiles, Geom::NT, stream, grids);
  case kPatEvenOdd:
    return asc ? launch_one(fused_ladder_batched<T, K, G, STRIDED, kPatEvenOdd, true, NTO>, p, bytes, tiles, Geom::NT, stream, grids)
               : launch_one(fused_ladder_batched<T, K, G, STRIDED, kPatEvenOdd, false, NTO>, p, bytes, tiles, Geom::NT, stream, grids);
  default:
    return hipErrorInvalidValue;
  }
}

template <typename T, int K, int G, bool STRIDED>
hipError_t launch_batched_layout(SweepParams const &p, ManyGrids grids,
                                 int64_t tiles, hipStream_t stream) {
  if (p.schedule != kScheduleLadderAsc && p.schedule != kScheduleLadderDesc)
    return hipErrorInvalidValue;
  if constexpr (width_at_launch<T, K, G, STRIDED>()) {
    if constexpr (wide_fits<T, K, G>())
      if (p.threads == 1024)
        return launch_batched_nt<T, K, G, STRIDED, 1024>(p, grids, tiles, stream);
    if (p.threads == 256)
      return launch_batched_nt<T, K, G, STRIDED, 256>(p, grids, tiles, stream);
    if (p.threads == 512)
      return launch_batched_nt<T, K, G, STRIDED, 512>(p, grids, tiles, stream);
  }
  if (p.threads != 0 && p.threads != TileGeom<K, G, STRIDED, T>::NT)
    return hipErrorInvalidValue; // a tile table built for another width
  return launch_batched_nt<T, K, G, STRIDED, 0>(p, grids, tiles, stream);
}

template <typename T, int K, int G>
hipError_t launch_batched_kg(SweepParams const &p, bool strided,
                             ManyGrids grids, int64_t tiles,
                             hipStream_t stream) {
  return strided ? launch_batched_layout<T, K, G, true>(p, grids, tiles, stream)
                 : launch_batched_layout<T, K, G, false>(p, grids, tiles, stream);
}

// Compile-time dispatch on the line period 2^L (1 <= L <= LMAX).
template <typename T, int L, int LMAX>
hipError_t dispatch_period(SweepParams const &p, hipStream_t stream) {
  if constexpr (L > LMAX) {
    return hipErrorInvalidValue;
  } else {
    if (p.L == L) {
      constexpr int K = fused_chunk_of(sizeof(T), L);
      return launch_kg<T, K, (1 << L) / K>(p, stream);
    }
    return dispatch_period<T, L + 1, LMAX>(p, stream);
  }
}

template <typename T, int L, int LMAX>
hipError_t dispatch_period_batched(SweepParams const &p, bool strided,
                                   ManyGrids grids, int64_t tiles,
                                   hipStream_t stream) {
  if constexpr (L > LMAX) {
    return hipErrorInvalidValue;
  } else {
    if (p.L == L) {
      constexpr int K = fused_chunk_of(sizeof(T), L);
      return launch_batched_kg<T, K, (1 << L) / K>(p, strided, grids, tiles,
                                                   stream);
    }
    return dispatch_period_batched<T, L + 1, LMAX>(p, strided, grids, tiles,
                                                   stream);
  }
}

} // namespace fused
} // namespace paliwa_hip
