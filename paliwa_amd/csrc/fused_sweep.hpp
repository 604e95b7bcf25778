// The fused axis-sweep kernel (device code, included by fused_f64.hip and
// fused_f32.hip, one translation unit per dtype so they compile in parallel).
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
// strided axis are copied row by row, TB consecutive columns per row.  All the
// loads of a thread are issued before the first LDS write so that K+1 loads
// per lane are in flight.  Along a strided axis with one lane per line (G == 1,
// short lines) each lane loads its own column straight into registers: the
// TB lanes of a row read TB consecutive elements, which is already coalesced.
//
// Level schedules.  fused_ladder covers level lists whose storage strides are
// consecutive powers of two in increasing (hierarchize) or decreasing
// (dehierarchize) order with the pass-parity patterns of the reference's
// tables ({odd}, {odd, even}, {even, odd}); every register index is then a
// compile-time constant.  fused_generic runs any level list and any filter
// list by dispatching on the stride at run time.
#pragma once

#include "lifting.hpp"

namespace paliwa_hip {
namespace fused {

constexpr int kThreads = 256;
constexpr int kXcds = 8;

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

// One pass at storage stride S < K on this lane's chunk.
template <typename T, int K, int G, int S, int ODD>
__device__ __forceinline__ void pass_in_lane(T (&v)[K], int gbase, int gi,
                                             T c0, T c1, T c2) {
  if constexpr (ODD) {
    T nb;
    if constexpr (G == 1)
      nb = v[0];
    else
      nb = __shfl(v[0], gbase + ((gi + 1) & (G - 1)), 64);
#pragma unroll
    for (int j = S; j < K; j += 2 * S) {
      T const hi = (j + S < K) ? v[j + S] : nb;
      v[j] = lift(c0, v[j - S], c1, v[j], c2, hi);
    }
  } else {
    T pb;
    if constexpr (G == 1)
      pb = v[K - S];
    else
      pb = __shfl(v[K - S], gbase + ((gi - 1) & (G - 1)), 64);
#pragma unroll
    for (int j = 0; j < K; j += 2 * S) {
      T const lo = (j >= S) ? v[j - S] : pb;
      v[j] = lift(c0, lo, c1, v[j], c2, v[j + S]);
    }
  }
}

// One pass at a stride that is q >= 1 chunks long.
template <typename T, int K, int G>
__device__ __forceinline__ void pass_cross_lane(T (&v)[K], int gbase, int gi,
                                                int q, int odd, T c0, T c1,
                                                T c2) {
  T const lo = __shfl(v[0], gbase + ((gi - q) & (G - 1)), 64);
  T const hi = __shfl(v[0], gbase + ((gi + q) & (G - 1)), 64);
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

// LDS geometry of a tile of TB lines.
//   contiguous axis: line-major, point r of line l at l*LP + r + (r >> LOGK)
//   strided axis:    row-major,  point r of column c at r*TB + c + (r >> LOGK)
// One pad word per chunk puts the K-strided register loads of the G lanes of
// a line on distinct banks.
template <int K, int G> struct TileGeom {
  static constexpr int TB = kThreads / G;
  static constexpr int LOGK = ilog2(K);
  static constexpr bool kDirect = false;
  __host__ __device__ static int line_pitch(int m) { return (m + G) | 1; }
  __host__ __device__ static int lds_elems(int m, bool strided) {
    if (strided && G == 1)
      return 0; // columns go straight to registers
    return strided ? m * TB + G + 2 : TB * line_pitch(m);
  }
};

// Kernel body; MODE 0 = ladder (PAT, ASC used), 1 = generic.
template <typename T, int K, int G, bool STRIDED, int MODE, int PAT, bool ASC>
__device__ __forceinline__ void sweep_tile(SweepParams const &p,
                                           uint32_t magic_m, T *lds) {
  using Geom = TileGeom<K, G>;
  constexpr int TB = Geom::TB;
  constexpr int LOGK = Geom::LOGK;
  constexpr int ITER = K + 1; // staged points per thread (image included)

  int const tid = threadIdx.x;
  int64_t const tile = xcd_tile(blockIdx.x, gridDim.x);
  int const m = p.m;
  int const n = 1 << p.L;
  T *const g = static_cast<T *>(p.data);
  int const li = tid / G;
  int const gi = tid % G;
  int const gbase = (tid & 63) - gi;

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

  T v[K];
  if constexpr (STRIDED && G == 1) {
    // One lane per line: each lane's column straight from HBM.
    int64_t const col = tile * TB + tid;
    if (col >= p.A * p.B)
      return;
    int64_t const a = col / p.B;
    int64_t const b = col - a * p.B;
    T *const colp = g + a * m * p.B + b;
#pragma unroll
    for (int j = 0; j < K; ++j)
      v[j] = colp[j * p.B];
    compute(v);
#pragma unroll
    for (int j = 0; j < K; ++j)
      colp[j * p.B] = v[j];
    if (m != n)
      colp[static_cast<int64_t>(n) * p.B] = v[0];
    return;
  } else {
    int const LP = Geom::line_pitch(m);
    auto lds_index = [&](int line, int r) {
      if constexpr (STRIDED)
        return r * TB + line + (r >> LOGK);
      else
        return line * LP + r + (r >> LOGK);
    };

    T buf[ITER];
    T *src = nullptr;  // contiguous: first point of the tile
    T *colp = nullptr; // strided: this thread's column
    int count = 0;     // contiguous: points in the tile
    int const col = tid % TB;
    int const row0 = tid / TB;
    if constexpr (!STRIDED) {
      int64_t const line0 = tile * TB;
      int64_t const left = p.A - line0;
      int const nl = static_cast<int>(left < TB ? left : TB);
      src = g + line0 * m;
      count = nl * m;
#pragma unroll
      for (int i = 0; i < ITER; ++i) {
        int const e = tid + i * kThreads;
        if (e < count)
          buf[i] = src[e];
      }
#pragma unroll
      for (int i = 0; i < ITER; ++i) {
        int const e = tid + i * kThreads;
        if (e < count) {
          int const l = __umulhi(static_cast<uint32_t>(e), magic_m);
          int const r = e - l * m;
          lds[lds_index(l, r)] = buf[i];
        }
      }
    } else {
      int64_t const c0 = tile * TB + col;
      if (c0 < p.A * p.B) {
        int64_t const a = c0 / p.B;
        int64_t const b = c0 - a * p.B;
        colp = g + a * m * p.B + b;
#pragma unroll
        for (int i = 0; i < ITER; ++i) {
          int const r = row0 + i * G;
          if (r < m)
            buf[i] = colp[r * p.B];
        }
#pragma unroll
        for (int i = 0; i < ITER; ++i) {
          int const r = row0 + i * G;
          if (r < m)
            lds[lds_index(col, r)] = buf[i];
        }
      }
    }
    __syncthreads();

#pragma unroll
    for (int j = 0; j < K; ++j)
      v[j] = lds[lds_index(li, gi * K + j)];
    compute(v);
#pragma unroll
    for (int j = 0; j < K; ++j)
      lds[lds_index(li, gi * K + j)] = v[j];
    if (m != n && gi == 0)
      lds[lds_index(li, n)] = v[0];
    __syncthreads();

    if constexpr (!STRIDED) {
#pragma unroll
      for (int i = 0; i < ITER; ++i) {
        int const e = tid + i * kThreads;
        if (e < count) {
          int const l = __umulhi(static_cast<uint32_t>(e), magic_m);
          int const r = e - l * m;
          buf[i] = lds[lds_index(l, r)];
        }
      }
#pragma unroll
      for (int i = 0; i < ITER; ++i) {
        int const e = tid + i * kThreads;
        if (e < count)
          src[e] = buf[i];
      }
    } else {
      if (colp != nullptr) {
#pragma unroll
        for (int i = 0; i < ITER; ++i) {
          int const r = row0 + i * G;
          if (r < m)
            buf[i] = lds[lds_index(col, r)];
        }
#pragma unroll
        for (int i = 0; i < ITER; ++i) {
          int const r = row0 + i * G;
          if (r < m)
            colp[r * p.B] = buf[i];
        }
      }
    }
  }
}

template <typename T, int K, int G, bool STRIDED, int PAT, bool ASC>
__global__ __launch_bounds__(kThreads) void fused_ladder(SweepParams p,
                                                         uint32_t magic_m) {
  extern __shared__ unsigned char smem_raw[];
  sweep_tile<T, K, G, STRIDED, 0, PAT, ASC>(p, magic_m,
                                            reinterpret_cast<T *>(smem_raw));
}

template <typename T, int K, int G, bool STRIDED>
__global__ __launch_bounds__(kThreads) void fused_generic(SweepParams p,
                                                          uint32_t magic_m) {
  extern __shared__ unsigned char smem_raw[];
  sweep_tile<T, K, G, STRIDED, 1, 0, true>(p, magic_m,
                                           reinterpret_cast<T *>(smem_raw));
}

template <typename Kernel>
hipError_t allow_lds(Kernel kernel, size_t bytes) {
  if (bytes <= 65536)
    return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(bytes));
}

template <typename Kernel>
hipError_t launch_one(Kernel kernel, SweepParams const &p, size_t bytes,
                      uint32_t tiles, hipStream_t stream) {
  hipError_t err = allow_lds(kernel, bytes);
  if (err != hipSuccess)
    return err;
  // e / m by multiply-high: exact while (points per tile) * m < 2^32.
  uint32_t const magic = static_cast<uint32_t>(0xffffffffu / p.m + 1);
  hipLaunchKernelGGL(kernel, dim3(tiles), dim3(kThreads), bytes, stream, p,
                     magic);
  return hipGetLastError();
}

template <typename T, int K, int G, bool STRIDED>
hipError_t launch_kg_layout(SweepParams const &p, hipStream_t stream) {
  using Geom = TileGeom<K, G>;
  int64_t const lines = p.A * p.B;
  int64_t const tiles = (lines + Geom::TB - 1) / Geom::TB;
  if (tiles > 0x7fffffffLL)
    return hipErrorInvalidValue;
  size_t const bytes = sizeof(T) * Geom::lds_elems(p.m, STRIDED);
  uint32_t const nt = static_cast<uint32_t>(tiles);
  if (p.schedule == kScheduleLadderAsc || p.schedule == kScheduleLadderDesc) {
    bool const asc = p.schedule == kScheduleLadderAsc;
    switch (p.pattern) {
    case kPatOdd:
      return asc ? launch_one(fused_ladder<T, K, G, STRIDED, kPatOdd, true>, p, bytes, nt, stream)
                 : launch_one(fused_ladder<T, K, G, STRIDED, kPatOdd, false>, p, bytes, nt, stream);
    case kPatOddEven:
      return asc ? launch_one(fused_ladder<T, K, G, STRIDED, kPatOddEven, true>, p, bytes, nt, stream)
                 : launch_one(fused_ladder<T, K, G, STRIDED, kPatOddEven, false>, p, bytes, nt, stream);
    case kPatEvenOdd:
      return asc ? launch_one(fused_ladder<T, K, G, STRIDED, kPatEvenOdd, true>, p, bytes, nt, stream)
                 : launch_one(fused_ladder<T, K, G, STRIDED, kPatEvenOdd, false>, p, bytes, nt, stream);
    default:
      break;
    }
  }
  return launch_one(fused_generic<T, K, G, STRIDED>, p, bytes, nt, stream);
}

template <typename T, int K, int G>
hipError_t launch_kg(SweepParams const &p, hipStream_t stream) {
  return p.B > 1 ? launch_kg_layout<T, K, G, true>(p, stream)
                 : launch_kg_layout<T, K, G, false>(p, stream);
}

} // namespace fused
} // namespace paliwa_hip
