// Axis sweeps of long lines (periods beyond the fused register kernel):
// float64 lines of 2^12 .. 2^15 points, float32 lines of 2^13 .. 2^16, for the
// ladder schedules of the named transforms (hierarchize_in's and
// dehierarchize_in's consecutive level lists with the reference's table
// parities, paliwa/paliwa_transform.hpp:346-399).  Also the column tiles of
// strided float64 lines of 2^8 points (below).
//
// A workgroup of NT threads owns a tile of TB lines; each line of period n is
// split into G = n / K chunks of K consecutive points, one chunk per thread in
// registers (TB * G = NT).  Every point is read from HBM once and written once
// per sweep:
//   * levels whose storage stride s is below K run inside the chunk, in
//     registers; each pass needs one value of a neighbouring chunk of the same
//     line (the next chunk's first point for an odd pass, the previous chunk's
//     point K - s for an even pass), exchanged through LDS with one barrier
//     per pass (two slot arrays alternate, so a pass's writes never meet the
//     previous pass's reads);
//   * the coarser levels touch only the chunks' first points: those go to
//     LDS, the remaining levels run there with one barrier per pass, and they
//     come back to the registers.
// Along a strided axis the lanes run across the tile's lines (the TB columns
// of a row are consecutive in memory, so each row access is TB points wide).
// Along the contiguous axis a line is one run of memory: the tile's lines go
// through an LDS staging buffer, a few lines at a time, with coalesced copies
// (consecutive threads, consecutive points), and each thread takes its chunk
// from a layout with one pad point after every chunk (chunks K points apart
// would put a wavefront's reads in one bank); lines too long for the buffer
// read their chunks straight from memory.
//
// Column tiles.  Strided float64 lines of 2^8 points (a 257^3 grid's first two
// axes) run here too, with chunks of 16 points: a 1024-thread tile holds 64
// columns, so each row access is 512 bytes (the fused kernel's 512-thread
// tiles hold 32, and at odd pitches every 256-byte row spans three 128-byte
// lines, the two at its ends shared with the neighbouring tiles), and 60
// VGPRs with 24 KiB of LDS let two such workgroups share a CU, so one
// computes while the other loads (the fused kernel's LDS transpose buffer
// for a tile that wide leaves room for one).
//
// Each update is c0*g[lower] + c1*g[x] + c2*g[upper],
// left to right without contraction, exactly the reference's expression, and
// a pass is a pure function of the state before it (written and read points
// have opposite parity), so the results are bit-identical to the reference's.
#include "fused_sweep.hpp"
#include "lifting.hpp"

#include <algorithm>
#include <type_traits>
#include <utility>

namespace paliwa_hip {
namespace {

constexpr int kLongThreads = 1024; // the widest tile; 256-thread tiles for few lines
// LDS staging of contiguous tiles: whole padded lines (n + n / K points), as
// many per round as fit
constexpr int kStageBytes = 68 * 1024;

template <typename T> constexpr int long_chunk() { return sizeof(T) == 8 ? 32 : 64; }
// column tiles: strided float64 lines of 2^kColumnL points in chunks of 128
// bytes (16 points).  Float32 column tiles (32-point chunks, 128 columns per
// tile) were bit-identical but gained at most 1-3% on the 257^3 float32 grid
// (516 tiles, one per workgroup) and lost on smaller grids: not built.
constexpr int kColumnL = 8;
template <typename T> constexpr int column_chunk() { return 128 / int(sizeof(T)); }
// waves per SIMD a kernel is compiled for (amdgpu_waves_per_eu, through
// __launch_bounds__'s second argument): 8 for 1024-thread tiles of chunks of
// at most 128 bytes (the column tiles: two workgroups per CU; 60 VGPRs for
// the two-pass tables, 58 for the hat table's dehierarchize), else 1 (no
// bound).  The hat table's hierarchize (fine levels before coarse ones)
// takes 72 VGPRs and held to 64 spills, so it runs unbounded, one workgroup
// per CU.
template <typename T, int K, int NT, int PAT, bool ASC> constexpr int long_waves_per_eu() {
  return NT == 1024 && K * int(sizeof(T)) <= 128 && (PAT != kPatOdd || !ASC) ? 8 : 1;
}

// Padded lines of period 2^L that one staging round holds (0: the line does
// not fit, its chunks are read from memory).
template <typename T> __host__ __device__ constexpr int stage_lines_of(int L) {
  int64_t const padded = ((int64_t(1) << L) + ((int64_t(1) << L) / long_chunk<T>())) *
                         int64_t(sizeof(T));
  return static_cast<int>(kStageBytes / padded);
}

template <typename T> struct LongCoefs {
  T c[2][3];
};

template <typename T>
__device__ __forceinline__ T lift3(T c0, T lo, T c1, T x, T c2, T hi) {
  return c0 * lo + c1 * x + c2 * hi;
}

// One pass at compile-time stride S (S < K) inside a chunk; `nb` is the
// neighbouring chunk's value (next chunk's point 0 for odd passes, previous
// chunk's point K - S for even passes).
template <typename T, int K, int S, bool ODD>
__device__ __forceinline__ void chunk_pass(T (&v)[K], T nb, T const (&c)[3]) {
  if constexpr (ODD) {
#pragma unroll
    for (int j = S; j < K; j += 2 * S) {
      T const hi = j + S < K ? v[j + S] : nb;
      v[j] = lift3(c[0], v[j - S], c[1], v[j], c[2], hi);
    }
  } else {
#pragma unroll
    for (int j = 0; j < K; j += 2 * S) {
      T const lo = j == 0 ? nb : v[j - S];
      v[j] = lift3(c[0], lo, c[1], v[j], c[2], v[j + S]);
    }
  }
}

template <int PAT> constexpr int pat_passes() { return PAT == kPatOdd ? 1 : 2; }
template <int PAT> constexpr bool pat_odd(int f) {
  return PAT == kPatOdd ? true : (PAT == kPatOddEven ? f == 0 : f == 1);
}

// Geometry of the tile held by thread `tid`: line li of the tile, chunk gi.
struct LongGeom {
  int G, TB, li, gi;
  int next, prev; // threads holding chunks gi + 1 and gi - 1 of the same line
  __device__ LongGeom(int tid, int G_, int TB_, bool strided) : G(G_), TB(TB_) {
    if (strided) {
      li = tid % TB;
      gi = tid / TB;
    } else {
      gi = tid % G;
      li = tid / G;
    }
    int const gn = gi + 1 == G ? 0 : gi + 1;
    int const gp = gi == 0 ? G - 1 : gi - 1;
    next = strided ? li + TB * gn : gn + G * li;
    prev = strided ? li + TB * gp : gp + G * li;
  }
};

template <typename T, int K, int NT, bool STRIDED, int PAT, bool ASC, typename Src>
__device__ __forceinline__ void long_tiles(SweepParams const &p, Src const &src,
                                           int64_t ntiles) {
  __shared__ T xch[2][NT];
  __shared__ T heads[NT];
  // staging buffer of contiguous tiles (one element in strided kernels)
  __shared__ T stage[STRIDED ? 1 : kStageBytes / sizeof(T)];
  constexpr int LOGK = fused::ilog2(K);
  int const tid = threadIdx.x;
  int const n = 1 << p.L;
  int const m = p.m;
  int const G = n / K;
  int const TB = NT / G;
  LongGeom const geo(tid, G, TB, STRIDED);
  LongCoefs<T> cf;
#pragma unroll
  for (int f = 0; f < pat_passes<PAT>(); ++f)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      cf.c[f][k] = static_cast<T>(p.coef[f][k]);
  int buf = 0; // exchange slot array of the next fine pass

  for (int64_t t = fused::xcd_tile(blockIdx.x, gridDim.x); t < ntiles; t += gridDim.x) {
    TileRef const ref = src(t);
    int64_t const line = ref.local * TB + geo.li;
    bool const active = line < ref.A * ref.B;
    T *base = nullptr;
    int64_t step = 1; // distance between consecutive points of the line
    if (active) {
      int64_t const a = line / ref.B;
      int64_t const b = line - a * ref.B;
      base = static_cast<T *>(ref.data) + a * int64_t(m) * ref.B + b;
      step = ref.B;
    }
    T v[K];
    // contiguous tiles through the staging buffer: rounds of R lines, line r
    // of a round at stage[r * P], chunk g of it at g * (K + 1)
    [[maybe_unused]] int const R = STRIDED ? 0 : min(TB, stage_lines_of<T>(p.L));
    [[maybe_unused]] int const P = n + G;
    [[maybe_unused]] int64_t const first = ref.local * TB; // the tile's first line
    [[maybe_unused]] int64_t const nlines = ref.A * ref.B;
    [[maybe_unused]] T *const lines0 = static_cast<T *>(ref.data) + first * int64_t(m);
    if constexpr (!STRIDED) {
      if (R > 0) {
#pragma unroll
        for (int j = 0; j < K; ++j) // lines past the grid's end stay 0
          v[j] = T(0);
        for (int r0 = 0; r0 < TB; r0 += R) {
          for (int r = 0; r < R && r0 + r < TB; ++r) {
            if (first + r0 + r >= nlines)
              break;
            T const *src = lines0 + int64_t(r0 + r) * m;
            for (int x = tid; x < n; x += NT)
              stage[r * P + x + (x >> LOGK)] = src[x];
          }
          __syncthreads();
          if (geo.li >= r0 && geo.li < r0 + R) {
            T const *c = stage + (geo.li - r0) * P + geo.gi * (K + 1);
#pragma unroll
            for (int j = 0; j < K; ++j)
              v[j] = c[j];
          }
          __syncthreads(); // the next round overwrites the buffer
        }
      }
    }
    if (STRIDED || R == 0) {
      if (active) {
        T const *q = base + int64_t(geo.gi) * K * step;
#pragma unroll
        for (int j = 0; j < K; ++j)
          v[j] = q[j * step];
      } else {
#pragma unroll
        for (int j = 0; j < K; ++j)
          v[j] = T(0);
      }
    }

    // fine levels (stride 2^e < K) in registers
    auto fine_level = [&](auto e_tag) {
      constexpr int E = decltype(e_tag)::value;
      constexpr int S = 1 << E;
#pragma unroll
      for (int f = 0; f < pat_passes<PAT>(); ++f) {
        bool const odd = pat_odd<PAT>(f);
        xch[buf][tid] = odd ? v[0] : v[K - S];
        __syncthreads();
        T const nb = xch[buf][odd ? geo.next : geo.prev];
        buf ^= 1;
        if (odd)
          chunk_pass<T, K, S, true>(v, nb, cf.c[f]);
        else
          chunk_pass<T, K, S, false>(v, nb, cf.c[f]);
      }
    };
    // the fine levels in the order of the tags (uniform conditions: every
    // thread of the workgroup takes the same branches to the barriers)
    auto fine_levels = [&](auto... e_tags) {
      if constexpr (pat_passes<PAT>() == 1 && pat_odd<PAT>(0)) {
        // odd passes only: every fine level needs the next chunk's point 0,
        // which no odd pass at a stride below K changes, so one exchange
        // serves them all
        bool const any = ((decltype(e_tags)::value >= p.lo &&
                           decltype(e_tags)::value <= p.hi) || ...);
        if (any) {
          xch[buf][tid] = v[0];
          __syncthreads();
          T const nb = xch[buf][geo.next];
          buf ^= 1;
          ((decltype(e_tags)::value >= p.lo && decltype(e_tags)::value <= p.hi
                ? chunk_pass<T, K, (1 << decltype(e_tags)::value), true>(v, nb, cf.c[0])
                : void()),
           ...);
        }
      } else {
        ((decltype(e_tags)::value >= p.lo && decltype(e_tags)::value <= p.hi
              ? fine_level(e_tags)
              : void()),
         ...);
      }
    };
    // coarse levels (stride 2^e >= K) on the chunks' first points, in LDS
    auto coarse_levels = [&]() {
      int const e0 = p.lo > LOGK ? p.lo : LOGK;
      if (p.hi < e0)
        return;
      heads[tid] = v[0];
      __syncthreads();
      for (int i = 0; i <= p.hi - e0; ++i) {
        int const e = ASC ? e0 + i : p.hi - i;
        int const S = 1 << (e - LOGK); // in chunks
#pragma unroll
        for (int f = 0; f < pat_passes<PAT>(); ++f) {
          int const off = pat_odd<PAT>(f) ? S : 0;
          if ((geo.gi & (2 * S - 1)) == off) {
            int const gl = geo.gi - S < 0 ? geo.gi - S + G : geo.gi - S;
            int const gh = geo.gi + S >= G ? geo.gi + S - G : geo.gi + S;
            int const tl = STRIDED ? geo.li + TB * gl : gl + G * geo.li;
            int const th = STRIDED ? geo.li + TB * gh : gh + G * geo.li;
            heads[tid] = lift3(cf.c[f][0], heads[tl], cf.c[f][1], heads[tid], cf.c[f][2],
                               heads[th]);
          }
          __syncthreads();
        }
      }
      v[0] = heads[tid];
      __syncthreads(); // heads may be rewritten by the next tile
    };
    if constexpr (ASC) {
      [&]<int... E>(std::integer_sequence<int, E...>) {
        fine_levels(std::integral_constant<int, E>{}...);
      }(std::make_integer_sequence<int, LOGK>{});
      coarse_levels();
    } else {
      coarse_levels();
      [&]<int... E>(std::integer_sequence<int, E...>) {
        fine_levels(std::integral_constant<int, LOGK - 1 - E>{}...);
      }(std::make_integer_sequence<int, LOGK>{});
    }

    if constexpr (!STRIDED) {
      if (R > 0) {
        for (int r0 = 0; r0 < TB; r0 += R) {
          if (geo.li >= r0 && geo.li < r0 + R) {
            T *c = stage + (geo.li - r0) * P + geo.gi * (K + 1);
#pragma unroll
            for (int j = 0; j < K; ++j)
              c[j] = v[j];
          }
          __syncthreads();
          for (int r = 0; r < R && r0 + r < TB; ++r) {
            if (first + r0 + r >= nlines)
              break;
            T *dst = lines0 + int64_t(r0 + r) * m;
            for (int x = tid; x < n; x += NT)
              dst[x] = stage[r * P + x + (x >> LOGK)];
            if (m != n && tid == 0) // periodic image: a copy of index 0
              dst[n] = stage[r * P];
          }
          __syncthreads(); // the next round (or tile) overwrites the buffer
        }
        continue;
      }
    }
    if (active) {
      T *q = base + int64_t(geo.gi) * K * step;
#pragma unroll
      for (int j = 0; j < K; ++j)
        q[j * step] = v[j];
      if (m != n && geo.gi == 0)
        base[int64_t(n) * step] = v[0];
    }
  }
}

template <typename T, int K, int NT, bool STRIDED, int PAT, bool ASC>
__global__ __launch_bounds__(NT, (long_waves_per_eu<T, K, NT, PAT, ASC>())) void long_ladder(SweepParams p,
                                                                             int64_t ntiles) {
  long_tiles<T, K, NT, STRIDED, PAT, ASC>(p, OneGrid{p.data, p.A, p.B}, ntiles);
}

template <typename T, int K, int NT, bool STRIDED, int PAT, bool ASC>
__global__ __launch_bounds__(NT, (long_waves_per_eu<T, K, NT, PAT, ASC>())) void long_ladder_batched(
    SweepParams p, int64_t ntiles, ManyGrids grids) {
  long_tiles<T, K, NT, STRIDED, PAT, ASC>(p, grids, ntiles);
}

template <int NT, typename Kernel, typename... Extra>
hipError_t launch_long_kernel(Kernel kernel, SweepParams const &p, int64_t tiles,
                              hipStream_t stream, Extra... extra) {
  int64_t const resident =
      fused::resident_workgroups(reinterpret_cast<const void *>(kernel), 0, NT);
  uint32_t const grid = static_cast<uint32_t>(tiles < resident ? tiles : resident);
  if (grid == 0)
    return hipSuccess;
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(NT), 0, stream, p, tiles, extra...);
  return hipGetLastError();
}

template <typename T, int K, bool STRIDED, bool BATCHED, int NT, typename... Extra>
hipError_t dispatch_long_nt(SweepParams const &p, int64_t tiles, hipStream_t stream,
                            Extra... extra) {
  bool const asc = p.schedule == kScheduleLadderAsc;
  auto go = [&](auto pat, auto asc_tag) -> hipError_t {
    constexpr int PAT = decltype(pat)::value;
    constexpr bool ASC = decltype(asc_tag)::value;
    if constexpr (BATCHED)
      return launch_long_kernel<NT>(long_ladder_batched<T, K, NT, STRIDED, PAT, ASC>, p,
                                    tiles, stream, extra...);
    else
      return launch_long_kernel<NT>(long_ladder<T, K, NT, STRIDED, PAT, ASC>, p, tiles,
                                    stream);
  };
  using F = std::false_type;
  using Tr = std::true_type;
  switch (p.pattern) {
  case kPatOdd:
    return asc ? go(std::integral_constant<int, kPatOdd>{}, Tr{})
               : go(std::integral_constant<int, kPatOdd>{}, F{});
  case kPatOddEven:
    return asc ? go(std::integral_constant<int, kPatOddEven>{}, Tr{})
               : go(std::integral_constant<int, kPatOddEven>{}, F{});
  case kPatEvenOdd:
    return asc ? go(std::integral_constant<int, kPatEvenOdd>{}, Tr{})
               : go(std::integral_constant<int, kPatEvenOdd>{}, F{});
  default:
    return hipErrorInvalidValue;
  }
}

// A line of G chunks fits a 256-thread tile when G <= 256.
template <typename T, int K, bool STRIDED, bool BATCHED, typename... Extra>
hipError_t dispatch_long(SweepParams const &p, int threads, int64_t tiles, hipStream_t stream,
                         Extra... extra) {
  if (threads == 256)
    return dispatch_long_nt<T, K, STRIDED, BATCHED, 256>(p, tiles, stream, extra...);
  return dispatch_long_nt<T, K, STRIDED, BATCHED, kLongThreads>(p, tiles, stream, extra...);
}

// Tile width for `lines` lines of G chunks: 1024 threads, or 256 where
// 1024-thread tiles would be fewer than the 256 CUs and a line fits 256
// threads.
int long_threads_of(int64_t G, int64_t lines) {
  if (G > 256)
    return kLongThreads;
  int64_t const wide_tiles = (lines * G + kLongThreads - 1) / kLongThreads;
  return wide_tiles < 256 ? 256 : kLongThreads;
}

bool long_columns_fit(int dtype, SweepParams const &p) {
  return dtype == PALIWA_FLOAT64 && p.B > 1 && p.L == kColumnL;
}

template <typename T>
hipError_t launch_long_typed(SweepParams const &p, hipStream_t stream) {
  int const dtype = sizeof(T) == 8 ? PALIWA_FLOAT64 : PALIWA_FLOAT32;
  int64_t const lines = p.A * p.B;
  if constexpr (sizeof(T) == 8) {
    if (long_columns_fit(dtype, p)) { // column tiles of 128-byte chunks
      constexpr int K = column_chunk<T>();
      int64_t const G = (int64_t(1) << p.L) / K;
      int const threads = long_threads_of(G, lines);
      int64_t const tb = threads / G;
      return dispatch_long<T, K, true, false>(p, threads, (lines + tb - 1) / tb, stream);
    }
  }
  int const threads = long_threads_for(dtype, p.L, lines);
  int64_t const tb = long_tile_lines(dtype, p.L, threads);
  int64_t const tiles = (lines + tb - 1) / tb;
  constexpr int K = long_chunk<T>();
  return p.B > 1 ? dispatch_long<T, K, true, false>(p, threads, tiles, stream)
                 : dispatch_long<T, K, false, false>(p, threads, tiles, stream);
}

} // namespace

bool long_ladder_fits(int dtype, SweepParams const &p) {
  int64_t const n = int64_t(1) << p.L;
  bool const ladder = (p.schedule == kScheduleLadderAsc || p.schedule == kScheduleLadderDesc) &&
                      p.pattern != kPatNone;
  return ladder && ((n > fused_max_period(dtype) && n <= long_ladder_max_period(dtype)) ||
                    long_columns_fit(dtype, p));
}

bool long_columns_preferred(int dtype, SweepParams const &p) {
  if (!long_ladder_fits(dtype, p) || !long_columns_fit(dtype, p))
    return false;
  int64_t const lines = p.A * p.B;
  int64_t const bytes = lines * p.m * int64_t(sizeof(double));
  if (lines < 16384)
    return false;
  if (p.pattern == kPatOdd) // the hat table: only an outer axis streaming HBM
    return p.B >= 4096 && bytes > (int64_t(256) << 20);
  return p.B >= 4096 || bytes <= (int64_t(160) << 20);
}

bool long_ladder_preferred(int dtype, SweepParams const &p) {
  return long_ladder_fits(dtype, p) &&
         (p.B > 1 || long_contiguous_staged(dtype, p.L) ||
          (int64_t(1) << p.L) > lds_line_max_period(dtype));
}

bool long_contiguous_staged(int dtype, int L) {
  return (dtype == PALIWA_FLOAT64 ? stage_lines_of<double>(L) : stage_lines_of<float>(L)) > 0;
}

int long_ladder_max_period(int dtype) {
  return dtype == PALIWA_FLOAT64 ? 32 * kLongThreads : 64 * kLongThreads;
}

int long_threads_for(int dtype, int L, int64_t lines) {
  return long_threads_of((int64_t(1) << L) / (dtype == PALIWA_FLOAT64 ? 32 : 64), lines);
}

int64_t long_tile_lines(int dtype, int L, int threads) {
  int const K = dtype == PALIWA_FLOAT64 ? 32 : 64;
  return threads / ((int64_t(1) << L) / K);
}

hipError_t launch_long_ladder(int dtype, SweepParams const &p, hipStream_t stream) {
  return dtype == PALIWA_FLOAT64 ? launch_long_typed<double>(p, stream)
                                 : launch_long_typed<float>(p, stream);
}

hipError_t launch_long_ladder_batched(int dtype, SweepParams const &p, bool strided,
                                      ManyGrids grids, int64_t tiles, hipStream_t stream) {
  int const threads = p.threads == 256 ? 256 : kLongThreads;
  constexpr int K64 = long_chunk<double>(), K32 = long_chunk<float>();
  if (dtype == PALIWA_FLOAT64)
    return strided ? dispatch_long<double, K64, true, true>(p, threads, tiles, stream, grids)
                   : dispatch_long<double, K64, false, true>(p, threads, tiles, stream, grids);
  return strided ? dispatch_long<float, K32, true, true>(p, threads, tiles, stream, grids)
                 : dispatch_long<float, K32, false, true>(p, threads, tiles, stream, grids);
}

} // namespace paliwa_hip
