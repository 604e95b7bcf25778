// Boxes: several consecutive axes of a small grid swept together in LDS, for
// the batched plans of combination-technique sets (many grids of a few
// thousand to a few hundred thousand points).  Sweeping axis by axis reads and
// writes every grid once per axis; a box holds the full extents of axes a..b
// for one index of the axes before a and one tile of the flattened axes after
// b, so the axes of the box cost one HBM read and one write together.  A
// 4-D grid of 2^13..2^16 points usually splits into two boxes (axes 0..1 in
// tiles, axes 2..3 whole), halving its traffic.
//
// Inside the box every axis runs, in axis order, every level of its named
// schedule (hierarchize_in's decreasing levels or dehierarchize_in's
// increasing ones, paliwa/paliwa_transform.hpp:346-399) with every pass of the
// filter table (paliwa/paliwa_wavelets.hpp:15-28), then rewrites its periodic
// image plane from index 0, as the per-axis sweep does.  Each update is
// c0*g[lower] + c1*g[x] + c2*g[upper], left to right without contraction,
// with the reference's periodic neighbours (paliwa/paliwa_transform.hpp:38-49);
// the points a pass writes and the points it reads have opposite parity, and
// a barrier separates passes, so the result is bit-identical to the axis-by-
// axis sweep and to the reference.
//
// Layout of a box in LDS: row-major over the box axes, then the tile's inner
// points (a row of `width` points: all inner points of the grid, when they
// fit, or a power-of-two tile of them, the last tile partly empty).  When the
// tile holds all inner points the box is one contiguous run of the grid, and
// a workgroup takes `batch` boxes of consecutive outer indices, one
// contiguous run again, so that small boxes still fill the workgroup.
#include "fused_sweep.hpp"
#include "lifting.hpp"

#include <algorithm>
#include <utility>

namespace paliwa_hip {
namespace {

template <typename T>
__device__ __forceinline__ T lift3(T c0, T lo, T c1, T x, T c2, T hi) {
  return c0 * lo + c1 * x + c2 * hi;
}

constexpr int kXcds = 8; // accelerator complex dies; workgroups go round robin

// One box of a launch resolved from its grid's descriptor: its points are
// base[src(f)] for f in [0, size), src(f) = f in a dense box (one contiguous
// run), (f / wl) * inner + f % wl in a tiled one (rows of wl, a power of two,
// of which the first `valid` are stored).  A box spans fewer than 2^28 points
// of its grid (the plan's limit), so offsets are 32-bit.
template <typename T> struct BoxItem {
  T *base;
  int inner;
  int size;     // points in LDS
  int box;      // points of one box (size / boxes of the workgroup)
  int valid;
  int wl_shift; // log2 wl (tiled boxes)
  int dense;    // 1: one contiguous run
};

template <typename T>
__device__ __forceinline__ BoxItem<T> resolve(BoxGrid const &d, int64_t blk) {
  int64_t const local = blk - d.block_base;
  int64_t const ob = local / d.tiles;
  int64_t const col0 = (local - ob * d.tiles) * d.width;
  int64_t const o0 = ob * d.batch; // first outer index of the item
  BoxItem<T> it;
  it.dense = int64_t(d.width) == d.inner; // then tiles == 1
  it.inner = int(d.inner);
  it.valid = int(min(int64_t(d.width), d.inner - col0));
  it.box = d.rows * d.width;
  it.size = int(min(int64_t(d.batch), d.outer - o0)) * it.box;
  it.wl_shift = it.dense ? 0 : __ffs(d.width) - 1;
  it.base = static_cast<T *>(d.data) + o0 * d.outer_stride + col0;
  return it;
}

// HBM -> registers: point tid + u * NT of the item in v[u], all loads issued
// back to back without branches (they stay in flight while the previous item
// is swept): a lane past the item's end, or on a column past a partial
// tile's stored ones, reads a stored point of the item instead and its value
// is not deposited.
template <typename T> using GlobalPtr = T __attribute__((address_space(1))) *;

template <typename T, int NT, int R>
__device__ __forceinline__ void fetch(BoxItem<T> const &it, T (&v)[R]) {
  int const tid = threadIdx.x;
  GlobalPtr<T> const base = (GlobalPtr<T>)(it.base);
  int const last = it.size - 1;
  if (it.dense) {
#pragma unroll
    for (int u = 0; u < R; ++u)
      v[u] = base[min(tid + u * NT, last)];
  } else {
    int const mask = (1 << it.wl_shift) - 1;
#pragma unroll
    for (int u = 0; u < R; ++u) {
      int const f = min(tid + u * NT, last);
      v[u] = base[(f >> it.wl_shift) * it.inner + min(f & mask, it.valid - 1)];
    }
  }
}

template <typename T, int NT, int R>
__device__ __forceinline__ void deposit(BoxItem<T> const &it, T const (&v)[R], T *g) {
  int const tid = threadIdx.x;
  int const mask = it.dense ? -1 : (1 << it.wl_shift) - 1;
  int const valid = it.dense ? 0x7fffffff : it.valid;
#pragma unroll
  for (int u = 0; u < R; ++u) {
    int const f = tid + u * NT;
    if (f < it.size && (f & mask) < valid)
      g[f] = v[u];
  }
}

template <typename T, int NT, int R>
__device__ __forceinline__ void write_back(BoxItem<T> const &it, T const *g) {
  int const tid = threadIdx.x;
  GlobalPtr<T> const base = (GlobalPtr<T>)(it.base);
  if (it.dense) {
#pragma unroll 4
    for (int u = 0; u < R; ++u)
      if (tid + u * NT < it.size)
        base[tid + u * NT] = g[tid + u * NT];
  } else {
    int const mask = (1 << it.wl_shift) - 1;
#pragma unroll 4
    for (int u = 0; u < R; ++u) {
      int const f = tid + u * NT;
      if (f < it.size && (f & mask) < it.valid)
        base[(f >> it.wl_shift) * it.inner + (f & mask)] = g[f];
    }
  }
}

// ---- lines in registers ---------------------------------------------------
// A line of period N = 2^L (L <= RL below) held by one thread: every
// level and pass runs on its registers at compile-time positions, the line
// costs one LDS read and one write per point, and the axis needs no barrier
// between its levels.  The direction and the passes' parities are template
// parameters and a level outside the axis's schedule is a uniform select, so
// the line stays in registers (no branch merges the array).

// lines of up to 2^RL (+1 image) points; float64 boxes of four workgroups
// per CU hold up to 2^4 (128 registers a thread)
template <typename T, int CAP_BYTES>
constexpr int register_levels_of() {
  return CAP_BYTES == kBoxCapacities[0] && sizeof(T) == 8 ? 4 : 5;
}

// One pass at storage stride S over the odd multiples of S (ODD) or the even
// ones, with the periodic neighbours of the N-lattice; `active` false leaves
// the line as it is.  Written and read points have opposite parity.
template <typename T, int N, int S, bool ODD, int... J>
__device__ __forceinline__ void register_pass(T (&v)[N], bool active, T c0, T c1, T c2,
                                              std::integer_sequence<int, J...>) {
  if constexpr (ODD)
    ((v[S + 2 * S * J] = active ? lift3(c0, v[S + 2 * S * J - S], c1, v[S + 2 * S * J], c2,
                                        v[(S + 2 * S * J + S) % N])
                                : v[S + 2 * S * J]),
     ...);
  else
    ((v[2 * S * J] = active ? lift3(c0, v[(2 * S * J + N - S) % N], c1, v[2 * S * J], c2,
                                    v[2 * S * J + S])
                            : v[2 * S * J]),
     ...);
}

template <typename T, int N, int S, int PAT>
__device__ __forceinline__ void register_level(T (&v)[N], bool active, BoxLaunch const &p) {
  constexpr auto updates = std::make_integer_sequence<int, N / (2 * S)>{};
  register_pass<T, N, S, PAT != kPatEvenOdd>(v, active, static_cast<T>(p.coef[0][0]),
                                             static_cast<T>(p.coef[0][1]),
                                             static_cast<T>(p.coef[0][2]), updates);
  if constexpr (PAT != kPatOdd)
    register_pass<T, N, S, PAT == kPatEvenOdd>(v, active, static_cast<T>(p.coef[1][0]),
                                               static_cast<T>(p.coef[1][1]),
                                               static_cast<T>(p.coef[1][2]), updates);
}

// Levels of storage strides 2^0 .. 2^(nl-1): increasing (forward) or
// decreasing (inverse).
template <typename T, int L, bool FWD, int PAT, int... K>
__device__ __forceinline__ void register_levels(T (&v)[1 << L], int nl, BoxLaunch const &p,
                                                std::integer_sequence<int, K...>) {
  if constexpr (FWD)
    (register_level<T, (1 << L), (1 << K), PAT>(v, K < nl, p), ...);
  else
    (register_level<T, (1 << L), (1 << (L - 1 - K)), PAT>(v, L - 1 - K < nl, p), ...);
}

template <typename T, int N, int... J>
__device__ __forceinline__ void load_line(T (&v)[N], T const *g, int at, int stride,
                                          std::integer_sequence<int, J...>) {
  ((v[J] = g[at + J * stride]), ...);
}

template <typename T, int N, int... J>
__device__ __forceinline__ void store_line(T const (&v)[N], T *g, int at, int stride,
                                           std::integer_sequence<int, J...>) {
  ((g[at + J * stride] = v[J]), ...);
}

// Every line of an axis (period 2^L, LDS stride I, O outer blocks of `line`
// points), one line per thread at a time; ends with a barrier.
template <typename T, int NT, int L, bool FWD, int PAT>
__device__ __forceinline__ void axis_in_registers(T *g, int O, int I, int line, bool image,
                                                  int nl, BoxLaunch const &p) {
  constexpr int N = 1 << L;
  int const lines = O * I;
  for (int l = int(threadIdx.x); l < lines; l += NT) {
    int const ol = l / I;
    int const at = ol * line + (l - ol * I);
    T v[N];
    load_line(v, g, at, I, std::make_integer_sequence<int, N>{});
    register_levels<T, L, FWD, PAT>(v, nl, p, std::make_integer_sequence<int, L>{});
    store_line(v, g, at, I, std::make_integer_sequence<int, N>{});
    if (image)
      g[at + N * I] = v[0];
  }
  __syncthreads();
}

// ---- longer lines: one pass at a time over the whole box --------------------

// Every line of an axis, pass by pass with a barrier after each, then the
// image plane; ends with a barrier.
template <typename T, int NT>
__device__ __forceinline__ void axis_by_passes(T *g, int O, int I, int line, int m, int Lk,
                                               int nl, BoxLaunch const &p) {
  int const tid = threadIdx.x;
  int const n = 1 << Lk;
  // the thread walks (q, i) = (line block and update index, position in the
  // block's inner run) in steps of NT from its own start
  int const i0 = tid % I, q0 = tid / I, di = NT % I, dq = NT / I;
  for (int lv = 0; lv < nl; ++lv) {
    int const sh = p.forward ? lv : nl - 1 - lv;
    int const s = 1 << sh;
    int const jb = Lk - sh - 1; // log2 of the updates per line and pass
    int const Q = O << jb;
#pragma unroll
    for (int f = 0; f < 2; ++f) {
      if (f >= p.npass)
        break;
      T const c0 = static_cast<T>(p.coef[f][0]);
      T const c1 = static_cast<T>(p.coef[f][1]);
      T const c2 = static_cast<T>(p.coef[f][2]);
      int const off = p.offset[f] * s;
      int i = i0, q = q0;
      while (q < Q) {
        int const j = q & ((1 << jb) - 1);
        int const x = off + 2 * s * j;
        int const lo = x == 0 ? n - s : x - s;
        int const hi = x + s == n ? 0 : x + s;
        int const b0 = (q >> jb) * line + i;
        g[b0 + x * I] = lift3(c0, g[b0 + lo * I], c1, g[b0 + x * I], c2, g[b0 + hi * I]);
        i += di;
        q += dq;
        if (i >= I) {
          i -= I;
          ++q;
        }
      }
      __syncthreads();
    }
  }
  if (m == n + 1) { // the image plane from index 0
    for (int u = tid; u < O * I; u += NT) {
      int const ol = u / I;
      int const at = ol * line + (u - ol * I);
      g[at + n * I] = g[at];
    }
    __syncthreads();
  }
}

// Every axis of the box, in order: all levels of its schedule, all passes,
// then its image plane; ends with a barrier.  axes[k] = {stored extent,
// levels to sweep, log2 period} in LDS.
template <typename T, int NT, int RL, bool FWD, int PAT>
__device__ __forceinline__ void sweep_box(BoxLaunch const &p, int naxes, int const (*axes)[3],
                                          int size, int box, T *g) {
  int I = box; // LDS stride of the current axis
  for (int k = 0; k < naxes; ++k) {
    int const m = axes[k][0];
    I /= m;
    int const nl = axes[k][1];
    if (nl == 0)
      continue;
    int const Lk = axes[k][2];
    int const line = m * I;    // points per outer line block of this axis
    int const O = size / line; // outer line blocks
    bool const image = m == (1 << Lk) + 1;
    switch (Lk) {
    case 1:
      axis_in_registers<T, NT, 1, FWD, PAT>(g, O, I, line, image, nl, p);
      break;
    case 2:
      axis_in_registers<T, NT, 2, FWD, PAT>(g, O, I, line, image, nl, p);
      break;
    case 3:
      axis_in_registers<T, NT, 3, FWD, PAT>(g, O, I, line, image, nl, p);
      break;
    case 4:
      axis_in_registers<T, NT, 4, FWD, PAT>(g, O, I, line, image, nl, p);
      break;
    case 5:
      if constexpr (RL >= 5) {
        axis_in_registers<T, NT, 5, FWD, PAT>(g, O, I, line, image, nl, p);
        break;
      }
      [[fallthrough]];
    default:
      axis_by_passes<T, NT>(g, O, I, line, m, Lk, nl, p);
    }
  }
}

// One box per workgroup; consecutive boxes (the tiles of one row range,
// which share the HBM lines at their edges) on one die.  All of the box's
// loads are issued back to back into registers before it lands in LDS.
template <typename T, int NT, int CAP_BYTES, bool FWD, int PAT>
__global__ __launch_bounds__(NT, kBoxCapacities[2] / CAP_BYTES) void box_sweep(BoxLaunch p) {
  constexpr int kPoints = int(box_points(CAP_BYTES, int(sizeof(T))));
  constexpr int R = (kPoints + NT - 1) / NT; // points per thread
  static_assert(kBoxMaxAxes * 3 * sizeof(int) <= kBoxTableBytes, "axis table");
  __shared__ T g[kPoints];
  __shared__ int axes[kBoxMaxAxes][3];
  int64_t const per_xcd = gridDim.x / kXcds;
  int64_t const blk = int64_t(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (blk >= p.blocks)
    return;
  // the grid index is uniform: scalar loads of the descriptor
  BoxGrid const &d = p.grids[__builtin_amdgcn_readfirstlane(p.block_grid[blk])];
  BoxItem<T> const it = resolve<T>(d, blk);
  {
    T v[R];
    fetch<T, NT, R>(it, v);
    deposit<T, NT, R>(it, v, g);
  }
  if (int(threadIdx.x) < d.naxes) {
    axes[threadIdx.x][0] = d.ext[threadIdx.x];
    axes[threadIdx.x][1] = d.nlev[threadIdx.x];
    axes[threadIdx.x][2] = d.L[threadIdx.x];
  }
  __syncthreads();
  sweep_box<T, NT, register_levels_of<T, CAP_BYTES>(), FWD, PAT>(p, d.naxes, axes, it.size,
                                                                  it.box, g);
  write_back<T, NT, R>(it, g);
}

// The parities of the passes as one of the named tables' patterns, or
// kPatNone.
int pattern_of(BoxLaunch const &p) {
  if (p.npass == 1 && p.offset[0] == 1)
    return kPatOdd;
  if (p.npass == 2 && p.offset[0] == 1 && p.offset[1] == 0)
    return kPatOddEven;
  if (p.npass == 2 && p.offset[0] == 0 && p.offset[1] == 1)
    return kPatEvenOdd;
  return kPatNone;
}

template <typename T, int NT, int CAP, bool FWD>
hipError_t launch_direction(BoxLaunch const &p, dim3 grid, hipStream_t stream) {
  switch (pattern_of(p)) {
  case kPatOdd:
    hipLaunchKernelGGL((box_sweep<T, NT, CAP, FWD, kPatOdd>), grid, dim3(NT), 0, stream, p);
    break;
  case kPatOddEven:
    hipLaunchKernelGGL((box_sweep<T, NT, CAP, FWD, kPatOddEven>), grid, dim3(NT), 0, stream, p);
    break;
  case kPatEvenOdd:
    hipLaunchKernelGGL((box_sweep<T, NT, CAP, FWD, kPatEvenOdd>), grid, dim3(NT), 0, stream, p);
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T, int NT, int CAP>
hipError_t launch_cap(BoxLaunch const &p, hipStream_t stream) {
  dim3 const grid(static_cast<uint32_t>((p.blocks + kXcds - 1) / kXcds * kXcds));
  return p.forward ? launch_direction<T, NT, CAP, true>(p, grid, stream)
                   : launch_direction<T, NT, CAP, false>(p, grid, stream);
}

template <typename T, int NT>
hipError_t launch_typed(BoxLaunch const &p, int capacity_bytes, hipStream_t stream) {
  switch (capacity_bytes) {
  case kBoxCapacities[0]:
    return launch_cap<T, NT, kBoxCapacities[0]>(p, stream);
  case kBoxCapacities[1]:
    return launch_cap<T, NT, kBoxCapacities[1]>(p, stream);
  case kBoxCapacities[2]:
    return launch_cap<T, NT, kBoxCapacities[2]>(p, stream);
  default:
    return hipErrorInvalidValue;
  }
}

} // namespace

hipError_t launch_box_batched(int dtype, BoxLaunch const &launch, int capacity_bytes,
                              int threads, hipStream_t stream) {
  if (launch.blocks <= 0)
    return hipSuccess;
  if (launch.blocks > int64_t(0x7fffffff) - kXcds)
    return hipErrorInvalidValue;
  bool const f64 = dtype == PALIWA_FLOAT64;
  switch (threads) {
  case 256:
    return f64 ? launch_typed<double, 256>(launch, capacity_bytes, stream)
               : launch_typed<float, 256>(launch, capacity_bytes, stream);
  case 512:
    return f64 ? launch_typed<double, 512>(launch, capacity_bytes, stream)
               : launch_typed<float, 512>(launch, capacity_bytes, stream);
  default:
    return hipErrorInvalidValue;
  }
}

} // namespace paliwa_hip
