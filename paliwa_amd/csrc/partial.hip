// Partial-output transform along one axis: the reference's transform_in with
// an `outputs` domain and its packed-pole execution from local plus remote
// storage (paliwa/paliwa_transform.hpp:253-344, 53-210).
//
// A plan fixes the axis (stored at `level`, period P = 2^level, storage
// indices), the levels and passes, the local run [local_start, local_start +
// extents[dim]) held by the caller's block and the output run inside it.  The
// planner (plan.cpp) gives the exact writes of every pass and the required
// inputs; the inputs outside the local run are the ghosts, which the caller
// supplies in a second array (the block's shape with extents[dim] = number of
// ghosts, ordered by storage index).  Every line of the block is then one
// window of W slots: slot s holds storage index (cover_min + s) mod P, where
// [cover_min, cover_max] is the smallest periodic interval holding every
// required input.  A workgroup gathers the windows of TB lines into LDS (from
// the local block or the ghosts), runs exactly the planned writes of every
// pass in LDS (g = c0 g[lower] + c1 g + c2 g[upper], left to right, no
// contraction: the arithmetic of the full transform, so the outputs are
// bit-identical to it), and writes back the output points only.  Ghosts and
// the other local points are not modified.
#include "common.hpp"
#include "fused_sweep.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <limits>
#include <new>
#include <vector>

namespace {

using namespace paliwa_hip;

// The whole-period register form runs when kFusedShare * window >= period.
constexpr int64_t kFusedShare = 2;

constexpr int kThreads = 256;
constexpr int32_t kNoSource = std::numeric_limits<int32_t>::min();
constexpr int16_t kNoRow = std::numeric_limits<int16_t>::min();
constexpr int64_t kMaxWindowBytes = 64 * 1024;

struct PartialPass {
  int32_t first;  // slot of the first write
  int32_t count;  // writes
  int32_t step;   // slots between writes (2 * stride)
  int32_t stride; // slots to the neighbours
  double c[3];
};

struct PartialArgs {
  int32_t const *slot_source; // >= 0 local index, < 0: ghost -(g+1), kNoSource
  PartialPass const *passes;
  int32_t npasses;
  int32_t W;      // window slots
  int32_t period; // P: slot arithmetic is mod P
  int32_t TB;     // lines per workgroup (a power of two dividing kThreads)
  int32_t logTB;
  uint32_t magic_W, magic_out; // e / W, e / out_count by multiply-high (B == 1)
  int64_t lines, B, m_local, m_ghost;
  int32_t out_slot, out_local, out_count;
};

// e / d from the host's magic = 2^32 / d rounded up (exact for the e and d
// here); magic 0 stands for d = 1, whose 2^32 does not fit the word.
__device__ __forceinline__ int divide(int e, uint32_t magic) {
  return magic != 0 ? static_cast<int>(__umulhi(static_cast<uint32_t>(e), magic)) : e;
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
    partial_lines(PartialArgs a, T *__restrict__ local, T const *__restrict__ ghosts) {
  extern __shared__ unsigned char smem[];
  T *t = reinterpret_cast<T *>(smem); // [W][TB]
  int const TB = a.TB;
  int const logTB = a.logTB;
  int64_t const line0 = int64_t(blockIdx.x) * TB;
  if (a.B > 1) {
    // across lines: thread -> (line tb, slots s0, s0 + rows, ...), so each
    // thread's line and its base pointers are computed once
    int const tb = threadIdx.x & (TB - 1);
    int const rows = kThreads >> logTB;
    int const s0 = threadIdx.x >> logTB;
    int64_t const line = line0 + tb;
    bool const valid = line < a.lines;
    int64_t const ai = valid ? line / a.B : 0;
    int64_t const bi = valid ? line - ai * a.B : 0;
    T *const lbase = local + ai * a.m_local * a.B + bi;
    T const *const gbase = ghosts + ai * a.m_ghost * a.B + bi;
    for (int s = s0; s < a.W; s += rows) {
      T v = T(0);
      int32_t const src = a.slot_source[s];
      if (valid) {
        if (src >= 0)
          v = lbase[src * a.B];
        else if (src != kNoSource)
          v = gbase[int64_t(-src - 1) * a.B];
      }
      t[s * TB + tb] = v;
    }
  } else {
    // along lines: consecutive threads read consecutive slots of a line
    int const n = a.W * TB;
    for (int idx = threadIdx.x; idx < n; idx += kThreads) {
      int const tb = divide(idx, a.magic_W);
      int const s = idx - tb * a.W;
      int64_t const line = line0 + tb;
      T v = T(0);
      if (line < a.lines) {
        int32_t const src = a.slot_source[s];
        if (src >= 0)
          v = local[line * a.m_local + src];
        else if (src != kNoSource)
          v = ghosts[line * a.m_ghost + (-src - 1)];
      }
      t[s * TB + tb] = v;
    }
  }
  __syncthreads();
  for (int p = 0; p < a.npasses; ++p) {
    PartialPass const ps = a.passes[p];
    T const c0 = static_cast<T>(ps.c[0]);
    T const c1 = static_cast<T>(ps.c[1]);
    T const c2 = static_cast<T>(ps.c[2]);
    int const work = ps.count << logTB;
    for (int idx = threadIdx.x; idx < work; idx += kThreads) {
      int const i = idx >> logTB, tb = idx & (TB - 1);
      // slots mod P (inside the window by the plan)
      int x = ps.first + i * ps.step;
      if (x >= a.period)
        x -= a.period;
      int lo = x - ps.stride;
      if (lo < 0)
        lo += a.period;
      int hi = x + ps.stride;
      if (hi >= a.period)
        hi -= a.period;
      T const lower = t[lo * TB + tb];
      T const self = t[x * TB + tb];
      T const upper = t[hi * TB + tb];
      T const head = c0 * lower + c1 * self;
      t[x * TB + tb] = head + c2 * upper;
    }
    __syncthreads();
  }
  if (a.B > 1) {
    int const tb = threadIdx.x & (TB - 1);
    int const rows = kThreads >> logTB;
    int64_t const line = line0 + tb;
    if (line >= a.lines)
      return;
    int64_t const ai = line / a.B;
    int64_t const bi = line - ai * a.B;
    T *const lbase = local + ai * a.m_local * a.B + bi;
    for (int i = threadIdx.x >> logTB; i < a.out_count; i += rows) {
      int s = a.out_slot + i; // mod P, inside the window
      if (s >= a.period)
        s -= a.period;
      int64_t j = a.out_local + i; // wraps only when the local run is the whole period
      if (j >= a.m_local)
        j -= a.m_local;
      lbase[j * a.B] = t[s * TB + tb];
    }
  } else {
    int const outs = a.out_count * TB;
    for (int idx = threadIdx.x; idx < outs; idx += kThreads) {
      int const tb = divide(idx, a.magic_out);
      int const i = idx - tb * a.out_count;
      int64_t const line = line0 + tb;
      if (line >= a.lines)
        continue;
      int s = a.out_slot + i;
      if (s >= a.period)
        s -= a.period;
      int64_t j = a.out_local + i;
      if (j >= a.m_local)
        j -= a.m_local;
      local[line * a.m_local + j] = t[s * TB + tb];
    }
  }
}

// Along a strided axis whose period the fused kernel holds (a ladder of
// consecutive levels): the whole period of each line in registers, as in the
// fused sweep.  Rows of the window come from the local block or the ghosts,
// every other row is zero; the ladder then runs on all rows, and only the
// outputs are stored.  Bit-identical on the outputs: the planner puts every
// output's dependency cone inside the window, and no point computed from the
// zero rows is read by a point of that cone.
struct FusedPartialArgs {
  int16_t const *row_source; // [P] per row: local index, -(ghost + 1), kNoRow
  int64_t lines, B, m_local, m_ghost;
  int32_t out_row, out_local, out_count;
  int32_t lo, hi;            // ladder: log2 storage strides lo .. hi
  double coef[2][3];
};

// LDS of partial_fused: the whole period of TB lines, row-major with PAD
// points after every chunk of K rows (the fused sweep's strided layout, staged
// here for every G, G = 1 included), then the row table (int16 per row).
template <typename T, int K, int G>
constexpr size_t fused_partial_table_at() {
  using Geom = fused::TileGeom<K, G, true, T>;
  return (sizeof(T) * size_t(K * G * Geom::TB + G * Geom::PAD + 2) + 15) / 16 * 16;
}

template <typename T, int K, int G, int PAT, bool ASC>
__global__ __launch_bounds__(fused::threads_for<T>(true, K, G)) void partial_fused(
    FusedPartialArgs a, T *__restrict__ local, T const *__restrict__ ghosts,
    int64_t ntiles) {
  using Geom = fused::TileGeom<K, G, true, T>;
  constexpr int TB = Geom::TB;
  constexpr int NT = Geom::NT;
  constexpr int P = K * G;
  constexpr int RS = NT / TB; // rows per staging step (= G): K rows per thread
  extern __shared__ unsigned char smem[];
  T *const lds = reinterpret_cast<T *>(smem);
  int const tid = threadIdx.x;
  int const c = tid % TB;
  int const r0 = tid / TB;
  uint32_t const nwg = gridDim.x;
  // the row of the block or of the ghost array behind row r0 + k * RS, at
  // [r0 * K + k] after the tile: a thread's K entries are adjacent
  int16_t *const table = reinterpret_cast<int16_t *>(smem + fused_partial_table_at<T, K, G>());
  for (int x = tid; x < P; x += NT)
    table[(x % RS) * K + x / RS] = a.row_source[x];
  __syncthreads();
  // this thread's column of a tile in the block and in the ghost array (the
  // ghost array has the block's shape with m_ghost rows); null: past the grid
  struct Column {
    T *l;
    T const *g;
  };
  auto column = [&](int64_t tile) -> Column {
    int64_t const line = tile * TB + c;
    if (tile >= ntiles || line >= a.lines)
      return {nullptr, nullptr};
    int64_t const ai = line / a.B;
    int64_t const bi = line - ai * a.B;
    return {local + ai * a.m_local * a.B + bi,
            ghosts == nullptr ? nullptr : ghosts + ai * a.m_ghost * a.B + bi};
  };
  // rows r0, r0 + RS, ... of the column into registers (zeros off the window)
  auto fetch = [&](Column col, T (&buf)[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      int const src = table[r0 * K + k];
      T const *const base = src >= 0 ? col.l : col.g;
      int64_t const row = src >= 0 ? src : -src - 1;
      T v = T(0);
      if (col.l != nullptr && src != kNoRow)
        v = base[row * a.B];
      buf[k] = v;
    }
  };
  fused::Coefs<T> cf;
#pragma unroll
  for (int f = 0; f < 2; ++f)
#pragma unroll
    for (int k = 0; k < 3; ++k)
      cf.c[f][k] = static_cast<T>(a.coef[f][k]);
  int const li = tid / G;
  int const gi = tid % G;
  int const lane_base = (tid & 63) - gi;

  int64_t t = fused::xcd_tile(blockIdx.x, nwg);
  T buf[K];
  Column cur = column(t);
  fetch(cur, buf);
  for (; t < ntiles; t += nwg) {
    __syncthreads(); // the previous tile's output reads are done
#pragma unroll
    for (int k = 0; k < K; ++k)
      lds[fused::lds_index<K, G, true, T>(c, r0 + k * RS, 0)] = buf[k];
    __syncthreads();
    // the next tile's loads stay in flight while this one computes
    Column const nxt = column(t + nwg);
    if (t + nwg < ntiles)
      fetch(nxt, buf);
    T v[K];
#pragma unroll
    for (int j = 0; j < K; ++j)
      v[j] = lds[fused::lds_index<K, G, true, T>(li, gi * K + j, 0)];
    fused::run_ladder<T, K, G, PAT, ASC>(v, lane_base, gi, a.lo, a.hi, cf);
#pragma unroll
    for (int j = 0; j < K; ++j)
      lds[fused::lds_index<K, G, true, T>(li, gi * K + j, 0)] = v[j];
    __syncthreads();
    if (cur.l != nullptr) {
      for (int i = r0; i < a.out_count; i += RS) {
        int x = a.out_row + i;
        if (x >= P)
          x -= P;
        int64_t j = a.out_local + i; // wraps only when the local run is the whole period
        if (j >= a.m_local)
          j -= a.m_local;
        cur.l[j * a.B] = lds[fused::lds_index<K, G, true, T>(c, x, 0)];
      }
    }
    cur = nxt;
  }
}

template <typename T, int K, int G, int PAT, bool ASC>
hipError_t launch_partial_fused_kg(FusedPartialArgs const &a, T *local, T const *ghosts,
                                   hipStream_t s) {
  using Geom = fused::TileGeom<K, G, true, T>;
  size_t const bytes = fused_partial_table_at<T, K, G>() + sizeof(int16_t) * K * G;
  auto kernel = partial_fused<T, K, G, PAT, ASC>;
  if (bytes > 65536) {
    hipError_t const err = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               static_cast<int>(bytes));
    if (err != hipSuccess)
      return err;
  }
  int64_t const tiles = (a.lines + Geom::TB - 1) / Geom::TB;
  int64_t const resident = fused::resident_workgroups(reinterpret_cast<const void *>(kernel),
                                                      bytes, Geom::NT);
  uint32_t const grid = static_cast<uint32_t>(std::min<int64_t>(tiles, resident));
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(Geom::NT), bytes, s, a, local, ghosts, tiles);
  return hipGetLastError();
}

template <typename T, int K, int G>
hipError_t launch_partial_fused_pat(FusedPartialArgs const &a, int pattern, bool asc,
                                    T *local, T const *ghosts, hipStream_t s) {
  switch (pattern) {
  case kPatOdd:
    return asc ? launch_partial_fused_kg<T, K, G, kPatOdd, true>(a, local, ghosts, s)
               : launch_partial_fused_kg<T, K, G, kPatOdd, false>(a, local, ghosts, s);
  case kPatOddEven:
    return asc ? launch_partial_fused_kg<T, K, G, kPatOddEven, true>(a, local, ghosts, s)
               : launch_partial_fused_kg<T, K, G, kPatOddEven, false>(a, local, ghosts, s);
  case kPatEvenOdd:
    return asc ? launch_partial_fused_kg<T, K, G, kPatEvenOdd, true>(a, local, ghosts, s)
               : launch_partial_fused_kg<T, K, G, kPatEvenOdd, false>(a, local, ghosts, s);
  default:
    return hipErrorInvalidValue;
  }
}

template <typename T, int L, int LMAX>
hipError_t launch_partial_fused(FusedPartialArgs const &a, int period_log2, int pattern,
                                bool asc, T *local, T const *ghosts, hipStream_t s) {
  if constexpr (L > LMAX) {
    return hipErrorInvalidValue;
  } else {
    if (period_log2 == L) {
      constexpr int K = fused_chunk_of(sizeof(T), L);
      return launch_partial_fused_pat<T, K, (1 << L) / K>(a, pattern, asc, local, ghosts, s);
    }
    return launch_partial_fused<T, L + 1, LMAX>(a, period_log2, pattern, asc, local, ghosts, s);
  }
}

} // namespace

struct paliwa_partial {
  int32_t dtype = 0;
  int32_t ndim = 0;
  int32_t device = 0;
  int32_t dim = 0;
  int32_t blocks = 0;
  // the whole-period register form (strided axis, ladder, period within the
  // fused kernel's): used instead of partial_lines when set
  bool fused = false;
  int32_t level = 0, pattern = 0;
  bool asc = false;
  FusedPartialArgs fargs{};
  int64_t lds_bytes = 0;
  std::vector<int64_t> ghost_index;
  PartialArgs args{};
  void *dev = nullptr;
};

namespace {

void destroy(paliwa_partial *p) {
  if (p == nullptr)
    return;
  if (p->dev != nullptr) {
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(p->device);
    (void)hipFree(p->dev);
    (void)hipSetDevice(prev);
  }
  delete p;
}

} // namespace

extern "C" {

int paliwa_partial_create(paliwa_partial **plan, int32_t dtype, int32_t ndim,
                          const int64_t *extents, int32_t dim, int32_t level,
                          const int32_t *levels, int32_t nlevels,
                          const paliwa_filter *filters, int32_t nfilters,
                          int64_t local_start, int64_t output_start,
                          int64_t output_count) {
  if (plan == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  *plan = nullptr;
  if (dtype != PALIWA_FLOAT64 && dtype != PALIWA_FLOAT32)
    return fail(PALIWA_ERR_DTYPE, "dtype must be PALIWA_FLOAT64 or PALIWA_FLOAT32");
  if (ndim < 1 || ndim > PALIWA_MAX_DIMS || extents == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "ndim in [1, 26] and extents are required");
  if (dim < 0 || dim >= ndim)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "dim out of range");
  if (level < 0 || level > 30)
    return fail(PALIWA_ERR_LEVEL, "level out of range [0, 30]");
  int64_t const P = int64_t(1) << level;
  int64_t A = 1, B = 1;
  for (int d = 0; d < ndim; ++d) {
    if (extents[d] < 1)
      return fail(PALIWA_ERR_SHAPE, "extents must be >= 1");
    if (d < dim)
      A *= extents[d];
    if (d > dim)
      B *= extents[d];
  }
  int64_t const m_local = extents[dim];
  if (m_local > P)
    return fail(PALIWA_ERR_SHAPE, "the local run is longer than the period 2^level");
  if (local_start < 0 || local_start >= P || output_start < 0 || output_start >= P ||
      output_count < 0 || output_count > m_local)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "local or output run outside the period");
  int64_t const out_local = ((output_start - local_start) % P + P) % P;
  // a local run of the whole period is circular: outputs may wrap around it
  if (output_count > 0 && m_local < P && out_local + output_count > m_local)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "the output run is not inside the local run");
  int rc = check_filters(filters, nfilters);
  if (rc != PALIWA_OK)
    return rc;
  if (nlevels < 0 || nlevels > PALIWA_MAX_LEVELS || (nlevels > 0 && levels == nullptr))
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "levels: count in [0, 64] and array required");

  // the plan in storage coordinates (maximum level = level)
  paliwa_periodic_bounds const outputs =
      output_count == 0 ? paliwa_periodic_bounds{-1, -1}
                        : paliwa_periodic_bounds{output_start, (output_start + output_count - 1) % P};
  std::vector<paliwa_periodic_bounds> writes(std::max<int64_t>(1, int64_t(nlevels) * nfilters));
  std::vector<paliwa_periodic_bounds> required(level + 1);
  rc = plan_transform_bounds(outputs, level, level, levels, nlevels, filters, nfilters,
                             writes.data(), required.data());
  if (rc != PALIWA_OK)
    return rc;

  auto *pl = new (std::nothrow) paliwa_partial;
  if (pl == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "out of host memory");
  pl->dtype = dtype;
  pl->ndim = ndim;
  pl->dim = dim;
  (void)hipGetDevice(&pl->device);
  PartialArgs &a = pl->args;
  a.lines = A * B;
  a.B = B;
  a.m_local = m_local;
  a.period = static_cast<int32_t>(P);
  if (output_count == 0 || a.lines == 0) {
    *plan = pl;
    return PALIWA_OK;
  }

  // required inputs: the lattice points of each level's interval (disjoint)
  std::vector<int64_t> req;
  for (int32_t h = 0; h <= level; ++h) {
    paliwa_periodic_bounds const b = required[h];
    if (b.min < 0)
      continue;
    int64_t const stride = h == 0 ? P : int64_t(2) << (level - h);
    int64_t const last = b.max + (b.min > b.max ? P : 0);
    for (int64_t x = b.min; x <= last; x += stride)
      req.push_back(x % P);
  }
  std::sort(req.begin(), req.end());
  paliwa_periodic_bounds cover{};
  if (paliwa_covering_periodic_bounds(req.data(), static_cast<int64_t>(req.size()), P,
                                      &cover) != PALIWA_OK) {
    destroy(pl);
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "internal: required inputs not distinct");
  }
  int64_t const W = ((cover.max - cover.min) % P + P) % P + 1;
  int64_t const elem = dtype == PALIWA_FLOAT64 ? 8 : 4;
  if (W * elem > kMaxWindowBytes) {
    destroy(pl);
    return fail(PALIWA_ERR_UNSUPPORTED, "the window of required inputs exceeds 64 KiB per line");
  }
  auto slot_of = [&](int64_t x) { return ((x - cover.min) % P + P) % P; };
  // slot sources; ghosts are the required inputs outside the local run
  std::vector<int32_t> source(W, kNoSource);
  for (int64_t x : req) {
    int64_t const rel = ((x - local_start) % P + P) % P;
    if (rel < m_local) {
      source[slot_of(x)] = static_cast<int32_t>(rel);
    } else {
      source[slot_of(x)] = -static_cast<int32_t>(pl->ghost_index.size()) - 1;
      pl->ghost_index.push_back(x);
    }
  }
  a.m_ghost = static_cast<int64_t>(pl->ghost_index.size());
  // the passes in execution order, in slot coordinates
  std::vector<PartialPass> passes;
  for (int32_t i = 0; i < nlevels; ++i) {
    int64_t const stride = int64_t(1) << (level - levels[i]);
    for (int32_t p = 0; p < nfilters; ++p) {
      paliwa_periodic_bounds const w = writes[int64_t(i) * nfilters + p];
      if (w.min < 0)
        continue;
      int64_t const span = ((w.max - w.min) % P + P) % P;
      int64_t const count = span / (2 * stride) + 1;
      int64_t const first = slot_of(w.min);
      // every write and both its neighbours must be window slots (slot
      // arithmetic mod P: a run of writes may pass from slot P - 1 to 0 when
      // its lattice points skip the gap of the window)
      for (int64_t k = 0; k < count; ++k) {
        int64_t const x = (first + k * 2 * stride) % P;
        if (x >= W || (x - stride + P) % P >= W || (x + stride) % P >= W) {
          destroy(pl);
          return fail(PALIWA_ERR_INVALID_ARGUMENT, "internal: planned pass leaves the window");
        }
      }
      PartialPass ps{};
      ps.first = static_cast<int32_t>(first);
      ps.count = static_cast<int32_t>(count);
      ps.step = static_cast<int32_t>(2 * stride);
      ps.stride = static_cast<int32_t>(stride);
      for (int j = 0; j < 3; ++j)
        ps.c[j] = filters[p].coefficients[j];
      passes.push_back(ps);
    }
  }
  for (int64_t k = 0; k < output_count; ++k)
    if (slot_of(output_start + k) >= W) {
      destroy(pl);
      return fail(PALIWA_ERR_INVALID_ARGUMENT, "internal: an output is outside the window");
    }
  a.npasses = static_cast<int32_t>(passes.size());
  a.W = static_cast<int32_t>(W);
  a.out_slot = static_cast<int32_t>(slot_of(output_start));
  a.out_local = static_cast<int32_t>(out_local);
  a.out_count = static_cast<int32_t>(output_count);
  // lines per workgroup: as many windows as fit in 16 KiB, at most 64, a
  // power of two (so that it divides the workgroup size); smaller tiles keep
  // more workgroups per CU to cover the per-pass barriers (one slab of a
  // 256^3 biorthogonal sweep: 16 KiB 0.070 ms, 32 KiB 0.082, 64 KiB 0.118)
  int64_t tb = std::max<int64_t>(1, (16 * 1024) / (W * elem));
  tb = std::min<int64_t>({tb, 64, a.lines});
  int log_tb = 0;
  while ((int64_t(2) << log_tb) <= tb)
    ++log_tb;
  tb = int64_t(1) << log_tb;
  a.TB = static_cast<int32_t>(tb);
  a.logTB = log_tb;
  // exact while e * divisor < 2^32: e < W * TB (or out_count * TB) <= 2^16 * 64
  auto magic = [](int64_t d) -> uint32_t { // 0 for d == 1 (see divide())
    return d > 1 ? static_cast<uint32_t>(0xffffffffu / static_cast<uint32_t>(d) + 1) : 0u;
  };
  a.magic_W = magic(W);
  a.magic_out = magic(output_count);
  int64_t const blocks = (a.lines + tb - 1) / tb;
  if (blocks > INT32_MAX) {
    destroy(pl);
    return fail(PALIWA_ERR_UNSUPPORTED, "too many lines");
  }
  pl->blocks = static_cast<int32_t>(blocks);
  pl->lds_bytes = W * tb * elem;
  // the whole-period register form: a strided axis, a ladder of consecutive
  // levels with a table pattern, a period the fused kernel holds, and a
  // window of at least half the period (it computes every row of the period;
  // on one slab of a 256^3 float64 grid: window = period 0.035 ms against
  // 0.070 for partial_lines, a window of 77 rows 0.025 against 0.024, of 33
  // rows 0.024 against 0.011)
  std::vector<int16_t> row_source;
  if (B > 1 && level >= 1 && nlevels >= 1 && P <= fused_max_period(dtype) &&
      kFusedShare * W >= P) {
    SweepParams sp{};
    sp.nlevels = nlevels;
    for (int32_t i = 0; i < nlevels; ++i)
      sp.shift[i] = static_cast<int8_t>(level - levels[i]);
    set_filters(sp, filters, nfilters);
    classify(sp);
    if (sp.schedule != kScheduleGeneric && sp.pattern != kPatNone) {
      pl->fused = true;
      pl->level = level;
      pl->pattern = sp.pattern;
      pl->asc = sp.schedule == kScheduleLadderAsc;
      FusedPartialArgs &f = pl->fargs;
      f.lines = a.lines;
      f.B = B;
      f.m_local = m_local;
      f.m_ghost = a.m_ghost;
      f.out_row = static_cast<int32_t>(output_start);
      f.out_local = static_cast<int32_t>(out_local);
      f.out_count = static_cast<int32_t>(output_count);
      f.lo = sp.lo;
      f.hi = sp.hi;
      for (int q = 0; q < 2; ++q)
        for (int j = 0; j < 3; ++j)
          f.coef[q][j] = q < sp.npass ? sp.coef[q][j] : 0.0;
      // local and ghost indices are below the period (at most 2^12 here)
      row_source.assign(static_cast<size_t>(P), kNoRow);
      for (int64_t x = 0; x < P; ++x) {
        int64_t const slot = slot_of(x);
        if (slot < W && source[size_t(slot)] != kNoSource)
          row_source[size_t(x)] = static_cast<int16_t>(source[size_t(slot)]);
      }
    }
  }
  // device tables
  size_t const src_bytes = source.size() * sizeof(int32_t);
  size_t const pass_at = (src_bytes + 255) / 256 * 256;
  size_t const pass_bytes = std::max<size_t>(1, passes.size()) * sizeof(PartialPass);
  size_t const rows_at = (pass_at + pass_bytes + 255) / 256 * 256;
  size_t const total = rows_at + std::max<size_t>(1, row_source.size()) * sizeof(int16_t);
  std::vector<char> buf(total, 0);
  std::memcpy(buf.data(), source.data(), src_bytes);
  if (!passes.empty())
    std::memcpy(buf.data() + pass_at, passes.data(), passes.size() * sizeof(PartialPass));
  if (!row_source.empty())
    std::memcpy(buf.data() + rows_at, row_source.data(), row_source.size() * sizeof(int16_t));
  if (hipMalloc(&pl->dev, total) != hipSuccess ||
      hipMemcpy(pl->dev, buf.data(), total, hipMemcpyHostToDevice) != hipSuccess) {
    if (pl->dev == nullptr)
      (void)hipGetLastError();
    destroy(pl);
    return fail(PALIWA_ERR_HIP, "partial: device tables could not be allocated");
  }
  a.slot_source = static_cast<int32_t const *>(pl->dev);
  pl->fargs.row_source = reinterpret_cast<int16_t const *>(static_cast<char *>(pl->dev) + rows_at);
  a.passes = reinterpret_cast<PartialPass const *>(static_cast<char *>(pl->dev) + pass_at);
  *plan = pl;
  return PALIWA_OK;
}

int paliwa_partial_ghosts(const paliwa_partial *plan, int64_t *count,
                          int64_t *indices) {
  if (plan == nullptr || count == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan and count are required");
  *count = static_cast<int64_t>(plan->ghost_index.size());
  if (indices != nullptr && !plan->ghost_index.empty())
    std::memcpy(indices, plan->ghost_index.data(), plan->ghost_index.size() * sizeof(int64_t));
  return PALIWA_OK;
}

int paliwa_partial_transform(const paliwa_partial *plan, void *local,
                             const void *ghosts, void *stream) {
  if (plan == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (plan->blocks == 0)
    return PALIWA_OK;
  if (local == nullptr || (ghosts == nullptr && !plan->ghost_index.empty()))
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "local (and ghosts, if the plan has any) are required");
  auto s = static_cast<hipStream_t>(stream);
  if (plan->fused) {
    hipError_t const err =
        plan->dtype == PALIWA_FLOAT64
            ? launch_partial_fused<double, 1, 11>(plan->fargs, plan->level, plan->pattern,
                                                  plan->asc, static_cast<double *>(local),
                                                  static_cast<double const *>(ghosts), s)
            : launch_partial_fused<float, 1, 12>(plan->fargs, plan->level, plan->pattern,
                                                 plan->asc, static_cast<float *>(local),
                                                 static_cast<float const *>(ghosts), s);
    if (err != hipSuccess)
      return fail(PALIWA_ERR_HIP, "partial_fused launch failed");
    return PALIWA_OK;
  }
  size_t const lds = static_cast<size_t>(plan->lds_bytes);
  if (plan->dtype == PALIWA_FLOAT64)
    partial_lines<double><<<plan->blocks, kThreads, lds, s>>>(
        plan->args, static_cast<double *>(local), static_cast<double const *>(ghosts));
  else
    partial_lines<float><<<plan->blocks, kThreads, lds, s>>>(
        plan->args, static_cast<float *>(local), static_cast<float const *>(ghosts));
  if (hipGetLastError() != hipSuccess)
    return fail(PALIWA_ERR_HIP, "partial_lines launch failed");
  return PALIWA_OK;
}

int paliwa_partial_destroy(paliwa_partial *plan) {
  destroy(plan);
  return PALIWA_OK;
}

} // extern "C"
