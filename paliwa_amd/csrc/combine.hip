// Combination-technique combine: the step after the batched hierarchize in
// the reference's workflow (test/test_paliwa_combination_technique.cpp:
// 176-313).  Every hierarchical subspace k (0 <= k_d <= l_d) of the component
// grids that at least `min_count` grids contain (the reference keeps count > 1,
// :217-233) gets sparse-grid storage holding, per point, the fold over the
// grids that contain it, in grid order, of  s += coefficient * grid  (product
// rounded, then the sum; no fused multiply-add, as the reference's host
// build).  The full grid at the maximum level then receives 0 + s at each
// point of a counted subspace and 0 elsewhere (:292-313).
//
// Work layout.  A prefix p = (k_0..k_{D-2}) with K_p = its largest counted
// k_{D-1} owns rows_p = prod 2^max(k_d - 1, 0) rows; row r holds the points
// of the subspaces (p, 0..K_p) at prefix point r, 2^K_p in all, and each
// component grid containing (p, 0) stores them in one of its rows.  Blocks of
// 1024 positions of one prefix hold R = 1024 / 2^K_p whole rows (grouped by
// subspace, rows outer, m inner) or, for longer rows, 1024 positions of one
// row in heap order (y = 0 -> k = 0; y >= 1 -> k = floor(log2 y) + 1,
// m = y - 2^(k-1)); either way the subspace level never decreases along a
// block.  A wavefront takes 64 consecutive positions of a block, one per lane,
// and folds over the grids that hold its smallest level (so the wavefronts of
// a block's fine levels walk short lists), in grid order: the fold order per
// point is the grid order.  Wavefronts are dispatched longest list first.
// The critical path is the longest list (1325 grids for the 4-D 1325-grid
// set); it is walked kUnroll grids per step with the next step's loads in
// flight (see combine_gather).
#include "common.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cstring>
#include <map>
#include <new>
#include <vector>

namespace {

using namespace paliwa_hip;

constexpr int kCombineDims = PALIWA_COMBINE_MAX_DIMS;
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kLogBlock = 10;
constexpr int64_t kBlockPositions = int64_t(1) << kLogBlock;
constexpr int kUnroll = 8;                      // grids per step
constexpr int kChunk = 64;                      // term records per chunk
constexpr int kSteps = kChunk / kUnroll;        // steps per chunk
static_assert(kChunk % kUnroll == 0, "a step never straddles two chunks");
// term records read past the end of a list (prefetch of the next chunk, the
// records of the step after the last): the table ends with this many
// padding records, so such reads stay inside it
constexpr int kPadRecords = 2 * kChunk + kUnroll;
constexpr int kLevelShift = 56; // last-axis level in the pointer word's top byte

struct PrefixDesc {
  int64_t rows;       // points of the prefix subspace
  int32_t first_sub;  // subspace id of (p, 0); (p, k) is first_sub + k
  int32_t K;          // largest counted k_{D-1}
  int32_t k[kCombineDims];
};

// One wavefront's work: 64 consecutive positions t0 .. t0+63 of a 1024-point
// block of a prefix, and the grids (in grid order) holding any of them.
struct WaveJob {
  int32_t prefix;
  int32_t local;      // block index within the prefix
  int32_t t0;         // first position within the block (a multiple of 64)
  int32_t list_begin; // first term record
  int32_t list_len;
  int32_t pad;
};

// A term record: one grid holding a prefix's points, D + 1 words
// {data pointer | last-axis level << 56, coefficient bits, mult_0..mult_{D-2}};
// the grid element of a point is sum_d (2 m_d + 1) mult_d over the prefix axes
// (mult_d = stride_d << (l_d - k_d), 0 where k_d = 0) plus, along the last
// axis, (2 m + 1) << (l - k) for k >= 1 and 0 for k = 0.
struct CombineArgs {
  PrefixDesc const *prefixes;
  int64_t const *sub_offset; // first sparse element of each subspace
  int64_t const *records;    // per (prefix, smallest last level), grid order
  WaveJob const *jobs;       // [njobs], longest grid list first
  int32_t const *prefix_of;  // dense prefix-level table -> prefix id or -1
  int32_t ndim;
  int32_t nprefix;
  int32_t njobs;
  int32_t narrow;                    // records of the narrow layout
  int32_t max_level[kCombineDims];
  int32_t prefix_mult[kCombineDims]; // dense index multipliers of prefix levels
};

__device__ __forceinline__ int shape_log(int k) { return k > 0 ? k - 1 : 0; }

template <typename T> struct Accumulate;
template <> struct Accumulate<double> {
  __device__ static double step(double s, double c, double v) {
    return __dadd_rn(s, __dmul_rn(c, v));
  }
};
template <> struct Accumulate<float> {
  // float grid, double coefficient: product and sum in double, stored as float
  __device__ static float step(float s, double c, float v) {
    return static_cast<float>(__dadd_rn(static_cast<double>(s),
                                        __dmul_rn(c, static_cast<double>(v))));
  }
};

// lane `src`'s 64-bit word, in scalar registers
__device__ __forceinline__ int64_t lane_word(int64_t v, int src) {
  uint32_t const lo = static_cast<uint32_t>(
      __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(v)), src));
  uint32_t const hi = static_cast<uint32_t>(__builtin_amdgcn_readlane(
      static_cast<int>(static_cast<uint32_t>(static_cast<uint64_t>(v) >> 32)), src));
  return static_cast<int64_t>((uint64_t(hi) << 32) | lo);
}

// Global-memory pointer: its loads are global_load (counted by vmcnt only),
// not flat loads, which also count against lgkmcnt and would make every
// wait on a record broadcast wait for the grid loads as well.
template <typename T> using global_ptr = T const __attribute__((address_space(1))) *;

// Words per term record.  Wide records hold each prefix multiplier in a
// 64-bit word; narrow records (every grid below 2^24 elements, so every
// element index, odd multiplier and prefix multiplier fits in 24 bits) pack
// two 32-bit multipliers per word and index with 24-bit multiply-adds.
template <int D, bool NARROW> constexpr int record_words() {
  return NARROW ? 2 + D / 2 : D + 1;
}
constexpr int64_t kNarrowElements = int64_t(1) << 24;

// The records of one step (kUnroll grids), wave-uniform.
template <typename T, int D, bool NARROW> struct StepRecords {
  using Mult = std::conditional_t<NARROW, uint32_t, int64_t>;
  global_ptr<T> ptr[kUnroll];
  double coef[kUnroll];
  int lg[kUnroll];
  Mult mult[kUnroll][D > 1 ? D - 1 : 1];
};

// Gather over the grids of each wavefront's list.  Lane = position; the
// records of a list arrive 64 at a time (lane j holds record j of the chunk)
// and are broadcast with v_readlane.  Every load is issued unconditionally
// (positions outside a grid read its element 0 and drop the value), so the
// loop has no branch inside a chunk and the loads of step s + 1 stay in
// flight while step s accumulates: the fold over a long list is paced by
// issue, not by one memory latency per step.
template <typename T, int D, bool NARROW>
__global__ void __launch_bounds__(kThreads)
    combine_gather(CombineArgs a, T *__restrict__ sparse) {
  using Records = StepRecords<T, D, NARROW>;
  constexpr int P = D > 1 ? D - 1 : 1;            // prefix axes (at least one slot)
  constexpr int W = record_words<D, NARROW>();    // words per term record
  constexpr int U = kUnroll;
  int const lane = threadIdx.x & 63;
  int const wj = static_cast<int>(blockIdx.x) * kWaves +
                 __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  if (wj >= a.njobs)
    return;
  WaveJob const job = a.jobs[wj];
  PrefixDesc const &pd = a.prefixes[job.prefix];
  int const K = pd.K;
  // this lane's point: subspace level k along the last axis, odd
  // multipliers 2 m_d + 1 along the prefix axes, output index
  int const t = job.t0 + lane;
  int k;
  int64_t row, m;
  if (K <= kLogBlock) {
    int const logR = kLogBlock - K;
    int const u = t >> logR;
    k = u == 0 ? 0 : 32 - __clz(u);
    int const logw = k == 0 ? 0 : k - 1;
    int const within = t - (k == 0 ? 0 : 1 << (logR + k - 1));
    row = (int64_t(job.local) << logR) + (within >> logw);
    m = within & ((1 << logw) - 1);
    if (row >= pd.rows)
      k = 64; // no point: never inside a grid
  } else {
    int const split = K - kLogBlock; // a row spans 2^split blocks
    row = int64_t(job.local) >> split;
    int64_t const y =
        ((int64_t(job.local) & ((int64_t(1) << split) - 1)) << kLogBlock) + t;
    k = y == 0 ? 0 : 64 - __clzll(y);
    m = y == 0 ? 0 : y - (int64_t(1) << (k - 1));
  }
  int64_t const col_odd = 2 * m + 1;
  int64_t o[P];
  {
    int64_t rest = row;
#pragma unroll
    for (int d = P - 1; d >= 0; --d) {
      int const sl = d < D - 1 ? shape_log(pd.k[d]) : 0;
      o[d] = 2 * (rest & ((int64_t(1) << sl) - 1)) + 1;
      rest >>= sl;
    }
  }
  int64_t const out = k <= K ? a.sub_offset[pd.first_sub + k] + (row << shape_log(k)) + m : -1;

  int64_t const *rec = a.records + int64_t(job.list_begin) * W;
  int const len = job.list_len;
  int64_t cur[W], nxt[W];
#pragma unroll
  for (int w = 0; w < W; ++w)
    cur[w] = rec[int64_t(lane) * W + w];

  auto extract = [&](int64_t const (&ch)[W], int base, Records &r) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      int64_t const pw = lane_word(ch[0], base + u);
      r.ptr[u] = (global_ptr<T>)(static_cast<uintptr_t>(pw & ((int64_t(1) << kLevelShift) - 1)));
      r.lg[u] = static_cast<int>(static_cast<uint64_t>(pw) >> kLevelShift);
      r.coef[u] = __longlong_as_double(lane_word(ch[1], base + u));
#pragma unroll
      for (int d = 0; d < D - 1; ++d) {
        if constexpr (NARROW) {
          uint64_t const word = static_cast<uint64_t>(ch[2 + d / 2]);
          r.mult[u][d] = static_cast<uint32_t>(__builtin_amdgcn_readlane(
              static_cast<int>(static_cast<uint32_t>(d % 2 ? word >> 32 : word)), base + u));
        } else {
          r.mult[u][d] = lane_word(ch[2 + d], base + u);
        }
      }
    }
  };
  // grid e + u holds this lane's point
  auto inside = [&](Records const &r, int e, int u) {
    return e + u < len && k <= r.lg[u];
  };
  auto issue = [&](Records const &r, int e, T (&v)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      bool const in = inside(r, e, u);
      if constexpr (NARROW) {
        uint32_t idx = k > 0 ? static_cast<uint32_t>(col_odd) << (in ? r.lg[u] - k : 0) : 0u;
#pragma unroll
        for (int d = 0; d < D - 1; ++d)
          idx += __umul24(static_cast<uint32_t>(o[d]), r.mult[u][d]);
        v[u] = r.ptr[u][in ? idx : 0u];
      } else {
        int64_t idx = k > 0 ? col_odd << (in ? r.lg[u] - k : 0) : 0;
#pragma unroll
        for (int d = 0; d < D - 1; ++d)
          idx += o[d] * r.mult[u][d];
        v[u] = r.ptr[u][in ? idx : 0];
      }
    }
  };
  T acc = T(0);
  auto consume = [&](Records const &r, int e, T const (&v)[U]) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      T const next = Accumulate<T>::step(acc, r.coef[u], v[u]);
      acc = inside(r, e, u) ? next : acc;
    }
  };

  Records r0, r1;
  T v0[U], v1[U];
  extract(cur, 0, r0);
  issue(r0, 0, v0);
  for (int c0 = 0;; c0 += kChunk) {
    // steps c0, c0 + U, ... of this chunk, unrolled: word w of the next
    // chunk's records is fetched at step w mod kSteps
#pragma unroll
    for (int j = 0; j < kSteps; ++j) {
      int const e = c0 + j * U;
#pragma unroll
      for (int w = j; w < W; w += kSteps)
        nxt[w] = rec[int64_t(c0 + kChunk + lane) * W + w];
      if (j + 1 < kSteps)
        extract(cur, (j + 1) * U, r1);
      else
        extract(nxt, 0, r1);
      issue(r1, e + U, v1);
      // keep the scheduler from sinking these loads to their uses a step
      // later, which would put one memory latency back on every step
      __builtin_amdgcn_sched_barrier(0);
      consume(r0, e, v0);
      if (e + U >= len)
        goto done;
      r0 = r1;
#pragma unroll
      for (int u = 0; u < U; ++u)
        v0[u] = v1[u];
    }
#pragma unroll
    for (int w = 0; w < W; ++w)
      cur[w] = nxt[w];
  }
done:
  if (out >= 0)
    sparse[out] = acc;
}

struct FullExtents {
  int64_t e[kCombineDims];
};

// Full grid at the maximum level (extent 2^M or 2^M+1 per axis; an image
// plane repeats index 0): 0 + the sparse value at points of counted
// subspaces, 0 elsewhere.  One workgroup per (row, chunk of the last axis).
template <typename T, int D>
__global__ void __launch_bounds__(kThreads)
    combine_scatter(CombineArgs a, T const *__restrict__ sparse,
                    T *__restrict__ full, int64_t chunks, FullExtents ext) {
  int64_t const b = blockIdx.x;
  int64_t row = b / chunks;
  int64_t const i_last = (b - row * chunks) * kThreads + threadIdx.x;
  if (i_last >= ext.e[D - 1])
    return;
  int64_t const row_id = row;
  int pk[D > 1 ? D - 1 : 1];
  int64_t pm[D > 1 ? D - 1 : 1];
  int32_t dense = 0;
#pragma unroll
  for (int d = D - 2; d >= 0; --d) {
    int64_t const i = row % ext.e[d];
    row /= ext.e[d];
    int const M = a.max_level[d];
    int64_t const x = i & ((int64_t(1) << M) - 1);
    int const kd = x == 0 ? 0 : M - __ffsll(static_cast<long long>(x)) + 1;
    pk[d] = kd;
    pm[d] = kd == 0 ? 0 : x >> (M - kd + 1);
    dense += kd * a.prefix_mult[d];
  }
  int const M = a.max_level[D - 1];
  int64_t const x = i_last & ((int64_t(1) << M) - 1);
  int const k = x == 0 ? 0 : M - __ffsll(static_cast<long long>(x)) + 1;
  int64_t const m = k == 0 ? 0 : x >> (M - k + 1);
  T value = T(0);
  int32_t const p = a.prefix_of[dense];
  if (p >= 0 && k <= a.prefixes[p].K) {
    int64_t const sub_begin = a.sub_offset[a.prefixes[p].first_sub + k];
    int64_t local = 0;
#pragma unroll
    for (int d = 0; d < D - 1; ++d)
      local = (local << shape_log(pk[d])) + pm[d];
    local = (local << shape_log(k)) + m;
    value = T(0) + sparse[sub_begin + local];
  }
  full[row_id * ext.e[D - 1] + i_last] = value;
}

template <typename T, int D>
void launch_gather(CombineArgs const &a, int32_t nblocks, void *sparse,
                   hipStream_t s) {
  if (a.narrow)
    combine_gather<T, D, true><<<nblocks, kThreads, 0, s>>>(a, static_cast<T *>(sparse));
  else
    combine_gather<T, D, false><<<nblocks, kThreads, 0, s>>>(a, static_cast<T *>(sparse));
}

template <typename T, int D>
void launch_scatter(CombineArgs const &a, unsigned blocks, void const *sparse,
                    void *full, int64_t chunks, FullExtents const &ext,
                    hipStream_t s) {
  combine_scatter<T, D><<<blocks, kThreads, 0, s>>>(
      a, static_cast<T const *>(sparse), static_cast<T *>(full), chunks, ext);
}

template <typename T>
void gather_dispatch(int D, CombineArgs const &a, int32_t nblocks, void *sparse,
                     hipStream_t s) {
  switch (D) {
  case 1: launch_gather<T, 1>(a, nblocks, sparse, s); break;
  case 2: launch_gather<T, 2>(a, nblocks, sparse, s); break;
  case 3: launch_gather<T, 3>(a, nblocks, sparse, s); break;
  case 4: launch_gather<T, 4>(a, nblocks, sparse, s); break;
  case 5: launch_gather<T, 5>(a, nblocks, sparse, s); break;
  case 6: launch_gather<T, 6>(a, nblocks, sparse, s); break;
  case 7: launch_gather<T, 7>(a, nblocks, sparse, s); break;
  default: launch_gather<T, 8>(a, nblocks, sparse, s); break;
  }
}

template <typename T>
void scatter_dispatch(int D, CombineArgs const &a, unsigned blocks,
                      void const *sparse, void *full, int64_t chunks,
                      FullExtents const &ext, hipStream_t s) {
  switch (D) {
  case 1: launch_scatter<T, 1>(a, blocks, sparse, full, chunks, ext, s); break;
  case 2: launch_scatter<T, 2>(a, blocks, sparse, full, chunks, ext, s); break;
  case 3: launch_scatter<T, 3>(a, blocks, sparse, full, chunks, ext, s); break;
  case 4: launch_scatter<T, 4>(a, blocks, sparse, full, chunks, ext, s); break;
  case 5: launch_scatter<T, 5>(a, blocks, sparse, full, chunks, ext, s); break;
  case 6: launch_scatter<T, 6>(a, blocks, sparse, full, chunks, ext, s); break;
  case 7: launch_scatter<T, 7>(a, blocks, sparse, full, chunks, ext, s); break;
  default: launch_scatter<T, 8>(a, blocks, sparse, full, chunks, ext, s); break;
  }
}

} // namespace

struct paliwa_combine {
  int32_t dtype = 0;
  int32_t ndim = 0;
  int32_t device = 0;
  int32_t nsub = 0;
  int32_t nblocks = 0;
  int64_t sparse_points = 0;
  int64_t terms = 0;
  std::vector<int32_t> sub_levels; // [nsub][ndim]
  std::vector<int64_t> sub_offsets; // [nsub + 1]
  std::vector<int32_t> sub_counts;
  CombineArgs args{};
  void *dev = nullptr;
  int64_t dev_bytes = 0;
};

namespace {

void destroy(paliwa_combine *c) {
  if (c == nullptr)
    return;
  if (c->dev != nullptr) {
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(c->device);
    (void)hipFree(c->dev);
    (void)hipSetDevice(prev);
  }
  delete c;
}

template <typename V> void append_bytes(std::vector<char> &buf, V const &v, int64_t &at) {
  at = static_cast<int64_t>((buf.size() + 255) / 256 * 256);
  buf.resize(at + v.size() * sizeof(typename V::value_type));
  if (!v.empty())
    std::memcpy(buf.data() + at, v.data(), v.size() * sizeof(typename V::value_type));
}

} // namespace

extern "C" {

int paliwa_combine_create(paliwa_combine **plan, int32_t dtype, int32_t ndim,
                          int32_t ngrids, void *const *data,
                          const int64_t *extents, const int32_t *levels,
                          const double *coefficients,
                          const int32_t *maximum_level, int32_t min_count) {
  if (plan == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  *plan = nullptr;
  if (ndim < 1 || ndim > kCombineDims)
    return fail(PALIWA_ERR_UNSUPPORTED, "combine: ndim must be in [1, 8]");
  if (ngrids < 1 || data == nullptr || extents == nullptr || levels == nullptr ||
      coefficients == nullptr || maximum_level == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT,
                "combine: ngrids >= 1 and every array argument are required");
  if (min_count < 1)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "combine: min_count must be >= 1");
  for (int d = 0; d < ndim; ++d)
    if (maximum_level[d] < 0 || maximum_level[d] > 30)
      return fail(PALIWA_ERR_LEVEL, "combine: maximum_level out of range [0, 30]");
  for (int32_t g = 0; g < ngrids; ++g) {
    int64_t n = 0;
    int rc = check_grid(data[g], dtype, ndim, extents + int64_t(g) * ndim,
                        levels + int64_t(g) * ndim, maximum_level, &n);
    if (rc != PALIWA_OK)
      return fail(rc, "grid " + std::to_string(g) + ": " + last_error());
  }
  int64_t dense_prefixes = 1;
  for (int d = 0; d + 1 < ndim; ++d)
    dense_prefixes *= maximum_level[d] + 1;
  if (dense_prefixes > (int64_t(1) << 24))
    return fail(PALIWA_ERR_UNSUPPORTED, "combine: too many prefix levels");

  // subspace counts: every k <= l of every grid (paliwa_utils.hpp:9-25)
  using Key = std::vector<int32_t>;
  std::map<Key, int32_t> count;
  for (int32_t g = 0; g < ngrids; ++g) {
    int32_t const *l = levels + int64_t(g) * ndim;
    Key k(ndim, 0);
    while (true) {
      ++count[k];
      int d = ndim - 1;
      while (d >= 0 && k[d] == l[d]) {
        k[d] = 0;
        --d;
      }
      if (d < 0)
        break;
      ++k[d];
    }
  }
  auto *c = new (std::nothrow) paliwa_combine;
  if (c == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "out of host memory");
  c->dtype = dtype;
  c->ndim = ndim;
  (void)hipGetDevice(&c->device);

  // counted subspaces in lexicographic order; prefixes (all but the last
  // level) with their largest counted last level
  std::vector<PrefixDesc> prefixes;
  std::vector<int64_t> sub_offset;
  int64_t offset = 0;
  for (auto const &[k, n] : count) {
    if (n < min_count)
      continue;
    Key prefix(k.begin(), k.end() - 1);
    bool const new_prefix =
        prefixes.empty() ||
        !std::equal(prefix.begin(), prefix.end(), prefixes.back().k);
    // counts fall with the last level: a prefix's counted last levels are
    // 0..K without gaps
    if (new_prefix ? k.back() != 0 : k.back() != prefixes.back().K + 1) {
      destroy(c);
      return fail(PALIWA_ERR_INVALID_ARGUMENT, "combine: inconsistent subspace counts");
    }
    if (new_prefix) {
      PrefixDesc pd{};
      pd.first_sub = static_cast<int32_t>(sub_offset.size());
      pd.K = 0;
      for (int d = 0; d + 1 < ndim; ++d)
        pd.k[d] = prefix[d];
      prefixes.push_back(pd);
    } else {
      prefixes.back().K = k.back();
    }
    int64_t size = 1;
    for (int d = 0; d < ndim; ++d)
      size <<= (k[d] > 0 ? k[d] - 1 : 0);
    sub_offset.push_back(offset);
    c->sub_levels.insert(c->sub_levels.end(), k.begin(), k.end());
    c->sub_offsets.push_back(offset);
    c->sub_counts.push_back(n);
    c->terms += size * n;
    offset += size;
  }
  c->sub_offsets.push_back(offset);
  c->nsub = static_cast<int32_t>(sub_offset.size());
  c->sparse_points = offset;
  // wavefront jobs: per prefix, blocks of 1024 positions (whole rows when a
  // row has at most 1024 points), 64 positions per wavefront; a wavefront's
  // list holds the grids containing (p, 0) whose last level reaches its
  // smallest last-axis subspace level, in grid order
  std::vector<WaveJob> jobs;
  std::vector<int64_t> records;
  bool narrow = true;
  for (int32_t g = 0; g < ngrids; ++g) {
    int64_t n = 1;
    for (int d = 0; d < ndim; ++d)
      n *= extents[int64_t(g) * ndim + d];
    narrow = narrow && n < kNarrowElements;
  }
  int64_t const W = narrow ? 2 + ndim / 2 : ndim + 1;
  std::vector<int32_t> holders;
  std::vector<int64_t> grid_stride(int64_t(ngrids) * ndim);
  for (int32_t g = 0; g < ngrids; ++g) {
    int64_t stride = 1;
    for (int d = ndim - 1; d >= 0; --d) {
      grid_stride[int64_t(g) * ndim + d] = stride;
      stride *= extents[int64_t(g) * ndim + d];
    }
    uint64_t const ptr = reinterpret_cast<uintptr_t>(data[g]);
    if ((ptr >> kLevelShift) != 0) {
      destroy(c);
      return fail(PALIWA_ERR_UNSUPPORTED, "combine: grid address above 2^56");
    }
  }
  for (size_t p = 0; p < prefixes.size(); ++p) {
    PrefixDesc &pd = prefixes[p];
    int64_t rows = 1;
    for (int d = 0; d + 1 < ndim; ++d)
      rows <<= (pd.k[d] > 0 ? pd.k[d] - 1 : 0);
    pd.rows = rows;
    holders.clear();
    for (int32_t g = 0; g < ngrids; ++g) {
      int32_t const *l = levels + int64_t(g) * ndim;
      bool inside = true;
      for (int d = 0; d + 1 < ndim; ++d)
        inside = inside && pd.k[d] <= l[d];
      if (inside)
        holders.push_back(g);
    }
    // list of the holders whose last level reaches kmin, built once per kmin
    std::vector<std::pair<int32_t, int32_t>> list_of(pd.K + 1, {-1, 0});
    auto list_for = [&](int kmin) -> std::pair<int32_t, int32_t> {
      if (list_of[kmin].first >= 0)
        return list_of[kmin];
      int64_t const begin = int64_t(records.size()) / W;
      for (int32_t g : holders) {
        int32_t const *l = levels + int64_t(g) * ndim;
        if (l[ndim - 1] < kmin)
          continue;
        uint64_t const ptr = reinterpret_cast<uintptr_t>(data[g]);
        records.push_back(static_cast<int64_t>(ptr | (uint64_t(l[ndim - 1]) << kLevelShift)));
        int64_t word = 0;
        std::memcpy(&word, coefficients + g, sizeof(double));
        records.push_back(word);
        int64_t pair = 0;
        for (int d = 0; d + 1 < ndim; ++d) {
          int64_t const mult =
              pd.k[d] > 0 ? grid_stride[int64_t(g) * ndim + d] << (l[d] - pd.k[d]) : 0;
          if (!narrow) {
            records.push_back(mult);
          } else if (d % 2 == 0) {
            pair = mult;
            if (d + 2 >= ndim) // last multiplier: the word's high half stays 0
              records.push_back(pair);
          } else {
            records.push_back(pair | (mult << 32));
          }
        }
      }
      int64_t const len = int64_t(records.size()) / W - begin;
      if (begin > INT32_MAX - kPadRecords)
        return {-1, -1};
      list_of[kmin] = {static_cast<int32_t>(begin), static_cast<int32_t>(len)};
      return list_of[kmin];
    };
    int64_t const nb = pd.K <= kLogBlock
                           ? (rows + (int64_t(1) << (kLogBlock - pd.K)) - 1) >> (kLogBlock - pd.K)
                           : rows << (pd.K - kLogBlock);
    for (int64_t b = 0; b < nb; ++b) {
      for (int32_t t0 = 0; t0 < kBlockPositions; t0 += 64) {
        // smallest subspace level among the wavefront's points
        int kmin = -1;
        for (int32_t t = t0; t < t0 + 64 && kmin < 0; ++t) {
          if (pd.K <= kLogBlock) {
            int const logR = kLogBlock - pd.K;
            int const u = t >> logR;
            int const k = u == 0 ? 0 : 32 - __builtin_clz(static_cast<unsigned>(u));
            int const logw = k == 0 ? 0 : k - 1;
            int const within = t - (k == 0 ? 0 : 1 << (logR + k - 1));
            if ((b << logR) + (within >> logw) < rows)
              kmin = k;
          } else {
            int const split = pd.K - kLogBlock;
            int64_t const y = ((b & ((int64_t(1) << split) - 1)) << kLogBlock) + t;
            kmin = y == 0 ? 0 : 64 - __builtin_clzll(static_cast<unsigned long long>(y));
          }
        }
        if (kmin < 0)
          continue; // no point in this wavefront
        auto const [begin, len] = list_for(kmin);
        if (begin < 0 || b > INT32_MAX || jobs.size() >= size_t(INT32_MAX) / kWaves) {
          destroy(c);
          return fail(PALIWA_ERR_UNSUPPORTED, "combine: sparse grid too large");
        }
        jobs.push_back({static_cast<int32_t>(p), static_cast<int32_t>(b), t0, begin, len, 0});
      }
    }
  }
  // padding records: a valid pointer, never inside (they lie past every list)
  for (int i = 0; i < kPadRecords; ++i) {
    records.push_back(static_cast<int64_t>(reinterpret_cast<uintptr_t>(data[0])));
    for (int64_t w = 1; w < W; ++w)
      records.push_back(0);
  }
  // longest lists first: they set the gather's critical path
  std::stable_sort(jobs.begin(), jobs.end(), [](WaveJob const &x, WaveJob const &y) {
    return x.list_len > y.list_len;
  });
  int64_t const njobs = static_cast<int64_t>(jobs.size());
  // dense prefix table
  std::vector<int32_t> prefix_of(dense_prefixes, -1);
  int32_t mult[kCombineDims] = {};
  {
    int32_t m = 1;
    for (int d = ndim - 2; d >= 0; --d) {
      mult[d] = m;
      m *= maximum_level[d] + 1;
    }
    for (size_t p = 0; p < prefixes.size(); ++p) {
      int64_t dense = 0;
      for (int d = 0; d + 1 < ndim; ++d)
        dense += int64_t(prefixes[p].k[d]) * mult[d];
      prefix_of[dense] = static_cast<int32_t>(p);
    }
  }
  // one device allocation holding every table
  std::vector<char> buf;
  int64_t at_prefixes, at_subs, at_records, at_jobs, at_dense;
  append_bytes(buf, prefixes, at_prefixes);
  append_bytes(buf, sub_offset, at_subs);
  append_bytes(buf, records, at_records);
  append_bytes(buf, jobs, at_jobs);
  append_bytes(buf, prefix_of, at_dense);
  c->dev_bytes = static_cast<int64_t>(buf.size());
  if (hipMalloc(&c->dev, buf.size()) != hipSuccess ||
      hipMemcpy(c->dev, buf.data(), buf.size(), hipMemcpyHostToDevice) != hipSuccess) {
    if (c->dev == nullptr)
      (void)hipGetLastError();
    destroy(c);
    return fail(PALIWA_ERR_HIP, "combine: device tables could not be allocated");
  }
  char *base = static_cast<char *>(c->dev);
  CombineArgs &a = c->args;
  a.prefixes = reinterpret_cast<PrefixDesc const *>(base + at_prefixes);
  a.sub_offset = reinterpret_cast<int64_t const *>(base + at_subs);
  a.records = reinterpret_cast<int64_t const *>(base + at_records);
  a.jobs = reinterpret_cast<WaveJob const *>(base + at_jobs);
  a.prefix_of = reinterpret_cast<int32_t const *>(base + at_dense);
  a.ndim = ndim;
  a.nprefix = static_cast<int32_t>(prefixes.size());
  for (int d = 0; d < ndim; ++d) {
    a.max_level[d] = maximum_level[d];
    a.prefix_mult[d] = mult[d];
  }
  a.njobs = static_cast<int32_t>(njobs);
  a.narrow = narrow ? 1 : 0;
  c->nblocks = static_cast<int32_t>((njobs + kWaves - 1) / kWaves);
  *plan = c;
  return PALIWA_OK;
}

int paliwa_combine_info(const paliwa_combine *plan, int32_t *subspaces,
                        int64_t *sparse_points, int64_t *terms,
                        int64_t *table_bytes) {
  if (plan == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (subspaces != nullptr)
    *subspaces = plan->nsub;
  if (sparse_points != nullptr)
    *sparse_points = plan->sparse_points;
  if (terms != nullptr)
    *terms = plan->terms;
  if (table_bytes != nullptr)
    *table_bytes = plan->dev_bytes;
  return PALIWA_OK;
}

int paliwa_combine_subspaces(const paliwa_combine *plan, int32_t *levels,
                             int64_t *offsets, int32_t *counts) {
  if (plan == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (levels != nullptr && !plan->sub_levels.empty())
    std::memcpy(levels, plan->sub_levels.data(), plan->sub_levels.size() * sizeof(int32_t));
  if (offsets != nullptr)
    std::memcpy(offsets, plan->sub_offsets.data(), plan->sub_offsets.size() * sizeof(int64_t));
  if (counts != nullptr && !plan->sub_counts.empty())
    std::memcpy(counts, plan->sub_counts.data(), plan->sub_counts.size() * sizeof(int32_t));
  return PALIWA_OK;
}

int paliwa_combine_gather(paliwa_combine *plan, void *sparse, void *stream) {
  if (plan == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (plan->nblocks == 0)
    return PALIWA_OK; // no subspace kept
  if (sparse == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "sparse is NULL");
  auto s = static_cast<hipStream_t>(stream);
  TraceRange const range("paliwa combine gather");
  if (plan->dtype == PALIWA_FLOAT64)
    gather_dispatch<double>(plan->ndim, plan->args, plan->nblocks, sparse, s);
  else
    gather_dispatch<float>(plan->ndim, plan->args, plan->nblocks, sparse, s);
  if (hipGetLastError() != hipSuccess)
    return fail(PALIWA_ERR_HIP, "combine_gather launch failed");
  return PALIWA_OK;
}

int paliwa_combine_scatter(const paliwa_combine *plan, const void *sparse,
                           void *full, const int64_t *full_extents,
                           void *stream) {
  if (plan == nullptr || full == nullptr || full_extents == nullptr ||
      (sparse == nullptr && plan->sparse_points > 0))
    return fail(PALIWA_ERR_INVALID_ARGUMENT,
                "plan, sparse, full and full_extents are required");
  int const D = plan->ndim;
  int64_t ext[kCombineDims] = {1, 1, 1, 1, 1, 1, 1, 1};
  int64_t rows = 1;
  for (int d = 0; d < D; ++d) {
    int64_t const n = int64_t(1) << plan->args.max_level[d];
    if (full_extents[d] != n && full_extents[d] != n + 1)
      return fail(PALIWA_ERR_SHAPE, "full grid extent[" + std::to_string(d) +
                                        "] is neither 2^maximum_level nor 2^maximum_level+1");
    ext[d] = full_extents[d];
    if (d + 1 < D)
      rows *= ext[d];
  }
  int64_t const chunks = (ext[D - 1] + kThreads - 1) / kThreads;
  if (rows > INT32_MAX / chunks)
    return fail(PALIWA_ERR_UNSUPPORTED, "combine: full grid too large");
  auto s = static_cast<hipStream_t>(stream);
  unsigned const blocks = static_cast<unsigned>(rows * chunks);
  FullExtents fe{};
  for (int d = 0; d < kCombineDims; ++d)
    fe.e[d] = ext[d];
  TraceRange const range("paliwa combine scatter");
  if (plan->dtype == PALIWA_FLOAT64)
    scatter_dispatch<double>(D, plan->args, blocks, sparse, full, chunks, fe, s);
  else
    scatter_dispatch<float>(D, plan->args, blocks, sparse, full, chunks, fe, s);
  if (hipGetLastError() != hipSuccess)
    return fail(PALIWA_ERR_HIP, "combine_scatter launch failed");
  return PALIWA_OK;
}

int paliwa_combine_destroy(paliwa_combine *plan) {
  destroy(plan);
  return PALIWA_OK;
}

} // extern "C"
