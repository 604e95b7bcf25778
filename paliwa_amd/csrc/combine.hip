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
// component grid containing (p, 0) stores them in one of its rows.  A
// workgroup covers 1024 positions of one prefix: R = 1024 / 2^K_p whole rows
// (grouped by subspace, rows outer, m inner) or, for longer rows, 1024
// positions of one row in heap order (y = 0 -> k = 0; y >= 1 -> k =
// floor(log2 y) + 1, m = y - 2^(k-1)).  It loops over the grids containing
// (p, 0), in grid order, with the grid's descriptor in scalar registers; in
// each iteration every thread whose 4 points lie in the grid (k <= l_{D-1})
// loads them and accumulates.  So the fold order per point is the grid order,
// the loads of one iteration cover the block's grid rows together (the
// strided subsets of the subspaces complete each other's cache lines), four
// loads per thread are in flight, and the writes are contiguous runs.
#include "common.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <new>
#include <vector>

namespace {

using namespace paliwa_hip;

constexpr int kCombineDims = PALIWA_COMBINE_MAX_DIMS;
constexpr int kThreads = 256;
constexpr int kPerThread = 4;
constexpr int kLogBlock = 10;
constexpr int64_t kBlockPositions = int64_t(1) << kLogBlock;
static_assert(kThreads * kPerThread == kBlockPositions, "block size");

struct PrefixDesc {
  int64_t rows;       // points of the prefix subspace
  int32_t first_sub;  // subspace id of (p, 0); (p, k) is first_sub + k
  int32_t K;          // largest counted k_{D-1}
  int32_t k[kCombineDims];
};

struct BlockJob {
  int32_t prefix;
  int32_t local;      // block index within the prefix
  int32_t list_begin; // first term record of the block
  int32_t list_len;
};

// A term record: one grid holding a prefix's points, D + 2 words
// {data pointer, coefficient bits, last-axis level, mult_0..mult_{D-2}}; the
// grid element of a point is sum_d (2 m_d + 1) mult_d over the prefix axes
// (mult_d = stride_d << (l_d - k_d), 0 where k_d = 0) plus, along the last
// axis, (2 m + 1) << (l - k) for k >= 1 and 0 for k = 0.
struct CombineArgs {
  PrefixDesc const *prefixes;
  int64_t const *sub_offset; // first sparse element of each subspace
  int64_t const *records;      // per prefix (and smallest last level), grid order
  BlockJob const *jobs;        // [nblocks]
  int32_t const *prefix_of;    // dense prefix-level table -> prefix id or -1
  int32_t ndim;
  int32_t nprefix;
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

template <typename T, int D>
__global__ void __launch_bounds__(kThreads)
    combine_gather(CombineArgs a, T *__restrict__ sparse) {
  constexpr int P = D > 1 ? D - 1 : 1; // prefix axes (at least one slot)
  BlockJob const job = a.jobs[blockIdx.x];
  PrefixDesc const &pd = a.prefixes[job.prefix];
  int const K = pd.K;
  int pk[P];
#pragma unroll
  for (int d = 0; d < P; ++d)
    pk[d] = d < D - 1 ? pd.k[d] : 0;
  // this thread's points: subspace level k along the last axis, odd
  // multipliers 2 m_d + 1 along the prefix axes, output index
  int k[kPerThread];
  int64_t col_odd[kPerThread];
  int64_t o[kPerThread][P];
  int64_t out[kPerThread];
#pragma unroll
  for (int r = 0; r < kPerThread; ++r) {
    int const t = r * kThreads + threadIdx.x;
    int64_t row, m;
    if (K <= kLogBlock) {
      int const logR = kLogBlock - K;
      int const u = t >> logR;
      k[r] = u == 0 ? 0 : 32 - __clz(u);
      int const logw = k[r] == 0 ? 0 : k[r] - 1;
      int const within = t - (k[r] == 0 ? 0 : 1 << (logR + k[r] - 1));
      row = (int64_t(job.local) << logR) + (within >> logw);
      m = within & ((1 << logw) - 1);
      if (row >= pd.rows)
        k[r] = 64; // no point: never inside a grid
    } else {
      int const split = K - kLogBlock; // a row spans 2^split workgroups
      row = int64_t(job.local) >> split;
      int64_t const y =
          ((int64_t(job.local) & ((int64_t(1) << split) - 1)) << kLogBlock) + t;
      k[r] = y == 0 ? 0 : 64 - __clzll(y);
      m = y == 0 ? 0 : y - (int64_t(1) << (k[r] - 1));
    }
    col_odd[r] = 2 * m + 1;
    int64_t rest = row;
#pragma unroll
    for (int d = D - 2; d >= 0; --d) {
      int const sl = shape_log(pk[d]);
      o[r][d] = 2 * (rest & ((int64_t(1) << sl) - 1)) + 1;
      rest >>= sl;
    }
    out[r] = k[r] <= K ? a.sub_offset[pd.first_sub + k[r]] + (row << shape_log(k[r])) + m : -1;
  }
  constexpr int W = D + 2;
  constexpr int U = 4; // grids per step: U * kPerThread loads in flight
  T s[kPerThread];
#pragma unroll
  for (int r = 0; r < kPerThread; ++r)
    s[r] = T(0);
  int64_t const *rec = a.records + int64_t(job.list_begin) * W;
  int const len = job.list_len;
  int e = 0;
  for (; e + U <= len; e += U) {
    T v[U][kPerThread];
    int lg[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      int64_t const *w = rec + int64_t(e + u) * W;
      T const *data = reinterpret_cast<T const *>(w[0]);
      lg[u] = static_cast<int>(w[2]);
#pragma unroll
      for (int r = 0; r < kPerThread; ++r) {
        if (k[r] <= lg[u]) {
          int64_t idx = k[r] > 0 ? col_odd[r] << (lg[u] - k[r]) : 0;
#pragma unroll
          for (int d = 0; d < D - 1; ++d)
            idx += o[r][d] * w[3 + d];
          v[u][r] = data[idx];
        }
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      double const c = __longlong_as_double(rec[int64_t(e + u) * W + 1]);
#pragma unroll
      for (int r = 0; r < kPerThread; ++r)
        if (k[r] <= lg[u])
          s[r] = Accumulate<T>::step(s[r], c, v[u][r]);
    }
  }
  for (; e < len; ++e) {
    int64_t const *w = rec + int64_t(e) * W;
    T const *data = reinterpret_cast<T const *>(w[0]);
    double const c = __longlong_as_double(w[1]);
    int const lgl = static_cast<int>(w[2]);
#pragma unroll
    for (int r = 0; r < kPerThread; ++r) {
      if (k[r] <= lgl) {
        int64_t idx = k[r] > 0 ? col_odd[r] << (lgl - k[r]) : 0;
#pragma unroll
        for (int d = 0; d < D - 1; ++d)
          idx += o[r][d] * w[3 + d];
        s[r] = Accumulate<T>::step(s[r], c, data[idx]);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kPerThread; ++r)
    if (out[r] >= 0)
      sparse[out[r]] = s[r];
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
  combine_gather<T, D><<<nblocks, kThreads, 0, s>>>(a, static_cast<T *>(sparse));
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
  // workgroups: per prefix, 1024 positions each (whole rows when a row has
  // at most 1024 points), each with the grids holding its points, in grid
  // order: the grids containing (p, 0) whose last level reaches the
  // workgroup's smallest last-axis subspace level
  std::vector<BlockJob> jobs;
  std::vector<int64_t> records;
  int64_t const W = ndim + 2;
  std::vector<int32_t> holders;
  std::vector<int64_t> grid_stride(int64_t(ngrids) * ndim);
  for (int32_t g = 0; g < ngrids; ++g) {
    int64_t stride = 1;
    for (int d = ndim - 1; d >= 0; --d) {
      grid_stride[int64_t(g) * ndim + d] = stride;
      stride *= extents[int64_t(g) * ndim + d];
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
    int64_t const nb = pd.K <= kLogBlock
                           ? (rows + (int64_t(1) << (kLogBlock - pd.K)) - 1) >> (kLogBlock - pd.K)
                           : rows << (pd.K - kLogBlock);
    if (int64_t(jobs.size()) + nb > INT32_MAX) {
      destroy(c);
      return fail(PALIWA_ERR_UNSUPPORTED, "combine: sparse grid too large");
    }
    int32_t whole_begin = -1; // list shared by the workgroups starting a row
    for (int64_t b = 0; b < nb; ++b) {
      int kmin = 0;
      if (pd.K > kLogBlock) {
        int64_t const y0 = (b & ((int64_t(1) << (pd.K - kLogBlock)) - 1)) << kLogBlock;
        kmin = y0 == 0 ? 0 : 64 - __builtin_clzll(static_cast<unsigned long long>(y0));
      }
      BlockJob job{static_cast<int32_t>(p), static_cast<int32_t>(b), 0, 0};
      if (kmin == 0 && whole_begin >= 0) {
        job.list_begin = whole_begin;
        job.list_len = static_cast<int32_t>(holders.size());
      } else {
        if (int64_t(records.size()) / W + int64_t(holders.size()) > INT32_MAX) {
          destroy(c);
          return fail(PALIWA_ERR_UNSUPPORTED, "combine: too many grid lists");
        }
        job.list_begin = static_cast<int32_t>(records.size() / W);
        for (int32_t g : holders) {
          int32_t const *l = levels + int64_t(g) * ndim;
          if (l[ndim - 1] < kmin)
            continue;
          int64_t word = 0;
          void const *ptr = data[g];
          std::memcpy(&word, &ptr, sizeof(ptr));
          records.push_back(word);
          std::memcpy(&word, coefficients + g, sizeof(double));
          records.push_back(word);
          records.push_back(l[ndim - 1]);
          for (int d = 0; d + 1 < ndim; ++d)
            records.push_back(pd.k[d] > 0 ? grid_stride[int64_t(g) * ndim + d] << (l[d] - pd.k[d]) : 0);
        }
        job.list_len = static_cast<int32_t>(records.size() / W) - job.list_begin;
        if (kmin == 0)
          whole_begin = job.list_begin;
      }
      jobs.push_back(job);
    }
  }
  int64_t const nblocks = static_cast<int64_t>(jobs.size());
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
  a.jobs = reinterpret_cast<BlockJob const *>(base + at_jobs);
  a.prefix_of = reinterpret_cast<int32_t const *>(base + at_dense);
  a.ndim = ndim;
  a.nprefix = static_cast<int32_t>(prefixes.size());
  for (int d = 0; d < ndim; ++d) {
    a.max_level[d] = maximum_level[d];
    a.prefix_mult[d] = mult[d];
  }
  c->nblocks = static_cast<int32_t>(nblocks);
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
