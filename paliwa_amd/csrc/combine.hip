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
// k_{D-1} owns rows_p = prod 2^max(k_d - 1, 0) rows of 2^K_p positions; a
// position y of a row is in heap order: y = 0 -> k_{D-1} = 0, y >= 1 ->
// k_{D-1} = floor(log2 y) + 1, m = y - 2^(k-1).  Threads walk the positions of
// all prefixes in order, so a wave past the first 64 positions of a row
// shares one subspace (one grid list, no divergence), and the subspaces of one
// row, which read interleaved elements of the same grid rows, run back to
// back in one workgroup: the strided reads of the coarse subspaces hit the
// lines the fine ones brought into L2.
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
constexpr int64_t kBlockPositions = int64_t(kThreads) * kPerThread;

struct GridDesc {
  void const *data;
  double coefficient;
  int64_t stride[kCombineDims]; // element stride per axis (row-major)
  int32_t level[kCombineDims];
};

struct PrefixDesc {
  int64_t pos_begin;  // first global position
  int32_t first_sub;  // subspace id of (p, 0); (p, k) is first_sub + k
  int32_t K;          // largest counted k_{D-1}
  int32_t k[kCombineDims];
};

struct SubDesc {
  int64_t offset; // first element in sparse storage
  int32_t list_begin;
  int32_t list_len;
};

struct CombineArgs {
  GridDesc const *grids;
  PrefixDesc const *prefixes;
  SubDesc const *subs;
  int32_t const *lists;
  int32_t const *block_prefix; // first prefix of each block, [nblocks + 1]
  int32_t const *prefix_of;    // dense prefix-level table -> prefix id or -1
  int64_t total_positions;
  int32_t ndim;
  int32_t nprefix;
  int32_t max_level[kCombineDims];
  int32_t prefix_mult[kCombineDims]; // dense index multipliers of prefix levels
};

__device__ __forceinline__ int shape_log(int k) { return k > 0 ? k - 1 : 0; }

// index along one axis of a grid at level lg of the point m of subspace level k
__device__ __forceinline__ int64_t axis_index(int k, int64_t m, int lg) {
  return k > 0 ? ((2 * m + 1) << (lg - k)) : 0;
}

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
  int32_t const b = blockIdx.x;
  int32_t const p_lo = a.block_prefix[b];
  int32_t const p_hi = a.block_prefix[b + 1];
#pragma unroll
  for (int r = 0; r < kPerThread; ++r) {
    int64_t const pos = int64_t(b) * kBlockPositions + r * kThreads + threadIdx.x;
    if (pos >= a.total_positions)
      return;
    // the prefix holding pos: the last one with pos_begin <= pos
    int32_t lo = p_lo, hi = p_hi;
    while (lo < hi) {
      int32_t const mid = (lo + hi + 1) >> 1;
      if (a.prefixes[mid].pos_begin <= pos)
        lo = mid;
      else
        hi = mid - 1;
    }
    PrefixDesc const &pd = a.prefixes[lo];
    int64_t const q = pos - pd.pos_begin;
    int const K = pd.K;
    int64_t const row = q >> K;
    int64_t const y = q & ((int64_t(1) << K) - 1);
    int const k = y == 0 ? 0 : 64 - __clzll(y);
    int64_t const m = y == 0 ? 0 : y - (int64_t(1) << (k - 1));
    SubDesc const sd = a.subs[pd.first_sub + k];
    // the row's point within the prefix levels (row-major over 2^(k_d - 1))
    int pk[D > 1 ? D - 1 : 1];
    int64_t mp[D > 1 ? D - 1 : 1];
    int64_t rest = row;
#pragma unroll
    for (int d = D - 2; d >= 0; --d) {
      pk[d] = pd.k[d];
      int const sl = shape_log(pk[d]);
      mp[d] = rest & ((int64_t(1) << sl) - 1);
      rest >>= sl;
    }
    T s = T(0);
    for (int e = 0; e < sd.list_len; ++e) {
      GridDesc const &g = a.grids[a.lists[sd.list_begin + e]];
      int64_t idx = axis_index(k, m, g.level[D - 1]) * g.stride[D - 1];
#pragma unroll
      for (int d = 0; d < D - 1; ++d)
        idx += axis_index(pk[d], mp[d], g.level[d]) * g.stride[d];
      T const v = static_cast<T const *>(g.data)[idx];
      s = Accumulate<T>::step(s, g.coefficient, v);
    }
    int64_t const width = int64_t(1) << shape_log(k);
    sparse[sd.offset + row * width + m] = s;
  }
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
    SubDesc const sd = a.subs[a.prefixes[p].first_sub + k];
    int64_t local = 0;
#pragma unroll
    for (int d = 0; d < D - 1; ++d)
      local = (local << shape_log(pk[d])) + pm[d];
    local = (local << shape_log(k)) + m;
    value = T(0) + sparse[sd.offset + local];
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
  std::vector<SubDesc> subs;
  std::vector<int32_t> lists;
  std::map<Key, int32_t> sub_id;
  int64_t offset = 0, positions = 0;
  for (auto const &[k, n] : count) {
    if (n < min_count)
      continue;
    Key prefix(k.begin(), k.end() - 1);
    bool const new_prefix =
        prefixes.empty() ||
        !std::equal(prefix.begin(), prefix.end(), prefixes.back().k);
    if (new_prefix) {
      if (k.back() != 0) {
        // a counted (p, k) implies a counted (p, 0): counts fall with k
        destroy(c);
        return fail(PALIWA_ERR_INVALID_ARGUMENT, "combine: inconsistent subspace counts");
      }
      PrefixDesc pd{};
      pd.first_sub = static_cast<int32_t>(subs.size());
      pd.K = 0;
      for (int d = 0; d + 1 < ndim; ++d)
        pd.k[d] = prefix[d];
      prefixes.push_back(pd);
    } else {
      if (k.back() != prefixes.back().K + 1) {
        destroy(c);
        return fail(PALIWA_ERR_INVALID_ARGUMENT, "combine: inconsistent subspace counts");
      }
      prefixes.back().K = k.back();
    }
    int64_t size = 1;
    for (int d = 0; d < ndim; ++d)
      size <<= (k[d] > 0 ? k[d] - 1 : 0);
    SubDesc sd{};
    sd.offset = offset;
    offset += size;
    sub_id[k] = static_cast<int32_t>(subs.size());
    subs.push_back(sd);
    c->sub_levels.insert(c->sub_levels.end(), k.begin(), k.end());
    c->sub_offsets.push_back(sd.offset);
    c->sub_counts.push_back(n);
  }
  c->sub_offsets.push_back(offset);
  c->nsub = static_cast<int32_t>(subs.size());
  c->sparse_points = offset;
  // positions: each prefix's rows of 2^K positions, one prefix after another
  for (auto &pd : prefixes) {
    int64_t rows = 1;
    for (int d = 0; d + 1 < ndim; ++d)
      rows <<= (pd.k[d] > 0 ? pd.k[d] - 1 : 0);
    pd.pos_begin = positions;
    positions += rows << pd.K;
  }
  // grid lists per subspace, in grid order
  std::vector<std::vector<int32_t>> per_sub(subs.size());
  for (int32_t g = 0; g < ngrids; ++g) {
    int32_t const *l = levels + int64_t(g) * ndim;
    Key k(ndim, 0);
    while (true) {
      auto it = sub_id.find(k);
      if (it != sub_id.end())
        per_sub[it->second].push_back(g);
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
  for (size_t s = 0; s < subs.size(); ++s) {
    subs[s].list_begin = static_cast<int32_t>(lists.size());
    subs[s].list_len = static_cast<int32_t>(per_sub[s].size());
    lists.insert(lists.end(), per_sub[s].begin(), per_sub[s].end());
    int64_t const size = c->sub_offsets[s + 1] - c->sub_offsets[s];
    c->terms += size * subs[s].list_len;
  }
  if (lists.size() > size_t(INT32_MAX)) {
    destroy(c);
    return fail(PALIWA_ERR_UNSUPPORTED, "combine: too many grid-subspace pairs");
  }
  // grids
  std::vector<GridDesc> grids(ngrids);
  for (int32_t g = 0; g < ngrids; ++g) {
    GridDesc gd{};
    gd.data = data[g];
    gd.coefficient = coefficients[g];
    int64_t stride = 1;
    for (int d = ndim - 1; d >= 0; --d) {
      gd.stride[d] = stride;
      stride *= extents[int64_t(g) * ndim + d];
      gd.level[d] = levels[int64_t(g) * ndim + d];
    }
    grids[g] = gd;
  }
  // blocks: first prefix of each block of positions
  int64_t const nblocks = (positions + kBlockPositions - 1) / kBlockPositions;
  if (nblocks > INT32_MAX) {
    destroy(c);
    return fail(PALIWA_ERR_UNSUPPORTED, "combine: sparse grid too large");
  }
  std::vector<int32_t> block_prefix(nblocks + 1);
  {
    size_t p = 0;
    for (int64_t b = 0; b <= nblocks; ++b) {
      int64_t const start = std::min(b * kBlockPositions, positions > 0 ? positions - 1 : 0);
      while (p + 1 < prefixes.size() && prefixes[p + 1].pos_begin <= start)
        ++p;
      block_prefix[b] = static_cast<int32_t>(p);
    }
  }
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
  int64_t at_grids, at_prefixes, at_subs, at_lists, at_blocks, at_dense;
  append_bytes(buf, grids, at_grids);
  append_bytes(buf, prefixes, at_prefixes);
  append_bytes(buf, subs, at_subs);
  append_bytes(buf, lists, at_lists);
  append_bytes(buf, block_prefix, at_blocks);
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
  a.grids = reinterpret_cast<GridDesc const *>(base + at_grids);
  a.prefixes = reinterpret_cast<PrefixDesc const *>(base + at_prefixes);
  a.subs = reinterpret_cast<SubDesc const *>(base + at_subs);
  a.lists = reinterpret_cast<int32_t const *>(base + at_lists);
  a.block_prefix = reinterpret_cast<int32_t const *>(base + at_blocks);
  a.prefix_of = reinterpret_cast<int32_t const *>(base + at_dense);
  a.total_positions = positions;
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
