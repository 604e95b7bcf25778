// Partial buffers: the CheckBounds branch of the reference's lifting pass
// (paliwa/paliwa_transform.hpp:24-51 with CheckBounds = true), which its
// transform_in takes for every buffer that is not the complete strided
// lattice (:224-228, 241-248).
//
// A buffer is a dense row-major array over the positions of each axis; axis d
// describes the finest-level coordinates of its positions (paliwa_axis_storage:
// front + p * step, or an ascending list in device memory).  A pass along
// `dim` at level l (spacing t = 2^(maximum_level[dim] - l), period P =
// 2^maximum_level[dim]) updates the stored points of its half lattice --
// coordinate offset*t mod 2t along dim, a multiple of
// 2^(maximum_level[d] - level[d]) along every other axis d -- whose two
// neighbours x -+ t (periodic wrap) are stored as well; every other point
// keeps its value.  Within a pass the written points and the points read have
// opposite parity, so the order of visits does not matter and the results are
// the reference's bit for bit (same expression, same order, no contraction).
//
// Kernels:
//   checked_tile_sweep  a workgroup stages a tile of up to 64 whole lines in
//                       LDS (lines of at most 64 KiB; an explicit coordinate
//                       list of at most 8192 entries is staged beside it), runs
//                       every level and pass there and writes the tile back:
//                       one HBM read and write per sweep;
//   checked_pass        one launch per (level, pass) straight on HBM, for
//                       longer lines.
// A buffer that is the complete lattice of every axis takes the ordinary
// sweep kernels, as the reference takes its unchecked branch.
#include "common.hpp"
#include "lifting.hpp"

#include <algorithm>
#include <string>

namespace {

using namespace paliwa_hip;

constexpr int kThreads = 256;
constexpr int64_t kMaxTileLines = 64;
constexpr int64_t kTileValueBytes = 64 * 1024;
constexpr int64_t kMaxStagedList = 8192;
constexpr int64_t kMaxAxisCount = (int64_t(1) << 30) + 1;

struct AxisDesc {
  int64_t count;
  int64_t front;
  int64_t step;
  int64_t const *coords; // device; nullptr: affine
  int64_t lattice_mask;  // 2^(maximum_level - level) - 1
};

struct CheckedParams {
  void *data;
  int64_t A, n, B; // the [A][n][B] view along dim
  int64_t P;       // period along dim, in finest points
  int64_t pitch;   // LDS row pitch of a staged line (tile kernel)
  int32_t ndim, dim;
  int32_t nlevels, npass;
  int32_t W;          // lines per tile (tile kernel)
  int32_t step_shift; // affine dim axis: log2(step) when a power of two, else -1
  int8_t tshift[PALIWA_MAX_LEVELS]; // log2 of each level's spacing t
  int8_t offset[PALIWA_MAX_FILTERS];
  double coef[PALIWA_MAX_FILTERS][3];
  AxisDesc axis[PALIWA_MAX_DIMS];
};

template <typename T>
__device__ __forceinline__ T lift(T c0, T lo, T c1, T x, T c2, T hi) {
  return c0 * lo + c1 * x + c2 * hi;
}

__device__ __forceinline__ int64_t coordinate(AxisDesc const &ax, int64_t p) {
  return ax.coords != nullptr ? ax.coords[p] : ax.front + p * ax.step;
}

// Whether every axis other than dim puts line `ln` on its operating lattice.
__device__ bool line_on_lattice(CheckedParams const &q, int64_t ln) {
  int64_t a = ln / q.B;
  int64_t b = ln - a * q.B;
  for (int d = q.ndim - 1; d > q.dim; --d) {
    AxisDesc const &ax = q.axis[d];
    int64_t const next = b / ax.count;
    if (coordinate(ax, b - next * ax.count) & ax.lattice_mask)
      return false;
    b = next;
  }
  for (int d = q.dim - 1; d >= 0; --d) {
    AxisDesc const &ax = q.axis[d];
    int64_t const next = a / ax.count;
    if (coordinate(ax, a - next * ax.count) & ax.lattice_mask)
      return false;
    a = next;
  }
  return true;
}

// Position along dim of coordinate x, or -1 when it is not stored.
__device__ __forceinline__ int64_t affine_position(CheckedParams const &q, int64_t x) {
  AxisDesc const &ax = q.axis[q.dim];
  int64_t const r = x - ax.front;
  if (r < 0)
    return -1;
  int64_t k;
  if (q.step_shift >= 0) {
    if (r & ((int64_t(1) << q.step_shift) - 1))
      return -1;
    k = r >> q.step_shift;
  } else {
    k = r / ax.step;
    if (k * ax.step != r)
      return -1;
  }
  return k < q.n ? k : -1;
}

__device__ __forceinline__ int64_t list_position(int64_t const *c, int64_t n, int64_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t const mid = (lo + hi) >> 1;
    if (c[mid] < x)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo < n && c[lo] == x ? lo : -1;
}

template <typename T, bool LIST>
__global__ __launch_bounds__(kThreads) void checked_tile_sweep(CheckedParams q) {
  extern __shared__ unsigned char smem_raw[];
  int64_t *const coords = reinterpret_cast<int64_t *>(smem_raw); // LIST: n entries
  T *const vals = reinterpret_cast<T *>(smem_raw + (LIST ? q.n * 8 : 0));
  int *const active = reinterpret_cast<int *>(vals + q.W * q.pitch);
  int const tid = threadIdx.x;
  int64_t const n = q.n;
  int64_t const lines = q.A * q.B;
  int64_t const tiles = (lines + q.W - 1) / q.W;
  if constexpr (LIST) {
    for (int64_t p = tid; p < n; p += kThreads)
      coords[p] = q.axis[q.dim].coords[p];
  }
  T *const g = static_cast<T *>(q.data);
  // lanes across lines (consecutive columns) along a strided axis, along the
  // line on the contiguous one: both read consecutive addresses
  bool const across = q.B > 1;
  for (int64_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    int64_t const first = tile * q.W;
    int64_t const wn = min(int64_t(q.W), lines - first);
    for (int64_t i = tid; i < wn * n; i += kThreads) {
      int64_t const w = across ? i % wn : i / n;
      int64_t const p = across ? i / wn : i - w * n;
      int64_t const ln = first + w;
      int64_t const a = ln / q.B;
      vals[w * q.pitch + p] = g[(a * n + p) * q.B + (ln - a * q.B)];
    }
    for (int64_t w = tid; w < wn; w += kThreads)
      active[w] = line_on_lattice(q, first + w);
    __syncthreads();
    for (int l = 0; l < q.nlevels; ++l) {
      int64_t const t = int64_t(1) << q.tshift[l];
      for (int f = 0; f < q.npass; ++f) {
        int64_t const parity = q.offset[f] * t;
        T const c0 = static_cast<T>(q.coef[f][0]);
        T const c1 = static_cast<T>(q.coef[f][1]);
        T const c2 = static_cast<T>(q.coef[f][2]);
        for (int64_t i = tid; i < wn * n; i += kThreads) {
          int64_t const p = i / wn;
          int64_t const w = i - p * wn;
          if (!active[w])
            continue;
          int64_t const x = LIST ? coords[p] : q.axis[q.dim].front + p * q.axis[q.dim].step;
          if ((x & (2 * t - 1)) != parity)
            continue;
          int64_t const xl = x == 0 ? q.P - t : x - t;
          int64_t const xh = x + t == q.P ? 0 : x + t;
          int64_t const pl = LIST ? list_position(coords, n, xl) : affine_position(q, xl);
          int64_t const ph = LIST ? list_position(coords, n, xh) : affine_position(q, xh);
          if (pl < 0 || ph < 0)
            continue;
          T *const row = vals + w * q.pitch;
          row[p] = lift(c0, row[pl], c1, row[p], c2, row[ph]);
        }
        __syncthreads();
      }
    }
    for (int64_t i = tid; i < wn * n; i += kThreads) {
      int64_t const w = across ? i % wn : i / n;
      int64_t const p = across ? i / wn : i - w * n;
      int64_t const ln = first + w;
      int64_t const a = ln / q.B;
      g[(a * n + p) * q.B + (ln - a * q.B)] = vals[w * q.pitch + p];
    }
    __syncthreads();
  }
}

template <typename T, bool LIST>
__global__ __launch_bounds__(kThreads) void checked_pass(CheckedParams q, int level_index,
                                                         int pass) {
  T *const g = static_cast<T *>(q.data);
  int64_t const lines = q.A * q.B;
  int64_t const total = lines * q.n;
  int64_t const t = int64_t(1) << q.tshift[level_index];
  int64_t const parity = q.offset[pass] * t;
  T const c0 = static_cast<T>(q.coef[pass][0]);
  T const c1 = static_cast<T>(q.coef[pass][1]);
  T const c2 = static_cast<T>(q.coef[pass][2]);
  AxisDesc const &ax = q.axis[q.dim];
  for (int64_t i = blockIdx.x * int64_t(kThreads) + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * kThreads) {
    int64_t const p = i / lines;
    int64_t const ln = i - p * lines;
    int64_t const x = coordinate(ax, p);
    if ((x & (2 * t - 1)) != parity)
      continue;
    int64_t const xl = x == 0 ? q.P - t : x - t;
    int64_t const xh = x + t == q.P ? 0 : x + t;
    int64_t const pl = LIST ? list_position(ax.coords, q.n, xl) : affine_position(q, xl);
    int64_t const ph = LIST ? list_position(ax.coords, q.n, xh) : affine_position(q, xh);
    if (pl < 0 || ph < 0 || !line_on_lattice(q, ln))
      continue;
    int64_t const a = ln / q.B;
    T *const base = g + a * q.n * q.B + (ln - a * q.B);
    base[p * q.B] = lift(c0, base[pl * q.B], c1, base[p * q.B], c2, base[ph * q.B]);
  }
}

template <typename Kernel> hipError_t allow_lds(Kernel kernel, size_t bytes) {
  if (bytes <= 65536)
    return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(bytes));
}

template <typename T, bool LIST>
hipError_t launch_checked(CheckedParams q, hipStream_t stream) {
  int64_t const lines = q.A * q.B;
  bool const tile_fits = q.n * int64_t(sizeof(T)) <= kTileValueBytes &&
                         (!LIST || q.n <= kMaxStagedList);
  if (tile_fits) {
    q.pitch = q.n | 1; // odd row pitch: lanes on neighbouring lines hit different banks
    q.W = static_cast<int32_t>(std::clamp<int64_t>(
        kTileValueBytes / (q.pitch * int64_t(sizeof(T))), 1, std::min(kMaxTileLines, lines)));
    size_t const bytes = (LIST ? q.n * 8 : 0) + size_t(q.W) * q.pitch * sizeof(T) +
                         size_t(q.W) * sizeof(int);
    hipError_t err = allow_lds(checked_tile_sweep<T, LIST>, bytes);
    if (err != hipSuccess)
      return err;
    int64_t const tiles = (lines + q.W - 1) / q.W;
    unsigned const blocks = static_cast<unsigned>(std::min<int64_t>(tiles, 2048));
    hipLaunchKernelGGL((checked_tile_sweep<T, LIST>), dim3(blocks), dim3(kThreads), bytes,
                       stream, q);
    return hipGetLastError();
  }
  int64_t const total = lines * q.n;
  unsigned const blocks =
      static_cast<unsigned>(std::min<int64_t>((total + kThreads - 1) / kThreads, 8192));
  for (int l = 0; l < q.nlevels; ++l)
    for (int f = 0; f < q.npass; ++f) {
      hipLaunchKernelGGL((checked_pass<T, LIST>), dim3(blocks), dim3(kThreads), 0, stream, q,
                         l, f);
      hipError_t const err = hipGetLastError();
      if (err != hipSuccess)
        return err;
    }
  return hipSuccess;
}

struct StorageInfo {
  int64_t elements = 0;
  bool complete = false; // every axis is its complete level lattice
};

// Validates a partial buffer; no device memory is read.
int check_storage(void const *data, int32_t dtype, int32_t ndim,
                  paliwa_axis_storage const *storage, int32_t const *level,
                  int32_t const *maximum_level, StorageInfo *info) {
  if (dtype != PALIWA_FLOAT64 && dtype != PALIWA_FLOAT32)
    return fail(PALIWA_ERR_DTYPE, "dtype must be PALIWA_FLOAT64 or PALIWA_FLOAT32");
  if (ndim < 1 || ndim > PALIWA_MAX_DIMS)
    return fail(PALIWA_ERR_UNSUPPORTED, "ndim must be in [1, 26]");
  if (storage == nullptr || level == nullptr || maximum_level == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT,
                "storage, level and maximum_level are required");
  int64_t total = 1;
  bool complete = true;
  for (int d = 0; d < ndim; ++d) {
    std::string const axis = "axis " + std::to_string(d);
    paliwa_axis_storage const &s = storage[d];
    if (level[d] < 0 || level[d] > maximum_level[d] || maximum_level[d] > 62)
      return fail(PALIWA_ERR_LEVEL, "require 0 <= level <= maximum_level <= 62 on " + axis);
    if (s.count < 0 || s.count > kMaxAxisCount)
      return fail(PALIWA_ERR_UNSUPPORTED, axis + ": count outside [0, 2^30 + 1]");
    int64_t const period = int64_t(1) << maximum_level[d];
    if (s.coordinates == nullptr) {
      if (s.step < 1 || s.front < 0)
        return fail(PALIWA_ERR_INVALID_ARGUMENT, axis + ": need step >= 1 and front >= 0");
      if (s.count > 0 &&
          (s.front >= period || (period - 1 - s.front) / s.step < s.count - 1))
        return fail(PALIWA_ERR_SHAPE, axis + ": coordinates leave [0, 2^maximum_level)");
      bool const full = s.front == 0 && s.step == (int64_t(1) << (maximum_level[d] - level[d])) &&
                        s.count == (int64_t(1) << level[d]);
      complete = complete && full;
    } else {
      complete = false;
    }
    if (s.count > 0 && total > INT64_MAX / s.count)
      return fail(PALIWA_ERR_UNSUPPORTED, "buffer too large");
    total *= s.count;
  }
  if (total > 0 && data == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "data is NULL");
  info->elements = total;
  info->complete = complete;
  return PALIWA_OK;
}

int run_checked(void *data, int32_t dtype, int32_t ndim, paliwa_axis_storage const *storage,
                int32_t const *level, int32_t const *maximum_level, int32_t dim,
                int32_t const *levels, int32_t nlevels, paliwa_filter const *filters,
                int32_t nfilters, StorageInfo const &info, hipStream_t stream) {
  if (info.elements == 0 || nlevels == 0 || nfilters == 0)
    return PALIWA_OK;
  if (info.complete) {
    int64_t extents[PALIWA_MAX_DIMS];
    for (int d = 0; d < ndim; ++d)
      extents[d] = storage[d].count;
    SweepParams p = make_params(data, ndim, extents, level, dim);
    p.nlevels = nlevels;
    for (int i = 0; i < nlevels; ++i)
      p.shift[i] = static_cast<int8_t>(level[dim] - levels[i]);
    set_filters(p, filters, nfilters);
    return sweep(dtype, p, PALIWA_PATH_AUTO, stream);
  }
  CheckedParams q{};
  q.data = data;
  q.ndim = ndim;
  q.dim = dim;
  q.A = 1;
  q.B = 1;
  for (int d = 0; d < ndim; ++d) {
    paliwa_axis_storage const &s = storage[d];
    q.axis[d] = {s.count, s.front, s.step, s.coordinates,
                 (int64_t(1) << (maximum_level[d] - level[d])) - 1};
    if (d < dim)
      q.A *= s.count;
    else if (d > dim)
      q.B *= s.count;
  }
  q.n = storage[dim].count;
  q.P = int64_t(1) << maximum_level[dim];
  int64_t const step = storage[dim].step;
  q.step_shift = -1;
  if (storage[dim].coordinates == nullptr && (step & (step - 1)) == 0)
    for (int k = 0; k < 63; ++k)
      if ((int64_t(1) << k) == step)
        q.step_shift = k;
  q.nlevels = nlevels;
  for (int i = 0; i < nlevels; ++i)
    q.tshift[i] = static_cast<int8_t>(maximum_level[dim] - levels[i]);
  q.npass = nfilters;
  for (int f = 0; f < nfilters; ++f) {
    q.offset[f] = static_cast<int8_t>(filters[f].offset);
    for (int c = 0; c < 3; ++c)
      q.coef[f][c] = filters[f].coefficients[c];
  }
  bool const list = storage[dim].coordinates != nullptr;
  hipError_t err;
  if (dtype == PALIWA_FLOAT64)
    err = list ? launch_checked<double, true>(q, stream) : launch_checked<double, false>(q, stream);
  else
    err = list ? launch_checked<float, true>(q, stream) : launch_checked<float, false>(q, stream);
  if (err != hipSuccess)
    return fail(PALIWA_ERR_HIP, std::string("HIP: ") + hipGetErrorString(err));
  return PALIWA_OK;
}

int named_checked(void *data, int32_t dtype, int32_t ndim, paliwa_axis_storage const *storage,
                  int32_t const *level, int32_t const *minimum_level,
                  int32_t const *maximum_level, int32_t first_dim, int32_t last_dim,
                  char const *wavelet_name, bool forward, void *stream) {
  StorageInfo info;
  int rc = check_storage(data, dtype, ndim, storage, level, maximum_level, &info);
  if (rc != PALIWA_OK)
    return rc;
  if ((rc = check_level_range(ndim, level, minimum_level, maximum_level)) != PALIWA_OK)
    return rc;
  WaveletEntry const *w = find_wavelet(wavelet_name);
  if (w == nullptr) // std::map::at in paliwa/paliwa_transform.hpp:370, 397
    return fail(PALIWA_ERR_UNKNOWN_WAVELET, "map::at");
  for (int d = first_dim; d <= last_dim; ++d) {
    int32_t levels[PALIWA_MAX_LEVELS]; // level <= maximum_level <= 62
    int32_t const nlevels = level[d] - minimum_level[d];
    for (int32_t i = 0; i < nlevels; ++i)
      levels[i] = forward ? level[d] - i : minimum_level[d] + 1 + i;
    rc = run_checked(data, dtype, ndim, storage, level, maximum_level, d, levels, nlevels,
                     forward ? w->forward : w->reconstruct,
                     forward ? w->nforward : w->nreconstruct, info,
                     static_cast<hipStream_t>(stream));
    if (rc != PALIWA_OK)
      return rc;
  }
  last_error().clear();
  return PALIWA_OK;
}

} // namespace

extern "C" {

int paliwa_transform_in_checked(void *data, int32_t dtype, int32_t ndim,
                                const paliwa_axis_storage *storage, const int32_t *level,
                                const int32_t *maximum_level, int32_t dim,
                                const int32_t *levels, int32_t nlevels,
                                const paliwa_filter *filters, int32_t nfilters,
                                void *stream) {
  StorageInfo info;
  int rc = check_storage(data, dtype, ndim, storage, level, maximum_level, &info);
  if (rc != PALIWA_OK)
    return rc;
  if (dim < 0 || dim >= ndim)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "dim out of range");
  if (nlevels < 0 || nlevels > PALIWA_MAX_LEVELS)
    return fail(PALIWA_ERR_UNSUPPORTED, "at most 64 levels per call");
  if (nlevels > 0 && levels == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "levels is NULL");
  if ((rc = check_filters(filters, nfilters)) != PALIWA_OK)
    return rc;
  for (int i = 0; i < nlevels; ++i)
    if (levels[i] < 1 || levels[i] > level[dim]) // paliwa/paliwa_transform.hpp:232
      return fail(PALIWA_ERR_LEVEL, "operating level " + std::to_string(levels[i]) +
                                        " outside [1, level[dim]]");
  rc = run_checked(data, dtype, ndim, storage, level, maximum_level, dim, levels, nlevels,
                   filters, nfilters, info, static_cast<hipStream_t>(stream));
  if (rc == PALIWA_OK)
    last_error().clear();
  return rc;
}

int paliwa_hierarchize_in_checked(void *data, int32_t dtype, int32_t ndim,
                                  const paliwa_axis_storage *storage, const int32_t *level,
                                  const int32_t *minimum_level,
                                  const int32_t *maximum_level, int32_t dim,
                                  const char *wavelet_name, void *stream) {
  if (dim < 0 || dim >= ndim)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "dim out of range");
  return named_checked(data, dtype, ndim, storage, level, minimum_level, maximum_level, dim,
                       dim, wavelet_name, true, stream);
}

int paliwa_dehierarchize_in_checked(void *data, int32_t dtype, int32_t ndim,
                                    const paliwa_axis_storage *storage, const int32_t *level,
                                    const int32_t *minimum_level,
                                    const int32_t *maximum_level, int32_t dim,
                                    const char *wavelet_name, void *stream) {
  if (dim < 0 || dim >= ndim)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "dim out of range");
  return named_checked(data, dtype, ndim, storage, level, minimum_level, maximum_level, dim,
                       dim, wavelet_name, false, stream);
}

int paliwa_hierarchize_checked(void *data, int32_t dtype, int32_t ndim,
                               const paliwa_axis_storage *storage, const int32_t *level,
                               const int32_t *minimum_level, const int32_t *maximum_level,
                               const char *wavelet_name, void *stream) {
  return named_checked(data, dtype, ndim, storage, level, minimum_level, maximum_level, 0,
                       ndim - 1, wavelet_name, true, stream);
}

int paliwa_dehierarchize_checked(void *data, int32_t dtype, int32_t ndim,
                                 const paliwa_axis_storage *storage, const int32_t *level,
                                 const int32_t *minimum_level, const int32_t *maximum_level,
                                 const char *wavelet_name, void *stream) {
  return named_checked(data, dtype, ndim, storage, level, minimum_level, maximum_level, 0,
                       ndim - 1, wavelet_name, false, stream);
}

} // extern "C"
