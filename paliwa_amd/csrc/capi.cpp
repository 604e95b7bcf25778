// C ABI of the lifting transform: argument validation, the filter tables and
// the per-axis dispatch onto the HIP kernels.  Declarations and the reference
// interface each entry point replaces: include/paliwa_hip.h.
#include "lifting.hpp"

#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

using paliwa_hip::SweepParams;

thread_local std::string g_last_error;

int fail(int code, std::string message) {
  g_last_error = std::move(message);
  return code;
}

// paliwa/paliwa_wavelets.hpp:15-28 -- forward and reconstruct lifting passes.
struct WaveletEntry {
  char const *name;
  int nforward;
  paliwa_filter forward[2];
  int nreconstruct;
  paliwa_filter reconstruct[2];
};

constexpr WaveletEntry kWavelets[] = {
    {"hat", 1, {{1, 0, {-0.5, 1.0, -0.5}}, {}}, 1, {{1, 0, {0.5, 1.0, 0.5}}, {}}},
    {"biorthogonal",
     2,
     {{1, 0, {-0.5, 1.0, -0.5}}, {0, 0, {0.25, 1.0, 0.25}}},
     2,
     {{0, 0, {-0.25, 1.0, -0.25}}, {1, 0, {0.5, 1.0, 0.5}}}},
    {"fullweighting",
     2,
     {{0, 0, {0.25, 0.5, 0.25}}, {1, 0, {-0.5, 1.0, -0.5}}},
     2,
     {{1, 0, {0.5, 1.0, 0.5}}, {0, 0, {-0.5, 2.0, -0.5}}}},
};

WaveletEntry const *find_wavelet(char const *name) {
  if (name == nullptr)
    return nullptr;
  for (auto const &w : kWavelets)
    if (std::strcmp(w.name, name) == 0)
      return &w;
  return nullptr;
}

// Checks dtype, rank, levels and extents of a grid; sets *elements.
int check_grid(void const *data, int32_t dtype, int32_t ndim,
               int64_t const *extents, int32_t const *level,
               int32_t const *maximum_level, int64_t *elements) {
  if (dtype != PALIWA_FLOAT64 && dtype != PALIWA_FLOAT32)
    return fail(PALIWA_ERR_DTYPE, "dtype must be PALIWA_FLOAT64 or PALIWA_FLOAT32");
  if (ndim < 1 || ndim > PALIWA_MAX_DIMS)
    return fail(PALIWA_ERR_UNSUPPORTED, "ndim must be in [1, 26]");
  if (extents == nullptr || level == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "extents and level are required");
  int64_t total = 1;
  for (int d = 0; d < ndim; ++d) {
    if (level[d] < 0 || level[d] > PALIWA_MAX_LEVEL)
      return fail(PALIWA_ERR_LEVEL, "level[" + std::to_string(d) +
                                        "] out of range [0, 40]");
    if (maximum_level != nullptr && level[d] > maximum_level[d])
      return fail(PALIWA_ERR_LEVEL, "level[" + std::to_string(d) +
                                        "] exceeds maximum_level");
    int64_t const n = int64_t(1) << level[d];
    if (extents[d] != n && extents[d] != n + 1)
      return fail(PALIWA_ERR_SHAPE,
                  "extent[" + std::to_string(d) + "] = " +
                      std::to_string(extents[d]) + " is neither 2^level = " +
                      std::to_string(n) + " nor 2^level+1");
    if (total > std::numeric_limits<int64_t>::max() / extents[d])
      return fail(PALIWA_ERR_UNSUPPORTED, "grid too large");
    total *= extents[d];
  }
  if (data == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "data is NULL");
  *elements = total;
  return PALIWA_OK;
}

int check_filters(paliwa_filter const *filters, int32_t nfilters) {
  if (nfilters < 0 || nfilters > PALIWA_MAX_FILTERS)
    return fail(PALIWA_ERR_UNSUPPORTED, "at most 16 filter passes");
  if (nfilters > 0 && filters == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "filters is NULL");
  for (int f = 0; f < nfilters; ++f)
    if (filters[f].offset != 0 && filters[f].offset != 1)
      // paliwa/paliwa_transform.hpp:239-240
      return fail(PALIWA_ERR_FILTER_OFFSET, "Filter offset not supported");
  return PALIWA_OK;
}

int check_level_range(int32_t ndim, int32_t const *level,
                      int32_t const *minimum_level,
                      int32_t const *maximum_level) {
  if (minimum_level == nullptr || maximum_level == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT,
                "minimum_level and maximum_level are required");
  // paliwa/paliwa_transform.hpp:359-362 (asserted there)
  for (int d = 0; d < ndim; ++d)
    if (minimum_level[d] < 0 || minimum_level[d] > level[d] ||
        level[d] > maximum_level[d])
      return fail(PALIWA_ERR_LEVEL,
                  "require 0 <= minimum_level <= level <= maximum_level on "
                  "axis " + std::to_string(d));
  return PALIWA_OK;
}

SweepParams make_params(void *data, int32_t ndim, int64_t const *extents,
                        int32_t const *level, int32_t dim) {
  SweepParams p{};
  p.data = data;
  p.A = 1;
  p.B = 1;
  for (int d = 0; d < dim; ++d)
    p.A *= extents[d];
  for (int d = dim + 1; d < ndim; ++d)
    p.B *= extents[d];
  p.m = static_cast<int32_t>(extents[dim]);
  p.L = level[dim];
  return p;
}

// One axis sweep through the chosen kernel schedule.
int sweep(int32_t dtype, SweepParams p, int32_t path, hipStream_t stream) {
  if (p.nlevels == 0 || p.npass == 0)
    return PALIWA_OK;
  paliwa_hip::classify(p);
  int64_t const n = int64_t(1) << p.L;
  if (path == PALIWA_PATH_AUTO) {
    if (n <= paliwa_hip::fused_max_period(dtype))
      path = PALIWA_PATH_FUSED;
    else if (n <= paliwa_hip::lds_line_max_period(dtype))
      path = PALIWA_PATH_LDS_LINE;
    else
      path = PALIWA_PATH_PER_PASS;
  }
  hipError_t err;
  switch (path) {
  case PALIWA_PATH_FUSED:
    if (n > paliwa_hip::fused_max_period(dtype))
      return fail(PALIWA_ERR_UNSUPPORTED, "line too long for the fused kernel");
    err = paliwa_hip::launch_fused(dtype, p, stream);
    break;
  case PALIWA_PATH_LDS_LINE:
    if (n > paliwa_hip::lds_line_max_period(dtype))
      return fail(PALIWA_ERR_UNSUPPORTED, "line too long for the LDS-line kernel");
    err = paliwa_hip::launch_lds_line(dtype, p, stream);
    break;
  case PALIWA_PATH_PER_PASS:
    err = paliwa_hip::launch_per_pass(dtype, p, stream);
    break;
  default:
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "unknown kernel path");
  }
  if (err != hipSuccess)
    return fail(PALIWA_ERR_HIP, std::string("HIP: ") + hipGetErrorString(err));
  return PALIWA_OK;
}

void set_filters(SweepParams &p, paliwa_filter const *filters,
                 int32_t nfilters) {
  p.npass = nfilters;
  for (int f = 0; f < nfilters; ++f) {
    p.offset[f] = static_cast<int8_t>(filters[f].offset);
    for (int c = 0; c < 3; ++c)
      p.coef[f][c] = filters[f].coefficients[c];
  }
}

// Level list of hierarchize_in (decreasing) or dehierarchize_in (increasing).
void set_named_levels(SweepParams &p, int32_t level, int32_t lmin,
                      bool forward) {
  p.nlevels = level - lmin;
  for (int i = 0; i < p.nlevels; ++i) {
    int const l = forward ? level - i : lmin + 1 + i;
    p.shift[i] = static_cast<int8_t>(level - l);
  }
}

int named_transform(void *data, int32_t dtype, int32_t ndim,
                    int64_t const *extents, int32_t const *level,
                    int32_t const *minimum_level, int32_t const *maximum_level,
                    int32_t first_dim, int32_t last_dim,
                    char const *wavelet_name, bool forward, void *stream) {
  int64_t elements = 0;
  int rc = check_grid(data, dtype, ndim, extents, level, maximum_level,
                      &elements);
  if (rc != PALIWA_OK)
    return rc;
  if ((rc = check_level_range(ndim, level, minimum_level, maximum_level)) !=
      PALIWA_OK)
    return rc;
  WaveletEntry const *w = find_wavelet(wavelet_name);
  if (w == nullptr)
    // std::map::at in paliwa/paliwa_transform.hpp:370, 397
    return fail(PALIWA_ERR_UNKNOWN_WAVELET, "map::at");
  for (int d = first_dim; d <= last_dim; ++d) {
    SweepParams p = make_params(data, ndim, extents, level, d);
    if (forward)
      set_filters(p, w->forward, w->nforward);
    else
      set_filters(p, w->reconstruct, w->nreconstruct);
    set_named_levels(p, level[d], minimum_level[d], forward);
    rc = sweep(dtype, p, PALIWA_PATH_AUTO, static_cast<hipStream_t>(stream));
    if (rc != PALIWA_OK)
      return rc;
  }
  g_last_error.clear();
  return PALIWA_OK;
}

} // namespace

extern "C" {

char const *paliwa_last_error(void) { return g_last_error.c_str(); }

char const *paliwa_version(void) { return "paliwa-hip 0.1.0 (gfx950)"; }

int paliwa_wavelet_filters(char const *wavelet_name, int32_t reconstruct,
                           paliwa_filter *out, int32_t capacity,
                           int32_t *count) {
  WaveletEntry const *w = find_wavelet(wavelet_name);
  if (w == nullptr)
    return fail(PALIWA_ERR_UNKNOWN_WAVELET, "map::at");
  int const n = reconstruct ? w->nreconstruct : w->nforward;
  paliwa_filter const *src = reconstruct ? w->reconstruct : w->forward;
  if (count != nullptr)
    *count = n;
  if (capacity < 0 || (capacity > 0 && out == nullptr))
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "bad output buffer");
  for (int i = 0; i < n && i < capacity; ++i)
    out[i] = src[i];
  return PALIWA_OK;
}

int paliwa_transform_in_path(void *data, int32_t dtype, int32_t ndim,
                             int64_t const *extents, int32_t const *level,
                             int32_t const *maximum_level, int32_t dim,
                             int32_t const *levels, int32_t nlevels,
                             paliwa_filter const *filters, int32_t nfilters,
                             int32_t path, void *stream) {
  int64_t elements = 0;
  int rc = check_grid(data, dtype, ndim, extents, level, maximum_level,
                      &elements);
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
  if (path < PALIWA_PATH_AUTO || path > PALIWA_PATH_PER_PASS)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "unknown kernel path");
  SweepParams p = make_params(data, ndim, extents, level, dim);
  p.nlevels = nlevels;
  for (int i = 0; i < nlevels; ++i) {
    // paliwa/paliwa_transform.hpp:232 asserts l > 0; a level above the
    // grid's would address points outside the storage.
    if (levels[i] < 1 || levels[i] > level[dim])
      return fail(PALIWA_ERR_LEVEL, "operating level " +
                                        std::to_string(levels[i]) +
                                        " outside [1, level[dim]]");
    p.shift[i] = static_cast<int8_t>(level[dim] - levels[i]);
  }
  set_filters(p, filters, nfilters);
  rc = sweep(dtype, p, path, static_cast<hipStream_t>(stream));
  if (rc == PALIWA_OK)
    g_last_error.clear();
  return rc;
}

int paliwa_transform_in(void *data, int32_t dtype, int32_t ndim,
                        int64_t const *extents, int32_t const *level,
                        int32_t const *maximum_level, int32_t dim,
                        int32_t const *levels, int32_t nlevels,
                        paliwa_filter const *filters, int32_t nfilters,
                        void *stream) {
  return paliwa_transform_in_path(data, dtype, ndim, extents, level,
                                  maximum_level, dim, levels, nlevels, filters,
                                  nfilters, PALIWA_PATH_AUTO, stream);
}

int paliwa_hierarchize_in(void *data, int32_t dtype, int32_t ndim,
                          int64_t const *extents, int32_t const *level,
                          int32_t const *minimum_level,
                          int32_t const *maximum_level, int32_t dim,
                          char const *wavelet_name, void *stream) {
  if (dim < 0 || dim >= ndim)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "dim out of range");
  return named_transform(data, dtype, ndim, extents, level, minimum_level,
                         maximum_level, dim, dim, wavelet_name, true, stream);
}

int paliwa_dehierarchize_in(void *data, int32_t dtype, int32_t ndim,
                            int64_t const *extents, int32_t const *level,
                            int32_t const *minimum_level,
                            int32_t const *maximum_level, int32_t dim,
                            char const *wavelet_name, void *stream) {
  if (dim < 0 || dim >= ndim)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "dim out of range");
  return named_transform(data, dtype, ndim, extents, level, minimum_level,
                         maximum_level, dim, dim, wavelet_name, false, stream);
}

int paliwa_hierarchize(void *data, int32_t dtype, int32_t ndim,
                       int64_t const *extents, int32_t const *level,
                       int32_t const *minimum_level,
                       int32_t const *maximum_level, char const *wavelet_name,
                       void *stream) {
  return named_transform(data, dtype, ndim, extents, level, minimum_level,
                         maximum_level, 0, ndim - 1, wavelet_name, true,
                         stream);
}

int paliwa_dehierarchize(void *data, int32_t dtype, int32_t ndim,
                         int64_t const *extents, int32_t const *level,
                         int32_t const *minimum_level,
                         int32_t const *maximum_level,
                         char const *wavelet_name, void *stream) {
  return named_transform(data, dtype, ndim, extents, level, minimum_level,
                         maximum_level, 0, ndim - 1, wavelet_name, false,
                         stream);
}

} // extern "C"
