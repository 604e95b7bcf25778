// C ABI of the lifting transform: argument validation, the filter tables and
// the per-axis dispatch onto the HIP kernels.  Declarations and the reference
// interface each entry point replaces: include/paliwa_hip.h.
#include "common.hpp"
#include "lifting.hpp"

#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

using namespace paliwa_hip;

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
    TraceRange const range(axis_range_name(forward, d));
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
  paliwa_hip::last_error().clear();
  return PALIWA_OK;
}

} // namespace

extern "C" {

char const *paliwa_last_error(void) { return paliwa_hip::last_error().c_str(); }

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
  if (path < PALIWA_PATH_AUTO || path > PALIWA_PATH_LONG_LADDER)
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
  TraceRange const range("paliwa transform_in");
  rc = sweep(dtype, p, path, static_cast<hipStream_t>(stream));
  if (rc == PALIWA_OK)
    paliwa_hip::last_error().clear();
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

int paliwa_auto_path(int32_t dtype, int32_t ndim, int64_t const *extents,
                     int32_t const *level, int32_t const *minimum_level, int32_t dim,
                     char const *wavelet_name, int32_t forward, int32_t *path) {
  if (path == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "path is NULL");
  int64_t elements = 0;
  int const unused = 0; // check_grid wants a data pointer; nothing is read
  int rc = check_grid(&unused, dtype, ndim, extents, level, nullptr, &elements);
  if (rc != PALIWA_OK)
    return rc;
  if (dim < 0 || dim >= ndim)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "dim out of range");
  if ((rc = check_level_range(ndim, level, minimum_level, level)) != PALIWA_OK)
    return rc;
  WaveletEntry const *w = find_wavelet(wavelet_name);
  if (w == nullptr)
    return fail(PALIWA_ERR_UNKNOWN_WAVELET, "map::at");
  SweepParams p = make_params(nullptr, ndim, extents, level, dim);
  if (forward)
    set_filters(p, w->forward, w->nforward);
  else
    set_filters(p, w->reconstruct, w->nreconstruct);
  set_named_levels(p, level[dim], minimum_level[dim], forward != 0);
  if (p.nlevels == 0) {
    *path = PALIWA_PATH_AUTO; // nothing to sweep
  } else {
    paliwa_hip::classify(p);
    *path = auto_path(dtype, p);
  }
  paliwa_hip::last_error().clear();
  return PALIWA_OK;
}

} // extern "C"
