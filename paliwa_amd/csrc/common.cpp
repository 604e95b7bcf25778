// Shared host helpers of the C ABI: error state, the filter tables, argument
// validation and the per-axis dispatch onto the kernels.
#include "common.hpp"

#include <cstring>
#include <limits>
#include <string>

namespace paliwa_hip {

char const *axis_range_name(bool forward, int dim) {
  static std::string const *names = [] {
    auto *n = new std::string[2 * PALIWA_MAX_DIMS];
    for (int d = 0; d < PALIWA_MAX_DIMS; ++d) {
      n[d] = "paliwa dehierarchize axis " + std::to_string(d);
      n[PALIWA_MAX_DIMS + d] = "paliwa hierarchize axis " + std::to_string(d);
    }
    return n;
  }();
  if (dim < 0 || dim >= PALIWA_MAX_DIMS)
    return "paliwa sweep";
  return names[(forward ? PALIWA_MAX_DIMS : 0) + dim].c_str();
}

std::string &last_error() {
  static thread_local std::string message;
  return message;
}

int fail(int code, std::string message) {
  last_error() = std::move(message);
  return code;
}

// paliwa/paliwa_wavelets.hpp:15-28 -- forward and reconstruct lifting passes.
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
    if (level[d] > PALIWA_MAX_AXIS_LEVEL)
      return fail(PALIWA_ERR_UNSUPPORTED,
                  "level[" + std::to_string(d) +
                      "] above 30: extents of 2^31 points or more are not supported");
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

int32_t auto_path(int32_t dtype, SweepParams const &p) {
  int64_t const n = int64_t(1) << p.L;
  if (paliwa_hip::long_columns_preferred(dtype, p))
    return PALIWA_PATH_LONG_LADDER; // column tiles
  if (n <= paliwa_hip::fused_max_period(dtype))
    return PALIWA_PATH_FUSED;
  if (paliwa_hip::long_ladder_preferred(dtype, p))
    return PALIWA_PATH_LONG_LADDER;
  if (n <= paliwa_hip::lds_line_max_period(dtype))
    return PALIWA_PATH_LDS_LINE;
  return PALIWA_PATH_PER_PASS;
}

// One axis sweep through the chosen kernel schedule.
int sweep(int32_t dtype, SweepParams p, int32_t path, hipStream_t stream) {
  if (p.nlevels == 0 || p.npass == 0)
    return PALIWA_OK;
  paliwa_hip::classify(p);
  int64_t const n = int64_t(1) << p.L;
  if (path == PALIWA_PATH_AUTO)
    path = auto_path(dtype, p);
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
  case PALIWA_PATH_LONG_LADDER:
    if (!paliwa_hip::long_ladder_fits(dtype, p))
      return fail(PALIWA_ERR_UNSUPPORTED,
                  "the long-line kernel takes lines above the fused kernel's periods "
                  "(and strided float64 lines of 256 points) and consecutive level "
                  "lists with the named tables' parities");
    err = paliwa_hip::launch_long_ladder(dtype, p, stream);
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

} // namespace paliwa_hip
