// Shared host helpers of the C ABI (common.cpp).
#pragma once

#include "lifting.hpp"

#include <rocprofiler-sdk-roctx/roctx.h>

#include <string>

namespace paliwa_hip {

// A roctx range (rocprofv3 --marker-trace shows it on the host timeline
// with the kernels launched inside it) for the lifetime of the object.
struct TraceRange {
  explicit TraceRange(char const *name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(TraceRange const &) = delete;
  TraceRange &operator=(TraceRange const &) = delete;
};
// "paliwa hierarchize axis d" / "paliwa dehierarchize axis d" (d < 26).
char const *axis_range_name(bool forward, int dim);

// Message of the last failed call on this thread.
std::string &last_error();
// Records `message` as the last error and returns `code`.
int fail(int code, std::string message);

// One named wavelet of the reference's tables (paliwa/paliwa_wavelets.hpp).
struct WaveletEntry {
  char const *name;
  int nforward;
  paliwa_filter forward[2];
  int nreconstruct;
  paliwa_filter reconstruct[2];
};
WaveletEntry const *find_wavelet(char const *name);

// Checks dtype, rank, levels and extents of a grid; sets *elements.
int check_grid(void const *data, int32_t dtype, int32_t ndim,
               int64_t const *extents, int32_t const *level,
               int32_t const *maximum_level, int64_t *elements);
int check_filters(paliwa_filter const *filters, int32_t nfilters);
// 0 <= minimum_level <= level <= maximum_level on every axis.
int check_level_range(int32_t ndim, int32_t const *level,
                      int32_t const *minimum_level,
                      int32_t const *maximum_level);
// The [A][m][B] view of axis `dim` of a dense row-major grid.
SweepParams make_params(void *data, int32_t ndim, int64_t const *extents,
                        int32_t const *level, int32_t dim);
// The schedule PALIWA_PATH_AUTO takes for a classified sweep (PALIWA_PATH_*).
int32_t auto_path(int32_t dtype, SweepParams const &p);
// One axis sweep through the chosen kernel schedule (PALIWA_PATH_*).
int sweep(int32_t dtype, SweepParams p, int32_t path, hipStream_t stream);
void set_filters(SweepParams &p, paliwa_filter const *filters,
                 int32_t nfilters);
// The exact dependency planner of a partial-output transform (plan.cpp).
int plan_transform_bounds(paliwa_periodic_bounds outputs, int32_t level,
                          int32_t maximum_level, int32_t const *levels,
                          int32_t nlevels, paliwa_filter const *filters,
                          int32_t nfilters, paliwa_periodic_bounds *writes,
                          paliwa_periodic_bounds *required);
// Level list of hierarchize_in (decreasing) or dehierarchize_in (increasing).
void set_named_levels(SweepParams &p, int32_t level, int32_t lmin,
                      bool forward);

} // namespace paliwa_hip
