// C entry points of the CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// Loaded with ctypes by tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg.  Never linked into, or called by, the HIP product library.
//
// Schedules:
//   0  serial direct  -- one sweep per (axis, level, pass), the order in which
//                        paliwa::hierarchize executes (paliwa/paliwa_transform
//                        .hpp:24-51, 215-251); this is the parity oracle.
//   1  OpenMP direct  -- the same sweeps, lines split across threads.
//   2  OpenMP packed poles -- per line: gather, every level and pass, scatter
//                        (PoleExecution::PackedPoles, paliwa/paliwa_transform
//                        .hpp:119-184); the reference's fastest CPU schedule.
// All three give bit-identical results: within one pass the written points
// and the points read have opposite parity, so no pass depends on the order
// in which its points are visited.
#include "paliwa_oracle.hpp"

#include <cstring>
#include <exception>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string last_error;

int fail(std::exception const &e, int code) {
  last_error = e.what();
  return code;
}

using paliwa_oracle::AxisView;
using paliwa_oracle::Filters;

template <typename T>
void lifting_pass_omp(T *g, AxisView v, int64_t n, int64_t s, int offset,
                      std::array<double, 3> const &f) {
  T const f0 = static_cast<T>(f[0]);
  T const f1 = static_cast<T>(f[1]);
  T const f2 = static_cast<T>(f[2]);
  int64_t const count = (n - offset * s + 2 * s - 1) / (2 * s);
#pragma omp parallel for collapse(2) schedule(static)
  for (int64_t a = 0; a < v.A; ++a) {
    for (int64_t k = 0; k < count; ++k) {
      int64_t const x = offset * s + 2 * s * k;
      T *block = g + a * v.m * v.B;
      int64_t const lower = x == 0 ? n - s : x - s;
      int64_t const upper = x + s == n ? 0 : x + s;
      T *gi = block + x * v.B;
      T const *gl = block + lower * v.B;
      T const *gu = block + upper * v.B;
      for (int64_t b = 0; b < v.B; ++b)
        gi[b] = f0 * gl[b] + f1 * gi[b] + f2 * gu[b];
    }
  }
}

template <typename T>
void copy_image_plane_omp(T *g, AxisView v, int64_t n) {
  if (v.m != n + 1)
    return;
#pragma omp parallel for schedule(static)
  for (int64_t a = 0; a < v.A; ++a) {
    T *block = g + a * v.m * v.B;
    for (int64_t b = 0; b < v.B; ++b)
      block[n * v.B + b] = block[b];
  }
}

// Per-line schedule: gather a pole into thread-private scratch, apply every
// level and pass, scatter (paliwa/paliwa_transform.hpp:173-182).
template <typename T>
void transform_packed_poles(T *g, AxisView v, int64_t n,
                            std::vector<int64_t> const &strides,
                            Filters const &filters) {
  std::vector<std::array<T, 3>> coef;
  for (auto const &[offset, f] : filters)
    coef.push_back({static_cast<T>(f[0]), static_cast<T>(f[1]),
                    static_cast<T>(f[2])});
  int64_t const poles = v.A * v.B;
#pragma omp parallel
  {
    std::vector<T> buf(n);
#pragma omp for schedule(static)
    for (int64_t p = 0; p < poles; ++p) {
      int64_t const a = p / v.B, b = p % v.B;
      T *line = g + a * v.m * v.B + b;
      for (int64_t x = 0; x < n; ++x)
        buf[x] = line[x * v.B];
      for (int64_t s : strides) {
        for (size_t q = 0; q < filters.size(); ++q) {
          int const offset = filters[q].first;
          T const f0 = coef[q][0], f1 = coef[q][1], f2 = coef[q][2];
          for (int64_t x = offset * s; x < n; x += 2 * s) {
            int64_t const lower = x == 0 ? n - s : x - s;
            int64_t const upper = x + s == n ? 0 : x + s;
            buf[x] = f0 * buf[lower] + f1 * buf[x] + f2 * buf[upper];
          }
        }
      }
      for (int64_t x = 0; x < n; ++x)
        line[x * v.B] = buf[x];
      if (v.m == n + 1)
        line[n * v.B] = buf[0];
    }
  }
}

template <typename T>
void transform_in_scheduled(T *g, std::vector<int64_t> const &extents,
                            std::vector<int> const &level, int dim,
                            std::vector<int> const &levels,
                            Filters const &filters, int schedule) {
  if (schedule == 0) {
    paliwa_oracle::transform_in(g, extents, level, dim, levels, filters);
    return;
  }
  paliwa_oracle::check_extents(extents, level);
  AxisView const v = paliwa_oracle::axis_view(extents, dim);
  int64_t const n = int64_t(1) << level[dim];
  std::vector<int64_t> strides;
  for (int l : levels) {
    if (l < 1 || l > level[dim])
      throw std::invalid_argument("operating level out of range");
    strides.push_back(int64_t(1) << (level[dim] - l));
  }
  for (auto const &[offset, f] : filters)
    if (offset != 0 && offset != 1)
      throw std::runtime_error("Filter offset not supported");
  if (strides.empty() || filters.empty())
    return;
  if (schedule == 2) {
    transform_packed_poles(g, v, n, strides, filters);
    return;
  }
  for (int64_t s : strides)
    for (auto const &[offset, f] : filters)
      lifting_pass_omp(g, v, n, s, offset, f);
  copy_image_plane_omp(g, v, n);
}

template <typename T>
void run_named(void *data, int ndim, int64_t const *extents,
               int32_t const *level, int32_t const *lmin, int32_t const *lmax,
               char const *name, bool forward, int schedule) {
  std::vector<int64_t> ext(extents, extents + ndim);
  std::vector<int> lv(level, level + ndim), mn(lmin, lmin + ndim),
      mx(lmax, lmax + ndim);
  paliwa_oracle::check_levels(lv, mn, mx);
  Filters const &filters = forward
                               ? paliwa_oracle::forward_filters().at(name)
                               : paliwa_oracle::reconstruct_filters().at(name);
  for (int d = 0; d < ndim; ++d) {
    auto const levels =
        forward ? paliwa_oracle::decreasing_levels(lv[d], mn[d])
                : paliwa_oracle::increasing_levels(lv[d], mn[d]);
    transform_in_scheduled(static_cast<T *>(data), ext, lv, d, levels,
                           filters, schedule);
  }
}

} // namespace

extern "C" {

// dtype: 0 = float64, 1 = float32.  Returns 0 on success, a negative code and
// a message (oracle_last_error) otherwise.
int oracle_transform_in(void *data, int dtype, int ndim,
                        int64_t const *extents, int32_t const *level, int dim,
                        int32_t const *levels, int nlevels,
                        int32_t const *offsets, double const *coefficients,
                        int nfilters, int schedule) {
  try {
    std::vector<int64_t> ext(extents, extents + ndim);
    std::vector<int> lv(level, level + ndim);
    std::vector<int> ls(levels, levels + nlevels);
    Filters filters;
    for (int p = 0; p < nfilters; ++p)
      filters.push_back({offsets[p],
                         {coefficients[3 * p], coefficients[3 * p + 1],
                          coefficients[3 * p + 2]}});
    if (dim < 0 || dim >= ndim)
      throw std::invalid_argument("dim out of range");
    if (dtype == 0)
      transform_in_scheduled(static_cast<double *>(data), ext, lv, dim, ls,
                             filters, schedule);
    else if (dtype == 1)
      transform_in_scheduled(static_cast<float *>(data), ext, lv, dim, ls,
                             filters, schedule);
    else
      throw std::invalid_argument("unknown dtype");
    return 0;
  } catch (std::out_of_range const &e) {
    return fail(e, -2);
  } catch (std::runtime_error const &e) {
    return fail(e, -3);
  } catch (std::exception const &e) {
    return fail(e, -1);
  }
}

int oracle_transform(void *data, int dtype, int ndim, int64_t const *extents,
                     int32_t const *level, int32_t const *lmin,
                     int32_t const *lmax, char const *name, int forward,
                     int schedule) {
  try {
    if (dtype == 0)
      run_named<double>(data, ndim, extents, level, lmin, lmax, name,
                        forward != 0, schedule);
    else if (dtype == 1)
      run_named<float>(data, ndim, extents, level, lmin, lmax, name,
                       forward != 0, schedule);
    else
      throw std::invalid_argument("unknown dtype");
    return 0;
  } catch (std::out_of_range const &e) {
    return fail(e, -2);
  } catch (std::runtime_error const &e) {
    return fail(e, -3);
  } catch (std::exception const &e) {
    return fail(e, -1);
  }
}

// Partial buffer: coords = the coordinates of every axis back to back
// (counts[d] of them for axis d).
int oracle_transform_in_checked(void *data, int dtype, int ndim,
                                int64_t const *counts, int64_t const *coords,
                                int32_t const *level,
                                int32_t const *maximum_level, int dim,
                                int32_t const *levels, int nlevels,
                                int32_t const *offsets,
                                double const *coefficients, int nfilters) {
  try {
    std::vector<std::vector<int64_t>> axes(ndim);
    for (int d = 0; d < ndim; ++d) {
      axes[d].assign(coords, coords + counts[d]);
      coords += counts[d];
    }
    std::vector<int> lv(level, level + ndim), mx(maximum_level, maximum_level + ndim);
    std::vector<int> ls(levels, levels + nlevels);
    Filters filters;
    for (int p = 0; p < nfilters; ++p)
      filters.push_back({offsets[p],
                         {coefficients[3 * p], coefficients[3 * p + 1],
                          coefficients[3 * p + 2]}});
    if (dtype == 0)
      paliwa_oracle::transform_in_checked(static_cast<double *>(data), axes, lv, mx,
                                          dim, ls, filters);
    else if (dtype == 1)
      paliwa_oracle::transform_in_checked(static_cast<float *>(data), axes, lv, mx,
                                          dim, ls, filters);
    else
      throw std::invalid_argument("unknown dtype");
    return 0;
  } catch (std::out_of_range const &e) {
    return fail(e, -2);
  } catch (std::runtime_error const &e) {
    return fail(e, -3);
  } catch (std::exception const &e) {
    return fail(e, -1);
  }
}

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

char const *oracle_last_error(void) { return last_error.c_str(); }

// ---- partial-output planner: brute force of the stencil graph -------------
// Restates the independent check of test/test_paliwa_plan.cpp:34-69: starting
// from the output points, walk the passes backwards; a pass at level l with
// parity `offset` writes the required points of its parity at that level and
// makes their neighbours at +-2^(maximum - l) required.  No intervals and no
// coefficients are involved.  oracle_plan_check compares a planner's periodic
// bounds (min, max pairs; min > max wraps; -1 empty) with these sets and
// returns 0 when every (level, pass) write set and the union of the required
// level sets agree, else 1 + the index of the first disagreeing write set, or
// -1 for the required set.
static void bounds_points(int64_t lo, int64_t hi, int64_t stride, int64_t period,
                          std::vector<char> &out) {
  std::fill(out.begin(), out.end(), 0);
  if (lo < 0)
    return;
  int64_t const last = hi + (lo > hi ? period : 0);
  for (int64_t i = lo; i <= last; i += stride)
    out[i % period] = 1;
}

int oracle_plan_check(int level, int maximum, int64_t start, int64_t count,
                      int const *levels, int nlevels, int const *offsets,
                      int nfilters, int64_t const *writes,
                      int64_t const *required) {
  int64_t const period = int64_t(1) << maximum;
  int64_t const spacing = int64_t(1) << (maximum - level);
  std::vector<char> req(period, 0);
  for (int64_t k = 0; k < count; ++k)
    req[((start + k) * spacing) % period] = 1;
  std::vector<std::vector<char>> expected(size_t(nlevels) * nfilters,
                                          std::vector<char>(period, 0));
  for (int l = nlevels; l-- > 0;) {
    int64_t const stride = int64_t(1) << (maximum - levels[l]);
    for (int p = nfilters; p-- > 0;) {
      auto &e = expected[size_t(l) * nfilters + p];
      for (int64_t i = 0; i < period; ++i)
        if (req[i] && i % (2 * stride) == offsets[p] * stride)
          e[i] = 1;
      for (int64_t i = 0; i < period; ++i)
        if (e[i]) {
          req[(i + period - stride) % period] = 1;
          req[(i + stride) % period] = 1;
        }
    }
  }
  std::vector<char> got(period);
  for (int l = 0; l < nlevels; ++l)
    for (int p = 0; p < nfilters; ++p) {
      size_t const k = size_t(l) * nfilters + p;
      bounds_points(writes[2 * k], writes[2 * k + 1],
                    int64_t(2) << (maximum - levels[l]), period, got);
      if (got != expected[k])
        return 1 + int(k);
    }
  std::vector<char> all(period, 0);
  for (int h = 0; h <= level; ++h) {
    int64_t const stride = h == 0 ? period : int64_t(2) << (maximum - h);
    bounds_points(required[2 * h], required[2 * h + 1], stride, period, got);
    for (int64_t i = 0; i < period; ++i)
      all[i] |= got[i];
  }
  return all == req ? 0 : -1;
}

} // extern "C"
