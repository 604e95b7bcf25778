// Drop-in for paliwa/paliwa_transform.hpp's full-grid interface on MI355X.
//
// Same function names, argument order, defaults and exception types as the
// reference's templates:
//   hierarchize / dehierarchize            paliwa/paliwa_transform.hpp:401-431
//   hierarchize_in<Dim> / dehierarchize_in  paliwa/paliwa_transform.hpp:346-399
//   transform_in<Dim> (complete grid)       paliwa/paliwa_transform.hpp:215-251
//   transform_in<Dim> (partial outputs)     paliwa/paliwa_transform.hpp:253-344
// The work runs in the HIP kernels of libpaliwa_hip.so through the C ABI of
// paliwa_hip.h.  Link with -lpaliwa_hip (paliwa_amd/lib).
//
// Partial buffers (the reference's checked branch, taken for every buffer that
// is not the complete strided lattice, paliwa/paliwa_transform.hpp:24-51,
// 224-248) go through the same names when the grid also has storage(): an
// array of paliwa_axis_storage, the finest-level coordinates of each axis's
// positions (paliwa::PartialSpan is one; include/paliwa/paliwa_ddc.hpp builds
// it from DDC's DiscreteDomain, StridedDiscreteDomain and SparseDiscreteDomain).
//
// Differences forced by the absence of DDC:
//   * the grid is any object with data_handle() (a device pointer to float or
//     double) and extents() (a std::array-like of per-axis extents, or an
//     mdspan-style extents object with extent(i)); paliwa::GridSpan is one.
//     Storage is dense row-major (DDC's layout_right), extent 2^level or
//     2^level+1 (periodic image) per axis;
//   * level vectors are std::array<long, D> (DDC: DiscreteVector<DDims...>);
//   * the template parameter Dim of the *_in functions is the axis index;
//   * the execution space is paliwa::HipExec, a HIP stream (the reference's
//     ExecSpace instance), passed explicitly as the reference's callers do
//     (its ExecSpace template parameter has no default); calls are
//     asynchronous on it and never synchronize, except the first call of the
//     partial-output transform_in with a given set of arguments, which
//     creates its plan (below).
// Errors: an unknown wavelet name throws std::out_of_range (std::map::at in
// the reference), a filter offset other than 0/1 throws
// std::runtime_error("Filter offset not supported"), violated level ranges
// (asserted by the reference) and malformed grids throw std::invalid_argument,
// a failed HIP call throws std::runtime_error.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "../paliwa_hip.h"
#include "paliwa_wavelets.hpp"

namespace paliwa {

// Non-owning view of a dense row-major grid in device memory.
template <typename T, std::size_t D> struct GridSpan {
  T *ptr = nullptr;
  std::array<std::int64_t, D> shape{};
  T *data_handle() const { return ptr; }
  std::array<std::int64_t, D> const &extents() const { return shape; }
};

// A partial buffer: dense row-major over the positions of each axis, whose
// finest-level coordinates `axes` describes (affine or a device list).
template <typename T, std::size_t D> struct PartialSpan {
  T *ptr = nullptr;
  std::array<paliwa_axis_storage, D> axes{};
  T *data_handle() const { return ptr; }
  std::array<std::int64_t, D> extents() const {
    std::array<std::int64_t, D> e{};
    for (std::size_t d = 0; d < D; ++d)
      e[d] = axes[d].count;
    return e;
  }
  std::array<paliwa_axis_storage, D> const &storage() const { return axes; }
};

// A HIP stream standing in for the reference's Kokkos execution-space instance.
struct HipExec {
  hipStream_t stream = nullptr;
};

template <std::size_t D> using LevelVector = std::array<long, D>;

namespace detail {

inline void throw_on_status(int rc) {
  if (rc == PALIWA_OK)
    return;
  std::string const message = paliwa_last_error();
  switch (rc) {
  case PALIWA_ERR_UNKNOWN_WAVELET:
    throw std::out_of_range(message);
  case PALIWA_ERR_FILTER_OFFSET:
  case PALIWA_ERR_HIP:
    throw std::runtime_error(message);
  default:
    throw std::invalid_argument(message);
  }
}

template <typename T> constexpr int32_t dtype_code() {
  using U = std::remove_cv_t<T>;
  static_assert(std::is_same_v<U, double> || std::is_same_v<U, float>,
                "paliwa grids hold double or float");
  return std::is_same_v<U, double> ? PALIWA_FLOAT64 : PALIWA_FLOAT32;
}

template <typename E>
auto extent_of(E const &e, std::size_t i) -> decltype(e.extent(i)) {
  return e.extent(i);
}
template <typename E>
auto extent_of(E const &e, std::size_t i) -> decltype(e[i]) {
  return e[i];
}

// number of axes of an extents object: rank() (mdspan style) or size()
template <typename E> auto rank_of(E const &e) -> decltype(e.rank()) { return e.rank(); }
template <typename E> auto rank_of(E const &e) -> decltype(e.size()) { return e.size(); }

template <std::size_t D, typename Span>
std::array<int64_t, D> grid_extents(Span const &grid) {
  std::array<int64_t, D> out{};
  auto const &e = grid.extents();
  for (std::size_t d = 0; d < D; ++d)
    out[d] = static_cast<int64_t>(extent_of(e, d));
  return out;
}

template <std::size_t D>
std::array<int32_t, D> to_i32(LevelVector<D> const &v) {
  std::array<int32_t, D> out{};
  for (std::size_t d = 0; d < D; ++d)
    out[d] = static_cast<int32_t>(v[d]);
  return out;
}

template <typename Span>
using value_type_of = std::remove_pointer_t<decltype(std::declval<Span>().data_handle())>;

// a grid with storage(): a partial buffer for the checked branch
template <typename Span, typename = void> struct has_storage : std::false_type {};
template <typename Span>
struct has_storage<Span, std::void_t<decltype(std::declval<Span const &>().storage())>>
    : std::true_type {};

template <std::size_t D, typename Span>
std::array<paliwa_axis_storage, D> storage_of(Span const &grid) {
  std::array<paliwa_axis_storage, D> out{};
  auto const &s = grid.storage();
  for (std::size_t d = 0; d < D; ++d)
    out[d] = s[d];
  return out;
}

inline std::vector<paliwa_filter>
to_filters(std::vector<std::pair<int, std::array<double, 3>>> const &filters) {
  std::vector<paliwa_filter> fs;
  for (auto const &[offset, c] : filters)
    fs.push_back(paliwa_filter{offset, 0, {c[0], c[1], c[2]}});
  return fs;
}

} // namespace detail

// paliwa/paliwa_transform.hpp:215-251: every level of `levels` in order, every
// filter pass in order, along axis Dim.  Returns true.
template <std::size_t Dim, typename Chunk, typename LevelRange, typename ExecSpace,
          std::size_t D>
bool transform_in(
    Chunk grid, LevelVector<D> const &level, LevelVector<D> const &maximum_level,
    LevelRange const &levels,
    std::vector<std::pair<int, std::array<double, 3>>> const &filters,
    ExecSpace instance = ExecSpace()) {
  static_assert(Dim < D, "axis out of range");
  auto const lv = detail::to_i32(level);
  auto const mx = detail::to_i32(maximum_level);
  std::vector<int32_t> ls;
  for (auto l : levels)
    ls.push_back(static_cast<int32_t>(l));
  auto const fs = detail::to_filters(filters);
  constexpr int32_t dtype = detail::dtype_code<detail::value_type_of<Chunk>>();
  if constexpr (detail::has_storage<Chunk>::value) {
    auto const st = detail::storage_of<D>(grid);
    detail::throw_on_status(paliwa_transform_in_checked(
        grid.data_handle(), dtype, static_cast<int32_t>(D), st.data(), lv.data(), mx.data(),
        static_cast<int32_t>(Dim), ls.data(), static_cast<int32_t>(ls.size()), fs.data(),
        static_cast<int32_t>(fs.size()), instance.stream));
  } else {
    auto const ext = detail::grid_extents<D>(grid);
    detail::throw_on_status(paliwa_transform_in(
        grid.data_handle(), dtype, static_cast<int32_t>(D), ext.data(), lv.data(), mx.data(),
        static_cast<int32_t>(Dim), ls.data(), static_cast<int32_t>(ls.size()), fs.data(),
        static_cast<int32_t>(fs.size()), instance.stream));
  }
  return true;
}

// The output points along the transformed axis for the partial-output form:
// `count` consecutive lattice points (at the grid's level) starting at the
// finest-level index `start` (the reference's `outputs` domain along Dim;
// along the other axes every line of the grid is transformed).
struct AxisRun {
  long start = 0;
  long count = 0;
};

// paliwa/paliwa_transform.hpp:331-344 (outputs held in a single allocation):
// the grid holds the whole axis Dim (2^level[Dim] points, no periodic image);
// only the output points are written, bit-identical to the full transform
// (the other points keep their values).  The levels must be consecutive and
// monotone (std::invalid_argument otherwise, as the reference's planner).
// The first call with a given set of arguments (device, dtype, shape, axis,
// levels, filters, outputs) plans on the host and creates the plan's small
// device tables, which allocates and copies synchronously; the plan is kept
// until the process ends, and later calls with the same arguments neither
// allocate nor synchronize.  The C ABI's paliwa_partial_* plan gives the
// caller control of that lifetime and also takes ghost values of a rank's
// block.
namespace detail {
struct PartialPlans {
  std::mutex mutex;
  std::map<std::vector<std::int64_t>, paliwa_partial *> plans;
};
inline PartialPlans &partial_plans() {
  static PartialPlans *plans = new PartialPlans; // never destroyed
  return *plans;
}
} // namespace detail

template <std::size_t Dim, typename Chunk, typename LevelRange, typename ExecSpace,
          std::size_t D>
bool transform_in(
    Chunk grid, LevelVector<D> const &level, LevelVector<D> const &maximum_level,
    LevelRange const &levels,
    std::vector<std::pair<int, std::array<double, 3>>> const &filters,
    AxisRun const &outputs, ExecSpace instance) {
  static_assert(Dim < D, "axis out of range");
  auto const ext = detail::grid_extents<D>(grid);
  long const spacing_log2 = maximum_level[Dim] - level[Dim];
  if (spacing_log2 < 0 || outputs.count < 0 ||
      (outputs.start & ((1L << spacing_log2) - 1)) != 0)
    throw std::invalid_argument("outputs are not on the grid's lattice");
  std::vector<int32_t> ls;
  for (auto l : levels)
    ls.push_back(static_cast<int32_t>(l));
  auto const fs = detail::to_filters(filters);
  constexpr int32_t dtype = detail::dtype_code<detail::value_type_of<Chunk>>();
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess)
    throw std::runtime_error("no HIP device");
  std::vector<std::int64_t> key{device, dtype, static_cast<std::int64_t>(D),
                                static_cast<std::int64_t>(Dim), level[Dim],
                                outputs.start >> spacing_log2, outputs.count};
  key.insert(key.end(), ext.begin(), ext.end());
  key.insert(key.end(), ls.begin(), ls.end());
  for (auto const &f : fs) {
    key.push_back(f.offset);
    for (double c : f.coefficients) {
      std::int64_t bits;
      std::memcpy(&bits, &c, sizeof bits);
      key.push_back(bits);
    }
  }
  paliwa_partial *plan = nullptr;
  {
    auto &cache = detail::partial_plans();
    std::lock_guard<std::mutex> lock(cache.mutex);
    auto const it = cache.plans.find(key);
    if (it != cache.plans.end()) {
      plan = it->second;
    } else {
      detail::throw_on_status(paliwa_partial_create(
          &plan, dtype, static_cast<int32_t>(D), ext.data(), static_cast<int32_t>(Dim),
          static_cast<int32_t>(level[Dim]), ls.data(), static_cast<int32_t>(ls.size()),
          fs.data(), static_cast<int32_t>(fs.size()), 0, outputs.start >> spacing_log2,
          outputs.count));
      cache.plans.emplace(std::move(key), plan);
    }
  }
  detail::throw_on_status(
      paliwa_partial_transform(plan, grid.data_handle(), nullptr, instance.stream));
  return true;
}

// paliwa/paliwa_transform.hpp:346-372
template <std::size_t Dim, typename ChunkSpanType, typename ExecSpace, std::size_t D>
bool hierarchize_in(ChunkSpanType const strided_grid, LevelVector<D> const &level,
                    LevelVector<D> const &minimum_level,
                    LevelVector<D> const &maximum_level,
                    std::string const &wavelet_name = "hat",
                    ExecSpace instance = ExecSpace()) {
  static_assert(Dim < D, "axis out of range");
  auto const lv = detail::to_i32(level), mn = detail::to_i32(minimum_level),
             mx = detail::to_i32(maximum_level);
  constexpr int32_t dtype = detail::dtype_code<detail::value_type_of<ChunkSpanType>>();
  if constexpr (detail::has_storage<ChunkSpanType>::value) {
    auto const st = detail::storage_of<D>(strided_grid);
    detail::throw_on_status(paliwa_hierarchize_in_checked(
        strided_grid.data_handle(), dtype, static_cast<int32_t>(D), st.data(), lv.data(),
        mn.data(), mx.data(), static_cast<int32_t>(Dim), wavelet_name.c_str(),
        instance.stream));
  } else {
    auto const ext = detail::grid_extents<D>(strided_grid);
    detail::throw_on_status(paliwa_hierarchize_in(
        strided_grid.data_handle(), dtype, static_cast<int32_t>(D), ext.data(), lv.data(),
        mn.data(), mx.data(), static_cast<int32_t>(Dim), wavelet_name.c_str(),
        instance.stream));
  }
  return true;
}

// paliwa/paliwa_transform.hpp:374-399
template <std::size_t Dim, typename ChunkSpanType, typename ExecSpace, std::size_t D>
bool dehierarchize_in(ChunkSpanType const strided_grid, LevelVector<D> const &level,
                      LevelVector<D> const &minimum_level,
                      LevelVector<D> const &maximum_level,
                      std::string const &wavelet_name = "hat",
                      ExecSpace instance = ExecSpace()) {
  static_assert(Dim < D, "axis out of range");
  auto const lv = detail::to_i32(level), mn = detail::to_i32(minimum_level),
             mx = detail::to_i32(maximum_level);
  constexpr int32_t dtype = detail::dtype_code<detail::value_type_of<ChunkSpanType>>();
  if constexpr (detail::has_storage<ChunkSpanType>::value) {
    auto const st = detail::storage_of<D>(strided_grid);
    detail::throw_on_status(paliwa_dehierarchize_in_checked(
        strided_grid.data_handle(), dtype, static_cast<int32_t>(D), st.data(), lv.data(),
        mn.data(), mx.data(), static_cast<int32_t>(Dim), wavelet_name.c_str(),
        instance.stream));
  } else {
    auto const ext = detail::grid_extents<D>(strided_grid);
    detail::throw_on_status(paliwa_dehierarchize_in(
        strided_grid.data_handle(), dtype, static_cast<int32_t>(D), ext.data(), lv.data(),
        mn.data(), mx.data(), static_cast<int32_t>(Dim), wavelet_name.c_str(),
        instance.stream));
  }
  return true;
}

// paliwa/paliwa_transform.hpp:401-415: axes 0, 1, ..., D-1.
template <typename ChunkSpanType, typename ExecSpace, std::size_t D>
void hierarchize(ChunkSpanType const strided_grid, LevelVector<D> const &level,
                 LevelVector<D> const &minimum_level,
                 LevelVector<D> const &maximum_level,
                 std::string const &wavelet_name = "hat",
                 ExecSpace instance = ExecSpace()) {
  auto const lv = detail::to_i32(level), mn = detail::to_i32(minimum_level),
             mx = detail::to_i32(maximum_level);
  constexpr int32_t dtype = detail::dtype_code<detail::value_type_of<ChunkSpanType>>();
  if constexpr (detail::has_storage<ChunkSpanType>::value) {
    auto const st = detail::storage_of<D>(strided_grid);
    detail::throw_on_status(paliwa_hierarchize_checked(
        strided_grid.data_handle(), dtype, static_cast<int32_t>(D), st.data(), lv.data(),
        mn.data(), mx.data(), wavelet_name.c_str(), instance.stream));
  } else {
    auto const ext = detail::grid_extents<D>(strided_grid);
    detail::throw_on_status(paliwa_hierarchize(
        strided_grid.data_handle(), dtype, static_cast<int32_t>(D), ext.data(), lv.data(),
        mn.data(), mx.data(), wavelet_name.c_str(), instance.stream));
  }
}

// paliwa/paliwa_transform.hpp:417-431: axes 0, 1, ..., D-1 (same order as
// the forward transform).
template <typename ChunkSpanType, typename ExecSpace, std::size_t D>
void dehierarchize(ChunkSpanType const strided_grid, LevelVector<D> const &level,
                   LevelVector<D> const &minimum_level,
                   LevelVector<D> const &maximum_level,
                   std::string const &wavelet_name = "hat",
                   ExecSpace instance = ExecSpace()) {
  auto const lv = detail::to_i32(level), mn = detail::to_i32(minimum_level),
             mx = detail::to_i32(maximum_level);
  constexpr int32_t dtype = detail::dtype_code<detail::value_type_of<ChunkSpanType>>();
  if constexpr (detail::has_storage<ChunkSpanType>::value) {
    auto const st = detail::storage_of<D>(strided_grid);
    detail::throw_on_status(paliwa_dehierarchize_checked(
        strided_grid.data_handle(), dtype, static_cast<int32_t>(D), st.data(), lv.data(),
        mn.data(), mx.data(), wavelet_name.c_str(), instance.stream));
  } else {
    auto const ext = detail::grid_extents<D>(strided_grid);
    detail::throw_on_status(paliwa_dehierarchize(
        strided_grid.data_handle(), dtype, static_cast<int32_t>(D), ext.data(), lv.data(),
        mn.data(), mx.data(), wavelet_name.c_str(), instance.stream));
  }
}

} // namespace paliwa
