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
// Differences forced by the absence of DDC:
//   * the grid is any object with data_handle() (a device pointer to float or
//     double) and extents() (a std::array-like of per-axis extents, or an
//     mdspan-style extents object with extent(i)); paliwa::GridSpan is one.
//     Storage is dense row-major (DDC's layout_right), extent 2^level or
//     2^level+1 (periodic image) per axis;
//   * level vectors are std::array<long, D> (DDC: DiscreteVector<DDims...>);
//   * the template parameter Dim of the *_in functions is the axis index;
//   * the execution space is paliwa::HipExec, a HIP stream (the reference's
//     ExecSpace instance); calls are asynchronous on it and never synchronize.
// Errors: an unknown wavelet name throws std::out_of_range (std::map::at in
// the reference), a filter offset other than 0/1 throws
// std::runtime_error("Filter offset not supported"), violated level ranges
// (asserted by the reference) and malformed grids throw std::invalid_argument,
// a failed HIP call throws std::runtime_error.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
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

} // namespace detail

// paliwa/paliwa_transform.hpp:215-251: every level of `levels` in order, every
// filter pass in order, along axis Dim.  Returns true.
template <std::size_t Dim, typename Chunk, typename LevelRange,
          typename ExecSpace = HipExec, std::size_t D>
bool transform_in(
    Chunk grid, LevelVector<D> const &level, LevelVector<D> const &maximum_level,
    LevelRange const &levels,
    std::vector<std::pair<int, std::array<double, 3>>> const &filters,
    ExecSpace instance = ExecSpace()) {
  static_assert(Dim < D, "axis out of range");
  auto const ext = detail::grid_extents<D>(grid);
  auto const lv = detail::to_i32(level);
  auto const mx = detail::to_i32(maximum_level);
  std::vector<int32_t> ls;
  for (auto l : levels)
    ls.push_back(static_cast<int32_t>(l));
  std::vector<paliwa_filter> fs;
  for (auto const &[offset, c] : filters)
    fs.push_back(paliwa_filter{offset, 0, {c[0], c[1], c[2]}});
  detail::throw_on_status(paliwa_transform_in(
      grid.data_handle(), detail::dtype_code<detail::value_type_of<Chunk>>(),
      static_cast<int32_t>(D), ext.data(), lv.data(), mx.data(),
      static_cast<int32_t>(Dim), ls.data(), static_cast<int32_t>(ls.size()),
      fs.data(), static_cast<int32_t>(fs.size()), instance.stream));
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
// The call plans on the host, allocates the plan's small device tables,
// runs, and synchronizes the stream before releasing them; the C ABI's
// paliwa_partial_* plan reuses one plan, also with ghost values of a rank's
// block.
template <std::size_t Dim, typename Chunk, typename LevelRange,
          typename ExecSpace = HipExec, std::size_t D>
bool transform_in(
    Chunk grid, LevelVector<D> const &level, LevelVector<D> const &maximum_level,
    LevelRange const &levels,
    std::vector<std::pair<int, std::array<double, 3>>> const &filters,
    AxisRun const &outputs, ExecSpace instance = ExecSpace()) {
  static_assert(Dim < D, "axis out of range");
  auto const ext = detail::grid_extents<D>(grid);
  long const spacing_log2 = maximum_level[Dim] - level[Dim];
  if (spacing_log2 < 0 || outputs.count < 0 ||
      (outputs.start & ((1L << spacing_log2) - 1)) != 0)
    throw std::invalid_argument("outputs are not on the grid's lattice");
  std::vector<int32_t> ls;
  for (auto l : levels)
    ls.push_back(static_cast<int32_t>(l));
  std::vector<paliwa_filter> fs;
  for (auto const &[offset, c] : filters)
    fs.push_back(paliwa_filter{offset, 0, {c[0], c[1], c[2]}});
  paliwa_partial *plan = nullptr;
  detail::throw_on_status(paliwa_partial_create(
      &plan, detail::dtype_code<detail::value_type_of<Chunk>>(),
      static_cast<int32_t>(D), ext.data(), static_cast<int32_t>(Dim),
      static_cast<int32_t>(level[Dim]), ls.data(), static_cast<int32_t>(ls.size()),
      fs.data(), static_cast<int32_t>(fs.size()), 0, outputs.start >> spacing_log2,
      outputs.count));
  int const rc = paliwa_partial_transform(plan, grid.data_handle(), nullptr, instance.stream);
  hipError_t const sync = hipStreamSynchronize(instance.stream);
  paliwa_partial_destroy(plan);
  detail::throw_on_status(rc);
  if (sync != hipSuccess)
    throw std::runtime_error(hipGetErrorString(sync));
  return true;
}

// paliwa/paliwa_transform.hpp:346-372
template <std::size_t Dim, typename ChunkSpanType, typename ExecSpace = HipExec,
          std::size_t D>
bool hierarchize_in(ChunkSpanType const strided_grid, LevelVector<D> const &level,
                    LevelVector<D> const &minimum_level,
                    LevelVector<D> const &maximum_level,
                    std::string const &wavelet_name = "hat",
                    ExecSpace instance = ExecSpace()) {
  static_assert(Dim < D, "axis out of range");
  auto const ext = detail::grid_extents<D>(strided_grid);
  auto const lv = detail::to_i32(level), mn = detail::to_i32(minimum_level),
             mx = detail::to_i32(maximum_level);
  detail::throw_on_status(paliwa_hierarchize_in(
      strided_grid.data_handle(),
      detail::dtype_code<detail::value_type_of<ChunkSpanType>>(),
      static_cast<int32_t>(D), ext.data(), lv.data(), mn.data(), mx.data(),
      static_cast<int32_t>(Dim), wavelet_name.c_str(), instance.stream));
  return true;
}

// paliwa/paliwa_transform.hpp:374-399
template <std::size_t Dim, typename ChunkSpanType, typename ExecSpace = HipExec,
          std::size_t D>
bool dehierarchize_in(ChunkSpanType const strided_grid,
                      LevelVector<D> const &level,
                      LevelVector<D> const &minimum_level,
                      LevelVector<D> const &maximum_level,
                      std::string const &wavelet_name = "hat",
                      ExecSpace instance = ExecSpace()) {
  static_assert(Dim < D, "axis out of range");
  auto const ext = detail::grid_extents<D>(strided_grid);
  auto const lv = detail::to_i32(level), mn = detail::to_i32(minimum_level),
             mx = detail::to_i32(maximum_level);
  detail::throw_on_status(paliwa_dehierarchize_in(
      strided_grid.data_handle(),
      detail::dtype_code<detail::value_type_of<ChunkSpanType>>(),
      static_cast<int32_t>(D), ext.data(), lv.data(), mn.data(), mx.data(),
      static_cast<int32_t>(Dim), wavelet_name.c_str(), instance.stream));
  return true;
}

// paliwa/paliwa_transform.hpp:401-415: axes 0, 1, ..., D-1.
template <typename ChunkSpanType, typename ExecSpace = HipExec, std::size_t D>
void hierarchize(ChunkSpanType const strided_grid, LevelVector<D> const &level,
                 LevelVector<D> const &minimum_level,
                 LevelVector<D> const &maximum_level,
                 std::string const &wavelet_name = "hat",
                 ExecSpace instance = ExecSpace()) {
  auto const ext = detail::grid_extents<D>(strided_grid);
  auto const lv = detail::to_i32(level), mn = detail::to_i32(minimum_level),
             mx = detail::to_i32(maximum_level);
  detail::throw_on_status(paliwa_hierarchize(
      strided_grid.data_handle(),
      detail::dtype_code<detail::value_type_of<ChunkSpanType>>(),
      static_cast<int32_t>(D), ext.data(), lv.data(), mn.data(), mx.data(),
      wavelet_name.c_str(), instance.stream));
}

// paliwa/paliwa_transform.hpp:417-431: axes 0, 1, ..., D-1 (same order as
// the forward transform).
template <typename ChunkSpanType, typename ExecSpace = HipExec, std::size_t D>
void dehierarchize(ChunkSpanType const strided_grid, LevelVector<D> const &level,
                   LevelVector<D> const &minimum_level,
                   LevelVector<D> const &maximum_level,
                   std::string const &wavelet_name = "hat",
                   ExecSpace instance = ExecSpace()) {
  auto const ext = detail::grid_extents<D>(strided_grid);
  auto const lv = detail::to_i32(level), mn = detail::to_i32(minimum_level),
             mx = detail::to_i32(maximum_level);
  detail::throw_on_status(paliwa_dehierarchize(
      strided_grid.data_handle(),
      detail::dtype_code<detail::value_type_of<ChunkSpanType>>(),
      static_cast<int32_t>(D), ext.data(), lv.data(), mn.data(), mx.data(),
      wavelet_name.c_str(), instance.stream));
}

} // namespace paliwa
