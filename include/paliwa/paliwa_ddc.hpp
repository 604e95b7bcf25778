// DDC-typed overloads of the drop-in: the reference's own argument types
// (paliwa/paliwa_transform.hpp:346-431) -- a ddc::ChunkSpan over a
// DiscreteDomain, StridedDiscreteDomain or SparseDiscreteDomain, level
// vectors as ddc::DiscreteVector<DDims...>, the axis of the *_in functions as
// a DDim tag, an execution-space instance -- mapped onto the C ABI.  Compiled
// only where DDC is installed (__has_include(<ddc/ddc.hpp>)); it uses only DDC
// names the reference itself uses.
//
// Each axis of the span's domain becomes a paliwa_axis_storage: a
// DiscreteDomain (stride 1) or StridedDiscreteDomain as front/stride/extent,
// a SparseDiscreteDomain as its coordinate list, copied to device memory with
// stream-ordered allocation (hipMallocAsync / hipFreeAsync on the instance's
// stream).  A domain that is the complete strided lattice of the levels runs
// the unchecked kernels, as the reference's transform_in does (:224-228);
// anything else runs the checked branch.  The instance is a HIP stream: an
// object with a `stream` member (paliwa::HipExec) or with hip_stream() (a
// Kokkos HIP execution space).
#pragma once

#if __has_include(<ddc/ddc.hpp>)

#include <ddc/ddc.hpp>

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "paliwa_transform.hpp"

namespace paliwa {
namespace ddc_adapter {

// position of Tag in Tags...
template <class Tag, class... Tags> constexpr std::size_t index_of() {
  std::size_t i = 0, found = sizeof...(Tags);
  ((std::is_same_v<Tag, Tags> ? (found = i, ++i) : ++i), ...);
  return found;
}

template <class E> auto stream_of(E const &e) -> decltype(e.stream) { return e.stream; }
template <class E> auto stream_of(E const &e) -> decltype(e.hip_stream()) {
  return e.hip_stream();
}

// The storage of one axis of a domain; a sparse axis appends its coordinates
// to `list` and leaves the device pointer to the caller.
template <class Tag, class... Tags>
paliwa_axis_storage axis(ddc::DiscreteDomain<Tags...> const &dom, std::vector<std::int64_t> &) {
  auto const d = ddc::select<Tag>(dom);
  return {static_cast<std::int64_t>(d.size()),
          static_cast<std::int64_t>(dom.front().template uid<Tag>()), 1, nullptr};
}

template <class Tag, class... Tags>
paliwa_axis_storage axis(ddc::StridedDiscreteDomain<Tags...> const &dom,
                         std::vector<std::int64_t> &) {
  return {static_cast<std::int64_t>(dom.extents().template get<Tag>()),
          static_cast<std::int64_t>(dom.front().template uid<Tag>()),
          static_cast<std::int64_t>(dom.strides().template get<Tag>()), nullptr};
}

template <class Tag, class... Tags>
paliwa_axis_storage axis(ddc::SparseDiscreteDomain<Tags...> const &dom,
                         std::vector<std::int64_t> &list) {
  ddc::SparseDiscreteDomain<Tag> const d(dom);
  ddc::host_for_each(d, [&list](ddc::DiscreteElement<Tag> i) {
    list.push_back(static_cast<std::int64_t>(i.template uid<Tag>()));
  });
  return {static_cast<std::int64_t>(list.size()), 0, 1, nullptr};
}

// A PartialSpan over a DDC chunk span, with device copies of its sparse
// coordinate lists (released on the stream when the call is done).
template <class ChunkSpanType, class... DDims> class Buffer {
public:
  using T = std::remove_pointer_t<decltype(std::declval<ChunkSpanType>().data_handle())>;
  static constexpr std::size_t D = sizeof...(DDims);

  Buffer(ChunkSpanType const &grid, hipStream_t stream) : stream_(stream) {
    auto const dom = grid.domain();
    std::size_t k = 0;
    ((span_.axes[k] = axis<DDims>(dom, lists_[k]), ++k), ...);
    span_.ptr = grid.data_handle();
    for (std::size_t d = 0; d < D; ++d) { // sparse axes only
      std::size_t const bytes = lists_[d].size() * sizeof(std::int64_t);
      if (bytes == 0)
        continue;
      void *dev = nullptr;
      if (hipMallocAsync(&dev, bytes, stream_) != hipSuccess ||
          hipMemcpyAsync(dev, lists_[d].data(), bytes, hipMemcpyHostToDevice, stream_) !=
              hipSuccess)
        throw std::runtime_error("could not place a sparse axis's coordinates");
      span_.axes[d].coordinates = static_cast<std::int64_t const *>(dev);
      owned_.push_back(dev);
    }
  }
  ~Buffer() {
    for (void *p : owned_)
      (void)hipFreeAsync(p, stream_);
  }
  Buffer(Buffer const &) = delete;
  Buffer &operator=(Buffer const &) = delete;
  PartialSpan<T, D> const &span() const { return span_; }

private:
  hipStream_t stream_;
  PartialSpan<T, D> span_{};
  std::array<std::vector<std::int64_t>, D> lists_{};
  std::vector<void *> owned_;
};

template <class... DDims>
LevelVector<sizeof...(DDims)> levels_of(ddc::DiscreteVector<DDims...> const &v) {
  auto const a = ddc::detail::array(v);
  LevelVector<sizeof...(DDims)> out{};
  for (std::size_t d = 0; d < out.size(); ++d)
    out[d] = static_cast<long>(a[d]);
  return out;
}

} // namespace ddc_adapter

// paliwa/paliwa_transform.hpp:346-372
template <class DDimInWhichToHierarchize, class ChunkSpanType, class ExecSpace,
          class... DDims>
bool hierarchize_in(ChunkSpanType const strided_grid, ddc::DiscreteVector<DDims...> const &level,
                    ddc::DiscreteVector<DDims...> const &minimum_level,
                    ddc::DiscreteVector<DDims...> const &maximum_level,
                    std::string const &wavelet_name = "hat", ExecSpace instance = ExecSpace()) {
  constexpr std::size_t dim = ddc_adapter::index_of<DDimInWhichToHierarchize, DDims...>();
  static_assert(dim < sizeof...(DDims), "the axis is not a dimension of the level vector");
  hipStream_t const s = ddc_adapter::stream_of(instance);
  ddc_adapter::Buffer<ChunkSpanType, DDims...> const buf(strided_grid, s);
  return hierarchize_in<dim>(buf.span(), ddc_adapter::levels_of(level),
                             ddc_adapter::levels_of(minimum_level),
                             ddc_adapter::levels_of(maximum_level), wavelet_name, HipExec{s});
}

// paliwa/paliwa_transform.hpp:374-399
template <class DDimInWhichToHierarchize, class ChunkSpanType, class ExecSpace,
          class... DDims>
bool dehierarchize_in(ChunkSpanType const strided_grid,
                      ddc::DiscreteVector<DDims...> const &level,
                      ddc::DiscreteVector<DDims...> const &minimum_level,
                      ddc::DiscreteVector<DDims...> const &maximum_level,
                      std::string const &wavelet_name = "hat", ExecSpace instance = ExecSpace()) {
  constexpr std::size_t dim = ddc_adapter::index_of<DDimInWhichToHierarchize, DDims...>();
  static_assert(dim < sizeof...(DDims), "the axis is not a dimension of the level vector");
  hipStream_t const s = ddc_adapter::stream_of(instance);
  ddc_adapter::Buffer<ChunkSpanType, DDims...> const buf(strided_grid, s);
  return dehierarchize_in<dim>(buf.span(), ddc_adapter::levels_of(level),
                               ddc_adapter::levels_of(minimum_level),
                               ddc_adapter::levels_of(maximum_level), wavelet_name,
                               HipExec{s});
}

// paliwa/paliwa_transform.hpp:401-415: the DDims in order
template <class ChunkSpanType, class ExecSpace, class... DDims>
void hierarchize(ChunkSpanType const strided_grid, ddc::DiscreteVector<DDims...> const &level,
                 ddc::DiscreteVector<DDims...> const &minimum_level,
                 ddc::DiscreteVector<DDims...> const &maximum_level,
                 std::string const &wavelet_name = "hat", ExecSpace instance = ExecSpace()) {
  hipStream_t const s = ddc_adapter::stream_of(instance);
  ddc_adapter::Buffer<ChunkSpanType, DDims...> const buf(strided_grid, s);
  hierarchize(buf.span(), ddc_adapter::levels_of(level), ddc_adapter::levels_of(minimum_level),
              ddc_adapter::levels_of(maximum_level), wavelet_name, HipExec{s});
}

// paliwa/paliwa_transform.hpp:417-431
template <class ChunkSpanType, class ExecSpace, class... DDims>
void dehierarchize(ChunkSpanType const strided_grid, ddc::DiscreteVector<DDims...> const &level,
                   ddc::DiscreteVector<DDims...> const &minimum_level,
                   ddc::DiscreteVector<DDims...> const &maximum_level,
                   std::string const &wavelet_name = "hat", ExecSpace instance = ExecSpace()) {
  hipStream_t const s = ddc_adapter::stream_of(instance);
  ddc_adapter::Buffer<ChunkSpanType, DDims...> const buf(strided_grid, s);
  dehierarchize(buf.span(), ddc_adapter::levels_of(level), ddc_adapter::levels_of(minimum_level),
                ddc_adapter::levels_of(maximum_level), wavelet_name, HipExec{s});
}

// paliwa/paliwa_transform.hpp:215-251
template <class DDimIn, class Chunk, class LevelRange, class ExecSpace, class... DDims>
bool transform_in(Chunk grid, ddc::DiscreteVector<DDims...> const &level,
                  ddc::DiscreteVector<DDims...> const &maximum_level, LevelRange const &levels,
                  std::vector<std::pair<int, std::array<double, 3>>> const &filters,
                  ExecSpace instance = ExecSpace()) {
  constexpr std::size_t dim = ddc_adapter::index_of<DDimIn, DDims...>();
  static_assert(dim < sizeof...(DDims), "the axis is not a dimension of the level vector");
  hipStream_t const s = ddc_adapter::stream_of(instance);
  ddc_adapter::Buffer<Chunk, DDims...> const buf(grid, s);
  return transform_in<dim>(buf.span(), ddc_adapter::levels_of(level),
                           ddc_adapter::levels_of(maximum_level), levels, filters, HipExec{s});
}

} // namespace paliwa

#endif // __has_include(<ddc/ddc.hpp>)
