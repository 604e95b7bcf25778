// TEST DOUBLE -- not DDC.  The few names of the DDC library (third-party,
// not installed here) that include/paliwa/paliwa_ddc.hpp uses, written from
// how the reference's own code calls them, so that the adapter compiles and
// its domain-to-storage mapping can be tested.  Compiling against this double
// does not show compatibility with a real DDC build.
#pragma once

#include <array>
#include <cstddef>
#include <type_traits>
#include <utility>
#include <vector>

namespace ddc {

namespace detail {
template <class T, class... Ts> constexpr std::size_t rank() {
  std::size_t i = 0, r = sizeof...(Ts);
  ((std::is_same_v<T, Ts> ? (r = i, ++i) : ++i), ...);
  return r;
}
} // namespace detail

template <class... Tags> struct DiscreteElement {
  std::array<long, sizeof...(Tags)> v{};
  DiscreteElement() = default;
  template <class... I, class = std::enable_if_t<sizeof...(I) == sizeof...(Tags)>>
  explicit DiscreteElement(I... i) : v{static_cast<long>(i)...} {}
  template <class T> long uid() const { return v[detail::rank<T, Tags...>()]; }
};

template <class... Tags> struct DiscreteVector {
  std::array<long, sizeof...(Tags)> v{};
  DiscreteVector() = default;
  template <class... I, class = std::enable_if_t<sizeof...(I) == sizeof...(Tags)>>
  explicit DiscreteVector(I... i) : v{static_cast<long>(i)...} {}
  template <class T> long get() const { return v[detail::rank<T, Tags...>()]; }
};

namespace detail {
template <class... Tags> std::array<long, sizeof...(Tags)> array(DiscreteVector<Tags...> const &x) {
  return x.v;
}
} // namespace detail

template <class... Tags> struct DiscreteDomain {
  DiscreteElement<Tags...> first;
  DiscreteVector<Tags...> count;
  DiscreteElement<Tags...> front() const { return first; }
  std::size_t size() const {
    std::size_t n = 1;
    for (long c : count.v)
      n *= std::size_t(c);
    return n;
  }
};

template <class T, class... Tags> DiscreteDomain<T> select(DiscreteDomain<Tags...> const &d) {
  constexpr std::size_t r = detail::rank<T, Tags...>();
  return {DiscreteElement<T>(d.first.v[r]), DiscreteVector<T>(d.count.v[r])};
}

template <class... Tags> struct StridedDiscreteDomain {
  DiscreteElement<Tags...> first;
  DiscreteVector<Tags...> count;
  DiscreteVector<Tags...> stride;
  DiscreteElement<Tags...> front() const { return first; }
  DiscreteVector<Tags...> extents() const { return count; }
  DiscreteVector<Tags...> strides() const { return stride; }
};

template <class... Tags> struct SparseDiscreteDomain {
  std::array<std::vector<long>, sizeof...(Tags)> lists;
  SparseDiscreteDomain() = default;
  // the product of one-dimensional sparse domains
  template <class... Parts, class = std::enable_if_t<(sizeof...(Parts) == sizeof...(Tags)) &&
                                                     (sizeof...(Tags) > 1)>>
  explicit SparseDiscreteDomain(Parts const &...parts) : lists{parts.lists[0]...} {}
  // one axis of a product
  template <class... Others, class = std::enable_if_t<sizeof...(Tags) == 1 &&
                                                      (sizeof...(Others) > 1)>>
  explicit SparseDiscreteDomain(SparseDiscreteDomain<Others...> const &d)
      : lists{d.lists[detail::rank<Tags..., Others...>()]} {}
  explicit SparseDiscreteDomain(std::vector<long> one) : lists{std::move(one)} {}
};

template <class Tag, class F> void host_for_each(SparseDiscreteDomain<Tag> const &d, F &&f) {
  for (long x : d.lists[0])
    f(DiscreteElement<Tag>(x));
}

template <class T, class Dom> struct ChunkSpan {
  T *ptr = nullptr;
  Dom dom;
  T *data_handle() const { return ptr; }
  Dom domain() const { return dom; }
};

} // namespace ddc
