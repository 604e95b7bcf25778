// CPU oracle for the paliwa lifting transform -- TEST INFRASTRUCTURE ONLY.
//
// This header restates, in plain serial C++, the arithmetic of the reference's
// full-grid lifting path so that the HIP implementation can be checked
// bit-for-bit.  Only the tests, the smoke check and the CPU-baseline leg of the
// benchmark may load it (through liboracle.so); the product path never does.
//
// Restated reference functions (file:line in freifrauvonbleifrei/paliwa):
//   * filter tables            paliwa/paliwa_wavelets.hpp:15-28
//   * detail::lifting_pass     paliwa/paliwa_transform.hpp:24-51
//       g(i) = f0*g(lower) + f1*g(i) + f2*g(upper), evaluated left to right,
//       lower = x==0 ? period-stride : x-stride, upper = x+stride==period ? 0
//       : x+stride (lines 38-41, 47-49).
//   * transform_in (full grid) paliwa/paliwa_transform.hpp:215-251
//   * hierarchize_in / dehierarchize_in  paliwa/paliwa_transform.hpp:346-399
//       (levels level..lmin+1 decreasing / lmin+1..level increasing)
//   * hierarchize / dehierarchize        paliwa/paliwa_transform.hpp:401-431
//       (fold over dimensions 0..D-1 for both directions)
//   * strided_domain_from_level + even/odd half lattices
//       paliwa/paliwa_domains.hpp:33-49, 85-111: at level l' of a grid stored
//       at level L along an axis, the storage stride is 2^(L-l') and the
//       period is 2^L.
//   * the CheckBounds=true branch of lifting_pass for partial buffers
//       (transform_in_checked; paliwa/paliwa_transform.hpp:31-32, 42-45)
//   * PoleExecution::PackedPoles CPU schedule (gather a line, run every level
//       and pass, scatter)  paliwa/paliwa_transform.hpp:119-184 -- used only as
//       the CPU timing baseline; its arithmetic is the same.
//
// Storage: dense row-major (last axis contiguous), extent 2^L (the reference's
// periodic lattice) or 2^L+1 (the lattice plus a periodic image plane at index
// 2^L, which is rewritten as a copy of index 0 after each axis sweep).
//
// How each function is pinned (see tests/test_oracle.py):
//   lifting_pass/transform_in: the 1-D brute-force periodic loop of
//     test/test_paliwa_plan.cpp:128-139 on its multi-pass filters (bit-exact),
//     the hierarchical-surplus closed form for "hat", the textbook LeGall 5/3
//     filter bank for "biorthogonal", the full-weighting stencil closed form,
//     single Fourier modes, constant fields;
//   hierarchize/dehierarchize: round trip, tensor (Kronecker) structure
//     checked against per-axis 1-D application, mass conservation and the
//     lmin-lattice invariance of test/test_paliwa_transform.cpp:57-64, 127-131;
//   packed-poles schedule: bit-identical to the direct schedule;
//   transform_in_checked: equal to transform_in on the complete lattice, a
//     brute force over a dictionary of stored points on tiny buffers, and
//     the property of test/test_paliwa_restrict.cpp:94-139 (a buffer of the
//     local points plus the planned required inputs reproduces the full
//     transform on the local points, bit for bit).
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace paliwa_oracle {

using Filter = std::pair<int, std::array<double, 3>>;
using Filters = std::vector<Filter>;

// paliwa/paliwa_wavelets.hpp:15-21
inline const std::map<std::string, Filters> &forward_filters() {
  static const std::map<std::string, Filters> table = {
      {"hat", {{1, {-0.5, 1.0, -0.5}}}},
      {"biorthogonal", {{1, {-0.5, 1.0, -0.5}}, {0, {0.25, 1.0, 0.25}}}},
      {"fullweighting", {{0, {0.25, 0.5, 0.25}}, {1, {-0.5, 1.0, -0.5}}}},
  };
  return table;
}

// paliwa/paliwa_wavelets.hpp:22-28
inline const std::map<std::string, Filters> &reconstruct_filters() {
  static const std::map<std::string, Filters> table = {
      {"hat", {{1, {0.5, 1.0, 0.5}}}},
      {"biorthogonal", {{0, {-0.25, 1.0, -0.25}}, {1, {0.5, 1.0, 0.5}}}},
      {"fullweighting", {{1, {0.5, 1.0, 0.5}}, {0, {-0.5, 2.0, -0.5}}}},
  };
  return table;
}

// A dense row-major grid viewed as lines along one axis: A outer blocks,
// m stored points along the axis, B inner points (the line's element stride).
struct AxisView {
  int64_t A = 1, m = 1, B = 1;
};

inline AxisView axis_view(std::vector<int64_t> const &extents, int dim) {
  AxisView v;
  for (int d = 0; d < dim; ++d)
    v.A *= extents[d];
  v.m = extents[dim];
  for (size_t d = dim + 1; d < extents.size(); ++d)
    v.B *= extents[d];
  return v;
}

// One lifting pass over every line of one axis.  n is the lattice period
// (2^L) in storage indices, s the storage stride of the operating level.
// Restates paliwa/paliwa_transform.hpp:24-51 (the CheckBounds=false branch).
template <typename T>
void lifting_pass(T *g, AxisView v, int64_t n, int64_t s, int offset,
                  std::array<double, 3> const &f) {
  T const f0 = static_cast<T>(f[0]);
  T const f1 = static_cast<T>(f[1]);
  T const f2 = static_cast<T>(f[2]);
  for (int64_t a = 0; a < v.A; ++a) {
    T *block = g + a * v.m * v.B;
    for (int64_t x = offset * s; x < n; x += 2 * s) {
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

// Rewrite the periodic image plane (index n along the axis) from index 0.
template <typename T> void copy_image_plane(T *g, AxisView v, int64_t n) {
  if (v.m != n + 1)
    return;
  for (int64_t a = 0; a < v.A; ++a) {
    T *block = g + a * v.m * v.B;
    for (int64_t b = 0; b < v.B; ++b)
      block[n * v.B + b] = block[b];
  }
}

inline void check_extents(std::vector<int64_t> const &extents,
                          std::vector<int> const &level) {
  if (extents.size() != level.size())
    throw std::invalid_argument("extents and level differ in length");
  for (size_t d = 0; d < extents.size(); ++d) {
    if (level[d] < 0 || level[d] > 40)
      throw std::invalid_argument("level out of range");
    int64_t const n = int64_t(1) << level[d];
    if (extents[d] != n && extents[d] != n + 1)
      throw std::invalid_argument("extent must be 2^level or 2^level+1");
  }
}

// paliwa/paliwa_transform.hpp:215-251: every level of `levels` in the given
// order, every filter pass in table order.
template <typename T>
void transform_in(T *g, std::vector<int64_t> const &extents,
                  std::vector<int> const &level, int dim,
                  std::vector<int> const &levels, Filters const &filters) {
  check_extents(extents, level);
  AxisView const v = axis_view(extents, dim);
  int64_t const n = int64_t(1) << level[dim];
  for (int l : levels) {
    if (l < 1 || l > level[dim])
      throw std::invalid_argument("operating level out of range");
    int64_t const s = int64_t(1) << (level[dim] - l);
    for (auto const &[offset, f] : filters) {
      if (offset != 0 && offset != 1)
        throw std::runtime_error("Filter offset not supported");
      lifting_pass(g, v, n, s, offset, f);
    }
  }
  if (!levels.empty() && !filters.empty())
    copy_image_plane(g, v, n);
}

inline std::vector<int> decreasing_levels(int level, int lmin) {
  std::vector<int> r;
  for (int l = level; l >= lmin + 1; --l)
    r.push_back(l);
  return r;
}

inline std::vector<int> increasing_levels(int level, int lmin) {
  std::vector<int> r;
  for (int l = lmin + 1; l <= level; ++l)
    r.push_back(l);
  return r;
}

inline void check_levels(std::vector<int> const &level,
                         std::vector<int> const &lmin,
                         std::vector<int> const &lmax) {
  // paliwa/paliwa_transform.hpp:359-362 (asserts in the reference)
  for (size_t d = 0; d < level.size(); ++d)
    if (lmin[d] < 0 || lmin[d] > level[d] || level[d] > lmax[d])
      throw std::invalid_argument("require 0 <= minimum <= level <= maximum");
}

// paliwa/paliwa_transform.hpp:346-372, 401-415
template <typename T>
void hierarchize(T *g, std::vector<int64_t> const &extents,
                 std::vector<int> const &level, std::vector<int> const &lmin,
                 std::vector<int> const &lmax, std::string const &name) {
  check_levels(level, lmin, lmax);
  Filters const &filters = forward_filters().at(name);
  for (int d = 0; d < static_cast<int>(level.size()); ++d)
    transform_in(g, extents, level, d, decreasing_levels(level[d], lmin[d]),
                 filters);
}

// paliwa/paliwa_transform.hpp:374-399, 417-431
template <typename T>
void dehierarchize(T *g, std::vector<int64_t> const &extents,
                   std::vector<int> const &level, std::vector<int> const &lmin,
                   std::vector<int> const &lmax, std::string const &name) {
  check_levels(level, lmin, lmax);
  Filters const &filters = reconstruct_filters().at(name);
  for (int d = 0; d < static_cast<int>(level.size()); ++d)
    transform_in(g, extents, level, d, increasing_levels(level[d], lmin[d]),
                 filters);
}

// ---- partial buffers: the CheckBounds=true branch -------------------------
// paliwa/paliwa_transform.hpp:24-51 with CheckBounds, taken by transform_in
// whenever the buffer is not the complete strided lattice (:224-228, 241-248).
// A buffer is a dense row-major array over the positions of each axis; axis d
// lists the finest-level coordinates of its positions (ascending, in
// [0, 2^maximum_level[d])).  A pass along `dim` at level l (spacing
// t = 2^(maximum_level[dim] - l), period P = 2^maximum_level[dim]) visits the
// half lattice of parity `offset` -- coordinates offset*t + 2t*k along dim
// and, along every other axis d, the multiples of 2^(maximum_level[d] -
// level[d]) (the operating lattice of strided_domain_from_level,
// paliwa/paliwa_domains.hpp:33-49, 85-111) -- and updates a point only when
// it and both neighbours (x -+ t with the periodic wrap of :38-41) are stored
// (:31-32, 42-45).  Written points and the points read have opposite parity,
// so the order of visits does not matter.
template <typename T>
void transform_in_checked(T *g, std::vector<std::vector<int64_t>> const &coords,
                          std::vector<int> const &level,
                          std::vector<int> const &maximum_level, int dim,
                          std::vector<int> const &levels, Filters const &filters) {
  size_t const D = coords.size();
  if (level.size() != D || maximum_level.size() != D || dim < 0 ||
      dim >= static_cast<int>(D))
    throw std::invalid_argument("bad rank or axis");
  for (size_t d = 0; d < D; ++d) {
    if (level[d] < 0 || level[d] > maximum_level[d] || maximum_level[d] > 62)
      throw std::invalid_argument("require 0 <= level <= maximum_level <= 62");
    for (size_t i = 0; i < coords[d].size(); ++i)
      if (coords[d][i] < 0 || coords[d][i] >= (int64_t(1) << maximum_level[d]) ||
          (i > 0 && coords[d][i] <= coords[d][i - 1]))
        throw std::invalid_argument("coordinates must ascend inside the period");
  }
  // lines: every combination of positions of the other axes, in row-major
  // order; a line takes part when each of those coordinates is on its
  // axis's operating lattice
  std::vector<int64_t> extent(D);
  for (size_t d = 0; d < D; ++d)
    extent[d] = static_cast<int64_t>(coords[d].size());
  AxisView const v = axis_view(extent, dim);
  if (v.A * v.m * v.B == 0)
    return;
  std::vector<char> line_on(v.A * v.B, 1);
  for (int64_t line = 0; line < v.A * v.B; ++line) {
    int64_t rest = line;
    for (size_t d = D; d-- > 0;) {
      if (static_cast<int>(d) == dim)
        continue;
      int64_t const pos = rest % extent[d];
      rest /= extent[d];
      int64_t const spacing = int64_t(1) << (maximum_level[d] - level[d]);
      if (coords[d][pos] % spacing != 0)
        line_on[line] = 0;
    }
  }
  std::vector<int64_t> const &c = coords[dim];
  auto position_of = [&c](int64_t x) -> int64_t { // -1: not stored
    size_t lo = 0, hi = c.size();
    while (lo < hi) {
      size_t const mid = (lo + hi) / 2;
      if (c[mid] < x)
        lo = mid + 1;
      else
        hi = mid;
    }
    return lo < c.size() && c[lo] == x ? static_cast<int64_t>(lo) : -1;
  };
  int64_t const P = int64_t(1) << maximum_level[dim];
  for (int l : levels) {
    if (l < 1 || l > level[dim])
      throw std::invalid_argument("operating level out of range");
    int64_t const t = int64_t(1) << (maximum_level[dim] - l);
    for (auto const &[offset, f] : filters) {
      if (offset != 0 && offset != 1)
        throw std::runtime_error("Filter offset not supported");
      T const f0 = static_cast<T>(f[0]);
      T const f1 = static_cast<T>(f[1]);
      T const f2 = static_cast<T>(f[2]);
      for (int64_t p = 0; p < v.m; ++p) {
        int64_t const x = c[p];
        if (x % (2 * t) != offset * t)
          continue;
        int64_t const lo = position_of(x == 0 ? P - t : x - t);
        int64_t const hi = position_of(x + t == P ? 0 : x + t);
        if (lo < 0 || hi < 0)
          continue;
        for (int64_t a = 0; a < v.A; ++a)
          for (int64_t b = 0; b < v.B; ++b) {
            if (!line_on[a * v.B + b])
              continue;
            T *block = g + a * v.m * v.B + b;
            block[p * v.B] = f0 * block[lo * v.B] + f1 * block[p * v.B] +
                             f2 * block[hi * v.B];
          }
      }
    }
  }
}

} // namespace paliwa_oracle
