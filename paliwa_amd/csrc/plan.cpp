// Exact dependency planner of a partial-output lifting transform along one
// axis (host code).  Semantics of the reference's transform_bounds and
// covering_periodic_bounds (paliwa/paliwa_domains.hpp:315-444, 472-487):
// given the output points, the levels in execution order and the passes'
// parities, it returns for every (level, pass) the points that pass must
// write and, per hierarchical level, the points whose input values are read.
// Coordinates are finest-level indices on the period 2^maximum_level; a set is
// an inclusive periodic interval on the lattice of its level or pass (min >
// max wraps, {-1, -1} is empty).
//
// How it works: the points of one axis are partitioned into hierarchical
// levels (level 0: the point 0; level h >= 1: odd multiples of
// 2^(maximum_level - h)), and the dependency state is one interval per level.
// Walking the passes backwards, a pass at level l with offset 1 writes the
// requested level-l points; with offset 0 it writes the requested points of
// every coarser level (the even points of level l).  Its writes read their
// neighbours at +-2^(maximum_level - l), which join the state: odd neighbours
// of an even write are level-l points, even neighbours of an odd write fall on
// the coarser levels.  With consecutive levels the state stays one interval per
// level and the sets are exact (no point is planned that the stencil graph
// does not need); the coefficients are never inspected.
#include "common.hpp"

#include <algorithm>
#include <vector>

namespace {

using namespace paliwa_hip;

// An interval on the unwrapped axis (min > max: empty).
struct Span {
  int64_t min = 1;
  int64_t max = 0;

  bool empty() const { return min > max; }

  // the part of the span on the lattice offset + stride * Z
  Span on_lattice(int64_t stride, int64_t offset) const {
    if (empty())
      return {};
    auto mod = [stride](int64_t x) { return ((x % stride) + stride) % stride; };
    return {min + mod(offset - min), max - mod(max - offset)};
  }

  void hull(Span const &other) {
    if (other.empty())
      return;
    if (empty()) {
      *this = other;
      return;
    }
    min = std::min(min, other.min);
    max = std::max(max, other.max);
  }

  paliwa_periodic_bounds wrap(int64_t period, int64_t stride, int64_t offset) const {
    if (empty())
      return {-1, -1};
    if ((max - min) / stride + 1 >= period / stride) // the whole lattice
      return {offset, period - stride + offset};
    auto mod = [period](int64_t x) { return ((x % period) + period) % period; };
    return {mod(min), mod(max)};
  }
};

} // namespace

namespace paliwa_hip {

int plan_transform_bounds(paliwa_periodic_bounds outputs, int32_t level,
                          int32_t maximum_level, int32_t const *levels,
                          int32_t nlevels, paliwa_filter const *filters,
                          int32_t nfilters, paliwa_periodic_bounds *writes,
                          paliwa_periodic_bounds *required_out) {
  if (nlevels < 0 || nfilters < 0 || (nlevels > 0 && levels == nullptr) ||
      (nfilters > 0 && filters == nullptr))
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "levels and filters are required");
  if (level < 0 || maximum_level < level || maximum_level > 62)
    return fail(PALIWA_ERR_LEVEL, "need 0 <= level <= maximum_level <= 62");
  if (nlevels > 1) {
    int32_t const step = levels[1] - levels[0];
    if (step != 1 && step != -1)
      return fail(PALIWA_ERR_INVALID_ARGUMENT, "Transform levels must be consecutive");
    for (int32_t i = 1; i < nlevels; ++i)
      if (levels[i] - levels[i - 1] != step)
        return fail(PALIWA_ERR_INVALID_ARGUMENT,
                    "Transform levels must be monotone and consecutive");
  }
  int64_t const period = int64_t(1) << maximum_level;
  if (!(outputs.min == -1 && outputs.max == -1) &&
      (outputs.min < 0 || outputs.max < 0 || outputs.min >= period || outputs.max >= period))
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "output bounds outside the period");
  auto lattice_stride = [&](int32_t h) -> int64_t {
    return h == 0 ? period : int64_t(2) << (maximum_level - h);
  };
  auto lattice_offset = [&](int32_t h) -> int64_t {
    return h == 0 ? 0 : int64_t(1) << (maximum_level - h);
  };
  Span out;
  if (outputs.min >= 0)
    out = {outputs.min, outputs.max + (outputs.min > outputs.max ? period : 0)};
  std::vector<Span> req(level + 1);
  for (int32_t h = 0; h <= level; ++h)
    req[h] = out.on_lattice(lattice_stride(h), lattice_offset(h));

  for (int32_t i = nlevels; i-- > 0;) {
    int32_t const l = levels[i];
    if (l < 1 || l > level)
      return fail(PALIWA_ERR_INVALID_ARGUMENT, "Invalid transform level");
    int64_t const stride = int64_t(1) << (maximum_level - l);
    for (int32_t p = nfilters; p-- > 0;) {
      int32_t const offset = filters[p].offset;
      if (offset != 0 && offset != 1)
        return fail(PALIWA_ERR_INVALID_ARGUMENT, "Filter offset not supported");
      Span w;
      if (offset == 1) {
        w = req[l];
      } else {
        for (int32_t h = 0; h < l; ++h)
          w.hull(req[h]);
      }
      if (writes != nullptr)
        writes[int64_t(i) * nfilters + p] = w.wrap(period, 2 * stride, offset * stride);
      if (w.empty())
        continue;
      Span const reads{w.min - stride, w.max + stride};
      if (offset == 0) {
        req[l].hull(reads.on_lattice(2 * stride, stride));
      } else {
        for (int32_t h = 0; h < l; ++h)
          req[h].hull(reads.on_lattice(lattice_stride(h), lattice_offset(h)));
      }
    }
  }
  if (required_out != nullptr)
    for (int32_t h = 0; h <= level; ++h)
      required_out[h] = req[h].wrap(period, lattice_stride(h), lattice_offset(h));
  return PALIWA_OK;
}

} // namespace paliwa_hip

extern "C" {

int paliwa_transform_bounds(paliwa_periodic_bounds outputs, int32_t level,
                            int32_t maximum_level, const int32_t *levels,
                            int32_t nlevels, const paliwa_filter *filters,
                            int32_t nfilters, paliwa_periodic_bounds *writes,
                            paliwa_periodic_bounds *required) {
  return plan_transform_bounds(outputs, level, maximum_level, levels, nlevels,
                               filters, nfilters, writes, required);
}

int paliwa_covering_periodic_bounds(const int64_t *indices, int64_t count,
                                    int64_t period,
                                    paliwa_periodic_bounds *cover) {
  if (cover == nullptr || count < 0 || (count > 0 && indices == nullptr) ||
      period < 1)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "bad arguments");
  if (count == 0) {
    *cover = {-1, -1};
    return PALIWA_OK;
  }
  // the complement of the largest gap between cyclically consecutive indices
  paliwa_periodic_bounds best{indices[0], indices[count - 1]};
  int64_t widest = indices[0] + period - indices[count - 1];
  for (int64_t i = 1; i < count; ++i) {
    if (indices[i] <= indices[i - 1])
      return fail(PALIWA_ERR_INVALID_ARGUMENT, "indices must be sorted and distinct");
    int64_t const gap = indices[i] - indices[i - 1];
    if (gap > widest) {
      widest = gap;
      best = {indices[i], indices[i - 1]};
    }
  }
  *cover = best;
  return PALIWA_OK;
}

} // extern "C"
