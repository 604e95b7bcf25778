// Exact dependency planner of a partial-output lifting transform along one
// axis (host code).  It answers what the reference's transform_bounds and
// covering_periodic_bounds answer (paliwa/paliwa_domains.hpp:337-444,
// 472-487): given the output points, the levels in execution order and the
// passes' parities, which points each (level, pass) must write and, per
// hierarchical level, which input values are read.  Results are inclusive
// periodic intervals of finest-level indices on the period P = 2^M (M =
// maximum_level): min > max wraps, {-1, -1} is empty, a whole lattice is
// written {offset, P - stride + offset}.
//
// Design.  Nothing here works in finest coordinates.  Each hierarchical level
// h has its own point index: level h >= 1 holds the points (2q + 1) 2^(M-h),
// level 0 the points q P, for every integer q (q outside one period names a
// periodic copy, so a dependency cone can cross the origin without special
// cases).  A pass at level l works on the level-l lattice of spacing
// t = 2^(M-l): its odd points are level l's points q (x = (2q + 1) t), its
// even points are indexed by j (x = 2 j t) and are exactly the points of the
// coarser levels, level h < l sitting at j = (2q + 1) 2^(l-1-h) (level 0 at
// j = q 2^(l-1)).  The dependency state is one run of indices per level.
// Walking the passes from the last to the first:
//   odd pass:  writes = the run of level l; they read the even j in
//              [first, last + 1], which land in coarser level runs by a
//              floor/ceil division per level;
//   even pass: writes = the j-span covered by the coarser levels' runs; they
//              read the odd q in [jfirst - 1, jlast], which join level l.
// A cone grows by one index per pass at the pass's level, so the runs stay
// contiguous and the sets are exact for consecutive levels (checked against a
// brute force of the stencil graph).  Indices are kept in 128-bit integers:
// with up to 16 passes per level a cone can grow to 16 periods, beyond 2^63
// finest points at the largest periods.  The filter coefficients are never
// read.
#include "common.hpp"

#include <algorithm>
#include <vector>

namespace {

using namespace paliwa_hip;
using Index = __int128;

// floor(a / 2^k) and ceil(a / 2^k) for any sign of a.
Index floor_shift(Index a, int k) { return a >> k; }
Index ceil_shift(Index a, int k) { return -((-a) >> k); }
Index floor_half(Index a) { return floor_shift(a, 1); }
Index ceil_half(Index a) { return ceil_shift(a, 1); }

// A run of consecutive point indices of one lattice; count 0 = empty.
struct Run {
  Index first = 0;
  Index count = 0;

  bool empty() const { return count <= 0; }
  Index last() const { return first + count - 1; }
  static Run between(Index a, Index b) { return b < a ? Run{} : Run{a, b - a + 1}; }
  Run merged(Run const &r) const {
    if (r.empty())
      return *this;
    if (empty())
      return r;
    return between(std::min(first, r.first), std::max(last(), r.last()));
  }
};

// The periodic form of a run on a lattice of `size` points per period whose
// point i lies at finest coordinate offset + i * spacing (< P for i < size).
paliwa_periodic_bounds periodic_form(Run const &r, int64_t size, int64_t offset,
                                     int64_t spacing) {
  if (r.empty())
    return {-1, -1};
  if (r.count >= size)
    return {offset, offset + (size - 1) * spacing};
  auto at = [&](Index i) {
    Index m = i % size;
    if (m < 0)
      m += size;
    return offset + static_cast<int64_t>(m) * spacing;
  };
  return {at(r.first), at(r.last())};
}

class Cone {
public:
  Cone(int32_t level, int32_t maximum_level)
      : M_(maximum_level), runs_(static_cast<size_t>(level) + 1) {}

  // Every level's points inside the finest-coordinate window [lo, hi].
  void seed(Index lo, Index hi) {
    for (size_t h = 0; h < runs_.size(); ++h) {
      int const h32 = static_cast<int>(h);
      if (h == 0) {
        runs_[0] = Run::between(ceil_shift(lo, M_), floor_shift(hi, M_));
      } else {
        int const k = M_ - h32; // x = (2q + 1) 2^k
        runs_[h] = Run::between(ceil_half(ceil_shift(lo, k) - 1),
                                floor_half(floor_shift(hi, k) - 1));
      }
    }
  }

  // An odd pass at level l: writes and then the dependencies of its reads.
  Run odd_pass(int l) {
    Run const w = runs_[l];
    if (!w.empty())
      for (int h = 0; h < l; ++h)
        runs_[h] = runs_[h].merged(coarse_points(h, l, w.first, w.last() + 1));
    return w;
  }

  // An even pass at level l: the even-index span of the coarser levels' runs.
  Run even_pass(int l) {
    Run w;
    for (int h = 0; h < l; ++h) {
      Run const &r = runs_[h];
      if (r.empty())
        continue;
      Index const unit = Index(1) << (h == 0 ? l - 1 : l - 1 - h);
      w = w.merged(h == 0 ? Run::between(r.first * unit, r.last() * unit)
                          : Run::between((2 * r.first + 1) * unit, (2 * r.last() + 1) * unit));
    }
    if (!w.empty())
      runs_[l] = runs_[l].merged(Run::between(w.first - 1, w.last()));
    return w;
  }

  Run const &run(int h) const { return runs_[h]; }

private:
  // The points of level h < l among the even points j in [j0, j1] of level l.
  Run coarse_points(int h, int l, Index j0, Index j1) const {
    if (h == 0)
      return Run::between(ceil_shift(j0, l - 1), floor_shift(j1, l - 1));
    int const k = l - 1 - h; // j = (2q + 1) 2^k
    return Run::between(ceil_half(ceil_shift(j0, k) - 1),
                        floor_half(floor_shift(j1, k) - 1));
  }

  int M_;
  std::vector<Run> runs_;
};

// The reference's validation order: the consecutive-level rule first, then,
// walking backwards, a level outside 1..level before any pass of that level
// runs and a parity other than 0/1 at the first pass it meets.
int validate(int32_t level, int32_t const *levels, int32_t nlevels,
             paliwa_filter const *filters, int32_t nfilters) {
  if (nlevels > 1) {
    int32_t const step = levels[1] - levels[0];
    if (step != 1 && step != -1)
      return fail(PALIWA_ERR_INVALID_ARGUMENT, "Transform levels must be consecutive");
    for (int32_t i = 2; i < nlevels; ++i)
      if (levels[i] - levels[i - 1] != step)
        return fail(PALIWA_ERR_INVALID_ARGUMENT,
                    "Transform levels must be monotone and consecutive");
  }
  bool parities_ok = true;
  for (int32_t p = 0; p < nfilters; ++p)
    parities_ok = parities_ok && (filters[p].offset == 0 || filters[p].offset == 1);
  for (int32_t i = nlevels; i-- > 0;) {
    if (levels[i] < 1 || levels[i] > level)
      return fail(PALIWA_ERR_INVALID_ARGUMENT, "Invalid transform level");
    if (!parities_ok)
      return fail(PALIWA_ERR_INVALID_ARGUMENT, "Filter offset not supported");
  }
  return PALIWA_OK;
}

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
  int64_t const P = int64_t(1) << maximum_level;
  bool const no_outputs = outputs.min == -1 && outputs.max == -1;
  if (!no_outputs && (outputs.min < 0 || outputs.max < 0 || outputs.min >= P ||
                      outputs.max >= P))
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "output bounds outside the period");
  if (int rc = validate(level, levels, nlevels, filters, nfilters); rc != PALIWA_OK)
    return rc;

  Cone cone(level, maximum_level);
  if (!no_outputs) // a wrapped request continues into the next period
    cone.seed(outputs.min, Index(outputs.max) + (outputs.max < outputs.min ? P : 0));
  for (int32_t i = nlevels; i-- > 0;) {
    int const l = levels[i];
    int64_t const t = int64_t(1) << (maximum_level - l); // lattice spacing
    int64_t const half = int64_t(1) << (l - 1);          // points per parity
    for (int32_t p = nfilters; p-- > 0;) {
      bool const odd = filters[p].offset == 1;
      Run const w = odd ? cone.odd_pass(l) : cone.even_pass(l);
      if (writes != nullptr)
        writes[int64_t(i) * nfilters + p] =
            periodic_form(w, half, odd ? t : 0, 2 * t);
    }
  }
  if (required_out != nullptr)
    for (int32_t h = 0; h <= level; ++h)
      required_out[h] =
          h == 0 ? periodic_form(cone.run(0), 1, 0, P)
                 : periodic_form(cone.run(h), int64_t(1) << (h - 1),
                                 int64_t(1) << (maximum_level - h),
                                 int64_t(2) << (maximum_level - h));
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

// The shortest periodic run holding every index: it leaves out the widest
// cyclic gap.  Gap i is the distance from index i-1 to index i, gap 0 the
// one that crosses the period's end; among equally wide gaps the lowest i
// is left out.
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
  for (int64_t i = 1; i < count; ++i)
    if (indices[i] <= indices[i - 1])
      return fail(PALIWA_ERR_INVALID_ARGUMENT, "indices must be sorted and distinct");
  auto gap = [&](int64_t i) {
    return i == 0 ? indices[0] + period - indices[count - 1] : indices[i] - indices[i - 1];
  };
  int64_t widest = 0;
  for (int64_t i = 1; i < count; ++i)
    if (gap(i) > gap(widest))
      widest = i;
  *cover = {indices[widest], indices[(widest + count - 1) % count]};
  return PALIWA_OK;
}

} // extern "C"
