// Drop-in for the partial-output planner of paliwa/paliwa_domains.hpp on
// MI355X: PeriodicBounds (:315-321), transform_bounds (:337-444) and
// covering_periodic_bounds (:472-487), with the reference's names, argument
// order and exceptions, computed by libpaliwa_hip.so (host code, no device
// work).  Coordinates are finest-level indices on the period
// 2^maximum_level; min > max wraps; {-1, -1} is empty.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../paliwa_hip.h"

namespace paliwa {

struct PeriodicBounds {
  long min = -1;
  long max = -1;
  bool empty() const { return min < 0; }
};

// The exact writes of every (level, pass) and the required inputs of every
// hierarchical level 0..level.  visit_pass(i, pass, writes) is called for
// every pass, backwards (last level and pass first), as the reference visits
// them.  Non-consecutive levels, invalid levels and filter offsets other than
// 0/1 throw std::invalid_argument.
template <typename LevelRange, typename VisitPass>
std::vector<PeriodicBounds> transform_bounds(
    PeriodicBounds outputs, long level, long maximum_level,
    LevelRange const &levels,
    std::vector<std::pair<int, std::array<double, 3>>> const &filters,
    VisitPass visit_pass) {
  std::vector<int32_t> ls;
  for (auto l : levels)
    ls.push_back(static_cast<int32_t>(l));
  std::vector<paliwa_filter> fs;
  for (auto const &[offset, c] : filters)
    fs.push_back(paliwa_filter{offset, 0, {c[0], c[1], c[2]}});
  std::vector<paliwa_periodic_bounds> writes(ls.size() * fs.size() + 1);
  std::vector<paliwa_periodic_bounds> required(static_cast<std::size_t>(level) + 1);
  int const rc = paliwa_transform_bounds(
      paliwa_periodic_bounds{outputs.min, outputs.max}, static_cast<int32_t>(level),
      static_cast<int32_t>(maximum_level), ls.data(), static_cast<int32_t>(ls.size()),
      fs.data(), static_cast<int32_t>(fs.size()), writes.data(), required.data());
  if (rc != PALIWA_OK)
    throw std::invalid_argument(paliwa_last_error());
  for (std::size_t i = ls.size(); i-- > 0;)
    for (std::size_t p = fs.size(); p-- > 0;) {
      auto const &w = writes[i * fs.size() + p];
      visit_pass(i, p, PeriodicBounds{static_cast<long>(w.min), static_cast<long>(w.max)});
    }
  std::vector<PeriodicBounds> out;
  for (auto const &b : required)
    out.push_back(PeriodicBounds{static_cast<long>(b.min), static_cast<long>(b.max)});
  return out;
}

// Smallest circular interval containing sorted, distinct indices.
inline PeriodicBounds covering_periodic_bounds(std::vector<long> const &indices,
                                               long period) {
  std::vector<int64_t> idx(indices.begin(), indices.end());
  paliwa_periodic_bounds cover{};
  if (paliwa_covering_periodic_bounds(idx.data(), static_cast<int64_t>(idx.size()),
                                      period, &cover) != PALIWA_OK)
    throw std::invalid_argument(paliwa_last_error());
  return PeriodicBounds{static_cast<long>(cover.min), static_cast<long>(cover.max)};
}

} // namespace paliwa
