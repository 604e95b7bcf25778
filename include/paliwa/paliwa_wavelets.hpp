// Drop-in for paliwa/paliwa_wavelets.hpp:15-28: the lifting filter tables,
// same names and types.  Each entry is an ordered list of passes
// (offset: 0 even / 1 odd points, {c0, c1, c2}); a pass sets
// g = c0*g[lower] + c1*g + c2*g[upper].
#pragma once

#include <array>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace paliwa {

static const std::map<std::string,
                      std::vector<std::pair<int, std::array<double, 3>>>>
    lifting_wavelet_filter_offsets_and_coefficients = {
        {"hat", {{1, {-0.5, 1.0, -0.5}}}},
        {"biorthogonal", {{1, {-0.5, 1.0, -0.5}}, {0, {0.25, 1.0, 0.25}}}},
        {"fullweighting", {{0, {0.25, 0.5, 0.25}}, {1, {-0.5, 1.0, -0.5}}}},
};

static const std::map<std::string,
                      std::vector<std::pair<int, std::array<double, 3>>>>
    lifting_wavelet_reconstruct_offsets_and_coefficients = {
        {"hat", {{1, {0.5, 1.0, 0.5}}}},
        {"biorthogonal", {{0, {-0.25, 1.0, -0.25}}, {1, {0.5, 1.0, 0.5}}}},
        {"fullweighting", {{1, {0.5, 1.0, 0.5}}, {0, {-0.5, 2.0, -0.5}}}},
};

} // namespace paliwa
