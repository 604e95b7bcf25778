// Host sanitizer run of the CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// Built together with oracle_capi.cpp under AddressSanitizer and
// UndefinedBehaviorSanitizer (oracle/Makefile, target `sanitize`) and run by
// tests/test_oracle_sanitized.py: every entry point of the oracle's C ABI on
// small grids -- the named transforms in both directions with every schedule,
// both dtypes, extents 2^l and 2^l + 1, generic level lists, partial
// buffers -- plus the argument errors, so that an out-of-bounds access or
// undefined behaviour in the oracle fails the test.  Prints "sanitizer checks
// OK" and exits 0 when every round trip closes.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

extern "C" {
int oracle_transform(void *data, int dtype, int ndim, int64_t const *extents,
                     int32_t const *level, int32_t const *lmin, int32_t const *lmax,
                     char const *name, int forward, int schedule);
int oracle_transform_in(void *data, int dtype, int ndim, int64_t const *extents,
                        int32_t const *level, int dim, int32_t const *levels, int nlevels,
                        int32_t const *offsets, double const *coefficients, int nfilters,
                        int schedule);
int oracle_transform_in_checked(void *data, int dtype, int ndim, int64_t const *counts,
                                int64_t const *coords, int32_t const *level,
                                int32_t const *maximum_level, int dim, int32_t const *levels,
                                int nlevels, int32_t const *offsets,
                                double const *coefficients, int nfilters);
char const *oracle_last_error(void);
}

namespace {

int failures = 0;

void expect(bool ok, char const *what) {
  if (!ok) {
    std::printf("FAILED: %s (%s)\n", what, oracle_last_error());
    ++failures;
  }
}

template <typename T>
void round_trips(std::mt19937 &rng) {
  int const dtype = sizeof(T) == 8 ? 0 : 1;
  double const tol = sizeof(T) == 8 ? 1e-12 : 1e-4;
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  int const shapes[][3] = {{3, 2, 4}, {1, 5, 2}, {4, 4, 1}, {2, 1, 3}};
  char const *names[] = {"hat", "biorthogonal", "fullweighting"};
  for (auto const &lv : shapes)
    for (int image = 0; image < 2; ++image)
      for (char const *name : names)
        for (int schedule = 0; schedule < 3; ++schedule) {
          int64_t ext[3];
          int64_t n = 1;
          for (int d = 0; d < 3; ++d) {
            ext[d] = (int64_t(1) << lv[d]) + image;
            n *= ext[d];
          }
          std::vector<T> g(static_cast<size_t>(n));
          for (auto &x : g)
            x = static_cast<T>(u(rng));
          std::vector<T> const start = g;
          int32_t const level[3] = {lv[0], lv[1], lv[2]};
          int32_t const lmin[3] = {0, lv[1] > 0 ? 1 : 0, 0};
          int32_t const lmax[3] = {6, 6, 6};
          expect(oracle_transform(g.data(), dtype, 3, ext, level, lmin, lmax, name, 1,
                                  schedule) == 0,
                 "forward");
          expect(oracle_transform(g.data(), dtype, 3, ext, level, lmin, lmax, name, 0,
                                  schedule) == 0,
                 "inverse");
          // the image planes are rewritten from index 0: compare the lattice
          double err = 0.0;
          for (int64_t i = 0; i < n; ++i) {
            int64_t r = i;
            bool lattice = true;
            for (int d = 2; d >= 0; --d) {
              lattice = lattice && (r % ext[d]) < (int64_t(1) << lv[d]);
              r /= ext[d];
            }
            if (lattice)
              err = std::fmax(err, std::fabs(double(g[i]) - double(start[i])));
          }
          expect(err <= tol, "round trip");
        }
}

void generic_levels(std::mt19937 &rng) {
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  int64_t const ext[2] = {17, 9};
  int32_t const level[2] = {4, 3};
  std::vector<double> g(17 * 9);
  for (auto &x : g)
    x = u(rng);
  int32_t const levels0[] = {4, 2, 2, 1, 3}; // repeated and out of order
  int32_t const levels1[] = {3, 1, 2};
  int32_t const offsets[] = {1, 0, 1};
  double const coefs[] = {-0.5, 1.0, -0.5, 0.25, 1.0, 0.25, 0.125, 0.75, 0.125};
  for (int schedule = 0; schedule < 3; ++schedule)
    for (int dim = 0; dim < 2; ++dim)
      expect(oracle_transform_in(g.data(), 0, 2, ext, level, dim, dim == 0 ? levels0 : levels1,
                                 dim == 0 ? 5 : 3, offsets, coefs, 3, schedule) == 0,
             "transform_in");
  for (double x : g)
    expect(std::isfinite(x), "finite");
}

void partial_buffers(std::mt19937 &rng) {
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  // axis 0: 5 of the 16 coordinates of level 4, axis 1: all 8 of level 3
  int64_t const counts[2] = {5, 8};
  int64_t const coords[] = {0, 3, 4, 8, 15, 0, 1, 2, 3, 4, 5, 6, 7};
  int32_t const level[2] = {4, 3};
  int32_t const maximum[2] = {4, 3};
  std::vector<float> g(5 * 8);
  for (auto &x : g)
    x = static_cast<float>(u(rng));
  int32_t const levels0[] = {4, 3, 2, 1};
  int32_t const levels1[] = {3, 2, 1};
  int32_t const offsets[] = {1, 0};
  double const coefs[] = {-0.5, 1.0, -0.5, 0.25, 1.0, 0.25};
  for (int dim = 0; dim < 2; ++dim)
    expect(oracle_transform_in_checked(g.data(), 1, 2, counts, coords, level, maximum, dim,
                                       dim == 0 ? levels0 : levels1, dim == 0 ? 4 : 3,
                                       offsets, coefs, 2) == 0,
           "transform_in_checked");
}

void argument_errors() {
  double g[4] = {0, 0, 0, 0};
  int64_t const ext[1] = {4};
  int64_t const bad_ext[1] = {5 + 2};
  int32_t const level[1] = {2}, lmin[1] = {0}, lmax[1] = {2};
  expect(oracle_transform(g, 0, 1, ext, level, lmin, lmax, "haar", 1, 0) != 0, "unknown name");
  expect(oracle_transform(g, 0, 1, bad_ext, level, lmin, lmax, "hat", 1, 0) != 0, "bad extent");
  expect(oracle_transform(g, 7, 1, ext, level, lmin, lmax, "hat", 1, 0) != 0, "bad dtype");
  int32_t const levels[] = {3};
  int32_t const offsets[] = {2};
  double const coefs[] = {0, 1, 0};
  expect(oracle_transform_in(g, 0, 1, ext, level, 0, levels, 1, offsets, coefs, 1, 0) != 0,
         "bad level and offset");
}

} // namespace

int main() {
  std::mt19937 rng(7);
  round_trips<double>(rng);
  round_trips<float>(rng);
  generic_levels(rng);
  partial_buffers(rng);
  argument_errors();
  if (failures != 0) {
    std::printf("%d sanitizer-run checks failed\n", failures);
    return 1;
  }
  std::printf("sanitizer checks OK\n");
  return 0;
}
