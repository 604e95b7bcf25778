// Exercises the C++ drop-in header include/paliwa/paliwa_transform.hpp the way
// the reference's own tests call paliwa (test/test_paliwa_transform.cpp):
// hierarchize / dehierarchize / *_in / transform_in on device grids, checked
// bit-for-bit against the CPU oracle (test infrastructure), plus the exception
// types of the reference.
//   test_dropin --host-only   exception checks that fail validation before any
//                             HIP call (runs without a GPU)
//   test_dropin               everything (needs a GPU)
#include <paliwa/paliwa_domains.hpp>
#include <paliwa/paliwa_io.hpp>
#include <paliwa/paliwa_transform.hpp>

#include "../../oracle/paliwa_oracle.hpp"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <cstring>
#include <random>
#include <vector>

#define CHECK(cond)                                                            \
  do {                                                                         \
    if (!(cond)) {                                                             \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);   \
      std::exit(1);                                                            \
    }                                                                          \
  } while (0)

#define HIP_CHECK(x) CHECK((x) == hipSuccess)

template <typename E, typename F> bool throws(F &&f) {
  try {
    f();
  } catch (E const &) {
    return true;
  } catch (...) {
    return false;
  }
  return false;
}

// test/test_paliwa_plan.cpp:199-221 through the drop-in planner header
static void planner_checks() {
  using Filters = std::vector<std::pair<int, std::array<double, 3>>>;
  Filters const one{{1, {0., 1., 0.}}};
  auto ignore = [](std::size_t, std::size_t, paliwa::PeriodicBounds) {};
  for (auto const &levels : std::vector<std::vector<long>>{{4, 1}, {1, 4}, {1, 2, 1}})
    CHECK(throws<std::invalid_argument>(
        [&] { paliwa::transform_bounds({0, 0}, 4, 4, levels, one, ignore); }));
  auto cover = [](std::vector<long> idx, long mn, long mx) {
    auto b = paliwa::covering_periodic_bounds(idx, 64);
    CHECK(b.min == mn && b.max == mx);
  };
  cover({}, -1, -1);
  cover({12}, 12, 12);
  cover({8, 12, 16}, 8, 16);
  cover({0, 4, 56, 60}, 56, 4);
  cover({0, 16, 32, 48}, 0, 48);
  // writes are visited backwards; a hat pass writes the odd outputs
  std::vector<std::pair<std::size_t, long>> seen;
  auto req = paliwa::transform_bounds(
      {4, 12}, 4, 4, std::vector<long>{4}, Filters{{1, {-.5, 1., -.5}}},
      [&](std::size_t i, std::size_t, paliwa::PeriodicBounds w) {
        seen.push_back({i, w.min});
        CHECK(w.min == 5 && w.max == 11);
      });
  CHECK(seen.size() == 1 && req.size() == 5 && req[0].empty());
  CHECK(req[2].min == 4 && req[2].max == 12 && req[4].min == 5 && req[4].max == 11);
}

static void host_only_checks() {
  planner_checks();
  // Never dereferenced: every call below fails argument validation.
  paliwa::GridSpan<double, 2> g{reinterpret_cast<double *>(0x1000), {17, 16}};
  paliwa::LevelVector<2> lv{4, 4}, mn{0, 0}, mx{4, 4};
  CHECK(throws<std::out_of_range>([&] { paliwa::hierarchize(g, lv, mn, mx, "haar", paliwa::HipExec{}); }));
  CHECK(throws<std::out_of_range>([&] { paliwa::dehierarchize_in<1>(g, lv, mn, mx, "db2", paliwa::HipExec{}); }));
  std::vector<std::pair<int, std::array<double, 3>>> bad{{2, {0.0, 1.0, 0.0}}};
  std::vector<long> levels{4, 3};
  bool got = false;
  try {
    paliwa::transform_in<0>(g, lv, mx, levels, bad, paliwa::HipExec{});
  } catch (std::runtime_error const &e) {
    got = std::strcmp(e.what(), "Filter offset not supported") == 0;
  }
  CHECK(got);
  paliwa::LevelVector<2> big{5, 0};
  CHECK(throws<std::invalid_argument>([&] { paliwa::hierarchize(g, lv, big, mx, "hat", paliwa::HipExec{}); }));
  paliwa::GridSpan<double, 2> wrong{reinterpret_cast<double *>(0x1000), {18, 16}};
  CHECK(throws<std::invalid_argument>([&] { paliwa::hierarchize(wrong, lv, mn, mx, "hat", paliwa::HipExec{}); }));
  CHECK(paliwa::lifting_wavelet_filter_offsets_and_coefficients.at("biorthogonal").size() == 2);
  // a partial buffer whose coordinates leave the period is refused on the host
  paliwa::PartialSpan<double, 2> part{reinterpret_cast<double *>(0x1000),
                                      {paliwa_axis_storage{8, 10, 1, nullptr},
                                       paliwa_axis_storage{4, 0, 1, nullptr}}};
  CHECK(throws<std::invalid_argument>(
      [&] { paliwa::hierarchize(part, lv, mn, mx, "hat", paliwa::HipExec{}); }));
  std::printf("host-only checks OK\n");
}

template <typename T>
static void grid_checks(hipStream_t stream) {
  // 3-D grid with images on two axes, the reference's level/minimum-level
  // arguments, every wavelet, forward then inverse.
  constexpr std::size_t D = 3;
  paliwa::LevelVector<D> lv{5, 3, 6}, mn{1, 0, 2}, mx{6, 3, 7};
  std::array<int64_t, D> ext{33, 8, 65};
  std::size_t const n = ext[0] * ext[1] * ext[2];
  std::mt19937 rng(20260928);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  std::vector<T> host(n);
  for (auto &x : host)
    x = static_cast<T>(u(rng));
  // close the images (index 2^l is a copy of index 0)
  for (int64_t i = 0; i < ext[0]; ++i)
    for (int64_t j = 0; j < ext[1]; ++j)
      host[(i * ext[1] + j) * ext[2] + 64] = host[(i * ext[1] + j) * ext[2]];
  for (int64_t j = 0; j < ext[1]; ++j)
    for (int64_t k = 0; k < ext[2]; ++k)
      host[(32 * ext[1] + j) * ext[2] + k] = host[j * ext[2] + k];
  T *dev = nullptr;
  HIP_CHECK(hipMalloc(&dev, n * sizeof(T)));
  paliwa::GridSpan<T, D> grid{dev, ext};
  std::vector<int> level(lv.begin(), lv.end()), lmin(mn.begin(), mn.end()),
      lmax(mx.begin(), mx.end());
  std::vector<int64_t> extents(ext.begin(), ext.end());
  for (char const *name : {"hat", "biorthogonal", "fullweighting"}) {
    HIP_CHECK(hipMemcpyAsync(dev, host.data(), n * sizeof(T), hipMemcpyHostToDevice, stream));
    paliwa::hierarchize(grid, lv, mn, mx, name, paliwa::HipExec{stream});
    std::vector<T> got(n), want = host;
    HIP_CHECK(hipMemcpyAsync(got.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    paliwa_oracle::hierarchize(want.data(), extents, level, lmin, lmax, name);
    CHECK(std::memcmp(got.data(), want.data(), n * sizeof(T)) == 0);
    // inverse axis by axis through the *_in entry points
    paliwa::dehierarchize_in<0>(grid, lv, mn, mx, name, paliwa::HipExec{stream});
    paliwa::dehierarchize_in<1>(grid, lv, mn, mx, name, paliwa::HipExec{stream});
    paliwa::dehierarchize_in<2>(grid, lv, mn, mx, name, paliwa::HipExec{stream});
    HIP_CHECK(hipMemcpyAsync(got.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    paliwa_oracle::dehierarchize(want.data(), extents, level, lmin, lmax, name);
    CHECK(std::memcmp(got.data(), want.data(), n * sizeof(T)) == 0);
    double err = 0.0;
    for (std::size_t i = 0; i < n; ++i)
      err = std::fmax(err, std::fabs(double(got[i]) - double(host[i])));
    CHECK(err < (sizeof(T) == 8 ? 1e-14 : 3e-6));
  }
  // transform_in with a custom 3-pass filter list and a non-consecutive
  // level list (the generic schedule)
  std::vector<std::pair<int, std::array<double, 3>>> filters{
      {1, {-0.5, 1.0, 0.125}}, {0, {0.0, 1.0, 0.25}}, {1, {0.3, 0.7, -0.1}}};
  std::vector<long> levels{6, 2, 4};
  HIP_CHECK(hipMemcpyAsync(dev, host.data(), n * sizeof(T), hipMemcpyHostToDevice, stream));
  paliwa::transform_in<2>(grid, lv, mx, levels, filters, paliwa::HipExec{stream});
  std::vector<T> got(n), want = host;
  HIP_CHECK(hipMemcpyAsync(got.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost, stream));
  HIP_CHECK(hipStreamSynchronize(stream));
  paliwa_oracle::transform_in(want.data(), extents, level, 2, {6, 2, 4}, filters);
  CHECK(std::memcmp(got.data(), want.data(), n * sizeof(T)) == 0);
  // partial outputs (paliwa/paliwa_transform.hpp:331-344) on a grid without
  // images along the axis: only the requested run changes, bit-exactly
  {
    constexpr std::size_t D2 = 2;
    paliwa::LevelVector<D2> lv2{6, 4}, mx2{7, 4};
    std::array<int64_t, D2> e2{64, 16};
    std::size_t const n2 = 64 * 16;
    std::vector<T> h2(n2);
    for (auto &x : h2)
      x = static_cast<T>(u(rng));
    T *d2 = nullptr;
    HIP_CHECK(hipMalloc(&d2, n2 * sizeof(T)));
    paliwa::GridSpan<T, D2> g2{d2, e2};
    std::vector<std::pair<int, std::array<double, 3>>> bio{{1, {-0.5, 1.0, -0.5}},
                                                           {0, {0.25, 1.0, 0.25}}};
    std::vector<long> lv_list{6, 5, 4, 3, 2, 1};
    // {120, 8} twice: the second call reuses the plan the first one made
    for (auto [start, count] :
         std::vector<std::pair<long, long>>{{120, 8}, {34, 5}, {0, 64}, {120, 8}}) {
      HIP_CHECK(hipMemcpyAsync(d2, h2.data(), n2 * sizeof(T), hipMemcpyHostToDevice, stream));
      paliwa::transform_in<0>(g2, lv2, mx2, lv_list, bio, paliwa::AxisRun{start, count},
                              paliwa::HipExec{stream});
      std::vector<T> got2(n2), want2 = h2;
      HIP_CHECK(hipMemcpyAsync(got2.data(), d2, n2 * sizeof(T), hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      paliwa_oracle::transform_in(want2.data(), {64, 16}, {6, 4}, 0, {6, 5, 4, 3, 2, 1}, bio);
      for (long i = 0; i < 64; ++i) {
        bool const out = ((i - start / 2) % 64 + 64) % 64 < count;
        for (long j = 0; j < 16; ++j) {
          T const w = out ? want2[i * 16 + j] : h2[i * 16 + j];
          CHECK(std::memcmp(&got2[i * 16 + j], &w, sizeof(T)) == 0);
        }
      }
    }
    CHECK(throws<std::invalid_argument>([&] {
      paliwa::transform_in<0>(g2, lv2, mx2, std::vector<long>{6, 4}, bio,
                              paliwa::AxisRun{0, 4}, paliwa::HipExec{stream});
    }));
    HIP_CHECK(hipFree(d2));
  }
  // a partial buffer through the same names: a sub-box of a grid stored at
  // its maximum level along axis 0 (a DDC DiscreteDomain block) and a strided
  // run along axis 1, against the oracle's checked branch
  {
    constexpr std::size_t D3 = 2;
    paliwa::LevelVector<D3> lv3{6, 4}, mn3{1, 0}, mx3{6, 5};
    std::vector<std::vector<int64_t>> coords(2);
    for (int64_t x = 5; x < 5 + 40; ++x)
      coords[0].push_back(x);
    for (int64_t y = 2; y < 2 + 2 * 12; y += 2)
      coords[1].push_back(y);
    std::size_t const n3 = coords[0].size() * coords[1].size();
    std::vector<T> h3(n3);
    for (auto &x : h3)
      x = static_cast<T>(u(rng));
    T *d3 = nullptr;
    HIP_CHECK(hipMalloc(&d3, n3 * sizeof(T)));
    paliwa::PartialSpan<T, D3> part{
        d3, {paliwa_axis_storage{int64_t(coords[0].size()), 5, 1, nullptr},
             paliwa_axis_storage{int64_t(coords[1].size()), 2, 2, nullptr}}};
    for (char const *name : {"hat", "biorthogonal", "fullweighting"}) {
      HIP_CHECK(hipMemcpyAsync(d3, h3.data(), n3 * sizeof(T), hipMemcpyHostToDevice, stream));
      paliwa::hierarchize(part, lv3, mn3, mx3, name, paliwa::HipExec{stream});
      std::vector<T> got3(n3), want3 = h3;
      HIP_CHECK(hipMemcpyAsync(got3.data(), d3, n3 * sizeof(T), hipMemcpyDeviceToHost, stream));
      HIP_CHECK(hipStreamSynchronize(stream));
      auto const &table = paliwa_oracle::forward_filters().at(name);
      paliwa_oracle::transform_in_checked(want3.data(), coords, {6, 4}, {6, 5}, 0,
                                          paliwa_oracle::decreasing_levels(6, 1), table);
      paliwa_oracle::transform_in_checked(want3.data(), coords, {6, 4}, {6, 5}, 1,
                                          paliwa_oracle::decreasing_levels(4, 0), table);
      CHECK(std::memcmp(got3.data(), want3.data(), n3 * sizeof(T)) == 0);
    }
    HIP_CHECK(hipFree(d3));
  }
  // raw dump (paliwa/paliwa_io.hpp): the elements, row-major, headerless
  char const *tmp = std::getenv("TMPDIR");
  std::string const path = std::string(tmp != nullptr ? tmp : "/tmp") + "/paliwa_dropin_dump.raw";
  paliwa::dump_chunk_span_to_binary_file(grid, path, paliwa::HipExec{stream});
  std::vector<T> file(n);
  std::FILE *f = std::fopen(path.c_str(), "rb");
  CHECK(f != nullptr);
  CHECK(std::fread(file.data(), sizeof(T), n, f) == n);
  CHECK(std::fgetc(f) == EOF);
  std::fclose(f);
  std::remove(path.c_str());
  CHECK(std::memcmp(file.data(), want.data(), n * sizeof(T)) == 0);
  CHECK(throws<std::runtime_error>(
      [&] { paliwa::dump_chunk_span_to_binary_file(grid, "/nonexistent-dir/x.raw", paliwa::HipExec{stream}); }));
  HIP_CHECK(hipFree(dev));
}

int main(int argc, char **argv) {
  host_only_checks();
  if (argc > 1 && std::strcmp(argv[1], "--host-only") == 0)
    return 0;
  hipStream_t stream;
  HIP_CHECK(hipStreamCreate(&stream));
  grid_checks<double>(stream);
  grid_checks<float>(stream);
  HIP_CHECK(hipStreamDestroy(stream));
  std::printf("drop-in header checks OK\n");
  return 0;
}
