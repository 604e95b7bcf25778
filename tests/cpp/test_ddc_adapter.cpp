// The DDC-typed overloads (include/paliwa/paliwa_ddc.hpp), compiled against
// the test double of DDC in tests/cpp/ddc_double: the domain of each DDC kind
// maps onto the intended per-axis storage (host-only checks), and through the
// adapter the GPU gives the oracle's results -- the checked branch for a
// DiscreteDomain block and a SparseDiscreteDomain of local points plus their
// dependencies, the ordinary transform for the complete strided lattice.
//   test_ddc_adapter --host-only | (no argument: everything, needs a GPU)
#include <paliwa/paliwa_ddc.hpp>

#include "../../oracle/paliwa_oracle.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
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

struct X {};
struct Y {};
using Vec = ddc::DiscreteVector<X, Y>;

static void host_only() {
  static_assert(paliwa::ddc_adapter::index_of<Y, X, Y>() == 1);
  static_assert(paliwa::ddc_adapter::index_of<X, X, Y>() == 0);
  std::vector<std::int64_t> list;
  ddc::DiscreteDomain<X, Y> box{ddc::DiscreteElement<X, Y>(5, 2), Vec(40, 12)};
  auto a = paliwa::ddc_adapter::axis<Y>(box, list);
  CHECK(a.count == 12 && a.front == 2 && a.step == 1 && a.coordinates == nullptr);
  ddc::StridedDiscreteDomain<X, Y> lat{ddc::DiscreteElement<X, Y>(0, 0), Vec(64, 16),
                                       Vec(2, 4)};
  a = paliwa::ddc_adapter::axis<X>(lat, list);
  CHECK(a.count == 64 && a.front == 0 && a.step == 2);
  ddc::SparseDiscreteDomain<X, Y> sp(ddc::SparseDiscreteDomain<X>(std::vector<long>{1, 4, 9}),
                                     ddc::SparseDiscreteDomain<Y>(std::vector<long>{0, 3}));
  a = paliwa::ddc_adapter::axis<Y>(sp, list);
  CHECK(a.count == 2 && list.size() == 2 && list[0] == 0 && list[1] == 3);
  auto const lv = paliwa::ddc_adapter::levels_of(Vec(6, 4));
  CHECK(lv[0] == 6 && lv[1] == 4);
  std::printf("adapter host-only checks OK\n");
}

template <class Dom>
static void run_case(Dom const &dom, std::vector<std::vector<int64_t>> const &coords,
                     hipStream_t stream, std::mt19937 &rng) {
  std::size_t const n = coords[0].size() * coords[1].size();
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  std::vector<double> h(n);
  for (auto &x : h)
    x = u(rng);
  double *d = nullptr;
  HIP_CHECK(hipMalloc(&d, n * sizeof(double)));
  ddc::ChunkSpan<double, Dom> span{d, dom};
  Vec const level(6, 4), lmin(1, 0), lmax(7, 4);
  for (char const *name : {"hat", "biorthogonal", "fullweighting"}) {
    HIP_CHECK(hipMemcpy(d, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
    paliwa::hierarchize(span, level, lmin, lmax, name, paliwa::HipExec{stream});
    paliwa::dehierarchize_in<Y>(span, level, lmin, lmax, name, paliwa::HipExec{stream});
    std::vector<double> got(n), want = h;
    HIP_CHECK(hipStreamSynchronize(stream));
    HIP_CHECK(hipMemcpy(got.data(), d, n * sizeof(double), hipMemcpyDeviceToHost));
    auto const &fwd = paliwa_oracle::forward_filters().at(name);
    auto const &inv = paliwa_oracle::reconstruct_filters().at(name);
    paliwa_oracle::transform_in_checked(want.data(), coords, {6, 4}, {7, 4}, 0,
                                        paliwa_oracle::decreasing_levels(6, 1), fwd);
    paliwa_oracle::transform_in_checked(want.data(), coords, {6, 4}, {7, 4}, 1,
                                        paliwa_oracle::decreasing_levels(4, 0), fwd);
    paliwa_oracle::transform_in_checked(want.data(), coords, {6, 4}, {7, 4}, 1,
                                        paliwa_oracle::increasing_levels(4, 0), inv);
    CHECK(std::memcmp(got.data(), want.data(), n * sizeof(double)) == 0);
  }
  HIP_CHECK(hipFree(d));
}

int main(int argc, char **argv) {
  host_only();
  if (argc > 1 && std::strcmp(argv[1], "--host-only") == 0)
    return 0;
  hipStream_t stream;
  HIP_CHECK(hipStreamCreate(&stream));
  std::mt19937 rng(5);
  // a DiscreteDomain block at the finest resolution along X (maximum level 7)
  {
    std::vector<std::vector<int64_t>> c(2);
    for (int64_t x = 9; x < 9 + 50; ++x)
      c[0].push_back(x);
    for (int64_t y = 0; y < 16; ++y)
      c[1].push_back(y);
    run_case(ddc::DiscreteDomain<X, Y>{ddc::DiscreteElement<X, Y>(9, 0), Vec(50, 16)}, c,
             stream, rng);
  }
  // the complete strided lattice (level 6 of maximum level 7 along X)
  {
    std::vector<std::vector<int64_t>> c(2);
    for (int64_t x = 0; x < 128; x += 2)
      c[0].push_back(x);
    for (int64_t y = 0; y < 16; ++y)
      c[1].push_back(y);
    run_case(ddc::StridedDiscreteDomain<X, Y>{ddc::DiscreteElement<X, Y>(0, 0), Vec(64, 16),
                                              Vec(2, 1)},
             c, stream, rng);
  }
  // a sparse product of irregular coordinate lists
  {
    std::vector<long> xs, ys;
    for (long x = 0; x < 128; x += 2)
      if (rng() % 4 != 0)
        xs.push_back(x);
    for (long y = 0; y < 16; ++y)
      if (rng() % 3 != 0)
        ys.push_back(y);
    std::vector<std::vector<int64_t>> c{std::vector<int64_t>(xs.begin(), xs.end()),
                                        std::vector<int64_t>(ys.begin(), ys.end())};
    run_case(ddc::SparseDiscreteDomain<X, Y>(ddc::SparseDiscreteDomain<X>(xs),
                                             ddc::SparseDiscreteDomain<Y>(ys)),
             c, stream, rng);
  }
  HIP_CHECK(hipStreamDestroy(stream));
  std::printf("ddc adapter checks OK\n");
  return 0;
}
