// Times the reference's per-grid loop over a combination-technique set in C++
// through the drop-in header, beside the batched plan, on the same grids.
//
// The set: every 4-D level vector with l_i >= 1 and |l|_1 in {13..16} (1325
// grids, extent 2^l_i + 1 per axis, float64), minimum level 0, maximum level
// 13.  The loop is the reference's
//   for each grid: paliwa::hierarchize(grid, level, minimum_level,
//                                      maximum_level, "hat",
//                                      instances[grid % instances.size()]);
//   fence all instances
// (test/test_paliwa_combination_technique.cpp:164-174, instances from
// Kokkos::Experimental::partition_space, at most 32: :358-366), here with HIP
// streams of default priority for the instances: one stream, and 32 streams
// round-robin.  The batched plan (paliwa_batch_*) runs on the same data;
// both results are compared bit for bit.  Prints one JSON line.
//   ct_per_grid [reps]
#include <paliwa/paliwa_transform.hpp>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#define HIP_OK(x)                                                              \
  do {                                                                         \
    hipError_t const e_ = (x);                                                 \
    if (e_ != hipSuccess) {                                                    \
      std::fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_),    \
                   __FILE__, __LINE__);                                        \
      std::exit(1);                                                            \
    }                                                                          \
  } while (0)

namespace {

constexpr int D = 4;
using Span = paliwa::GridSpan<double, D>;

struct Grid {
  std::array<long, D> level;
  std::array<std::int64_t, D> shape;
  std::int64_t elements;
  std::size_t offset; // into the shared allocation, in elements
};

std::vector<Grid> ct_set() {
  std::vector<Grid> out;
  std::size_t off = 0;
  for (int sum = 13; sum <= 16; ++sum)
    for (long a = 1; a <= sum; ++a)
      for (long b = 1; a + b <= sum; ++b)
        for (long c = 1; a + b + c < sum; ++c) {
          long const d = sum - a - b - c;
          Grid g{{a, b, c, d}, {}, 1, off};
          for (int k = 0; k < D; ++k) {
            g.shape[k] = (std::int64_t(1) << g.level[k]) + 1;
            g.elements *= g.shape[k];
          }
          off += std::size_t(g.elements);
          out.push_back(g);
        }
  return out;
}

// ms per repetition of fn over `reps` repetitions after one warm-up, on `s`
template <typename F> double timed(hipStream_t s, int reps, F &&fn) {
  fn();
  HIP_OK(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  HIP_OK(hipEventRecord(e0, s));
  for (int r = 0; r < reps; ++r)
    fn();
  HIP_OK(hipEventRecord(e1, s));
  HIP_OK(hipDeviceSynchronize());
  float ms = 0.0f;
  HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  HIP_OK(hipEventDestroy(e0));
  HIP_OK(hipEventDestroy(e1));
  return ms / reps;
}

} // namespace

int main(int argc, char **argv) {
  int const reps = argc > 1 ? std::max(1, std::atoi(argv[1])) : 3;
  std::vector<Grid> const grids = ct_set();
  std::size_t total = 0;
  for (auto const &g : grids)
    total += std::size_t(g.elements);

  // deterministic values (a hash of the index), copied to two device buffers
  std::vector<double> host(total);
  for (std::size_t i = 0; i < total; ++i) {
    std::uint64_t z = (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    host[i] = double(z >> 11) * 0x1.0p-53 * 2.0 - 1.0;
  }
  double *loop_data = nullptr, *batch_data = nullptr;
  HIP_OK(hipMalloc(&loop_data, total * sizeof(double)));
  HIP_OK(hipMalloc(&batch_data, total * sizeof(double)));
  HIP_OK(hipMemcpy(loop_data, host.data(), total * sizeof(double), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(batch_data, host.data(), total * sizeof(double), hipMemcpyHostToDevice));

  std::array<long, D> const lmin{0, 0, 0, 0}, lmax{13, 13, 13, 13};
  std::vector<hipStream_t> instances(32);
  for (auto &s : instances)
    HIP_OK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipStream_t const main_stream = instances[0];

  auto loop = [&](double *base, std::size_t n_instances) {
    for (std::size_t i = 0; i < grids.size(); ++i) {
      Span span{base + grids[i].offset, grids[i].shape};
      paliwa::hierarchize(span, grids[i].level, lmin, lmax, "hat",
                          paliwa::HipExec{instances[i % n_instances]});
    }
    for (std::size_t k = 1; k < n_instances; ++k) { // fence all instances
      hipEvent_t e;
      HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      HIP_OK(hipEventRecord(e, instances[k]));
      HIP_OK(hipStreamWaitEvent(main_stream, e, 0));
      HIP_OK(hipEventDestroy(e));
    }
  };
  double const one_stream_ms = timed(main_stream, reps, [&] { loop(loop_data, 1); });
  double const streams_ms = timed(main_stream, reps, [&] { loop(loop_data, 32); });

  std::vector<void *> ptrs;
  std::vector<std::int64_t> ext;
  std::vector<std::int32_t> lv;
  for (auto const &g : grids) {
    ptrs.push_back(batch_data + g.offset);
    for (int k = 0; k < D; ++k) {
      ext.push_back(g.shape[k]);
      lv.push_back(std::int32_t(g.level[k]));
    }
  }
  std::int32_t const mn[D] = {0, 0, 0, 0}, mx[D] = {13, 13, 13, 13};
  paliwa_batch *plan = nullptr;
  if (paliwa_batch_create(&plan, PALIWA_FLOAT64, D, std::int32_t(grids.size()), ptrs.data(),
                          ext.data(), lv.data(), mn, mx) != PALIWA_OK) {
    std::fprintf(stderr, "batch plan: %s\n", paliwa_last_error());
    return 1;
  }
  double const batched_ms = timed(main_stream, reps + 2, [&] {
    if (paliwa_batch_hierarchize(plan, "hat", main_stream) != PALIWA_OK) {
      std::fprintf(stderr, "batch: %s\n", paliwa_last_error());
      std::exit(1);
    }
  });

  // equal results: the same number of forward transforms on both copies
  // (warm-up + reps, twice for the loop, once + 2 more for the plan), so
  // compare fresh ones: reset both and transform once each
  HIP_OK(hipMemcpy(loop_data, host.data(), total * sizeof(double), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(batch_data, host.data(), total * sizeof(double), hipMemcpyHostToDevice));
  loop(loop_data, 32);
  if (paliwa_batch_hierarchize(plan, "hat", main_stream) != PALIWA_OK)
    return 1;
  HIP_OK(hipDeviceSynchronize());
  std::vector<double> a(total), b(total);
  HIP_OK(hipMemcpy(a.data(), loop_data, total * sizeof(double), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(b.data(), batch_data, total * sizeof(double), hipMemcpyDeviceToHost));
  bool const identical = std::memcmp(a.data(), b.data(), total * sizeof(double)) == 0;

  std::printf("{\"grids\": %zu, \"points\": %zu, \"per_grid_one_stream_ms\": %.4f, "
              "\"per_grid_32_streams_ms\": %.4f, \"batched_ms\": %.4f, "
              "\"speedup_vs_one_stream\": %.2f, \"speedup_vs_32_streams\": %.2f, "
              "\"per_grid_launches\": %zu, \"bit_identical\": %s}\n",
              grids.size(), total, one_stream_ms, streams_ms, batched_ms,
              one_stream_ms / batched_ms, streams_ms / batched_ms, 4 * grids.size(),
              identical ? "true" : "false");
  paliwa_batch_destroy(plan);
  for (auto s : instances)
    HIP_OK(hipStreamDestroy(s));
  HIP_OK(hipFree(loop_data));
  HIP_OK(hipFree(batch_data));
  return identical ? 0 : 2;
}
