// HIP kernels of the lifting transform for gfx950 (MI355X) other than the
// fused sweep (fused_sweep.hpp), and the per-axis launch dispatch.
//
// One axis sweep applies, to every line along the axis, every level of a
// level list and every pass of a filter list -- the work the reference issues
// as one Kokkos launch per (level, pass) in detail::lifting_pass
// (paliwa/paliwa_transform.hpp:24-51, 215-251).  The update of a point x of
// parity `offset` at storage stride s, period n, is
//     g[x] = c0*g[(x-s) mod n] + c1*g[x] + c2*g[(x+s) mod n]
// evaluated left to right with no fused multiply-add (-ffp-contract=off), so
// every result is bit-identical to the reference's.  Within one pass the
// written points and the points read have opposite parity, so a pass is a
// pure function of the state before it and may visit its points in any order.
//
// Kernels here:
//   lds_line_sweep  one whole line per workgroup in LDS, one barrier per pass;
//                   for lines longer than the fused kernel holds in registers.
//   per_pass_sweep  one launch per (level, pass) directly on HBM, the launch
//                   structure of the reference; for any length.
#include "fused_sweep.hpp"
#include "lifting.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>

namespace paliwa_hip {
namespace {

constexpr int kThreads = 256;

template <typename T>
__device__ __forceinline__ T lift(T c0, T lo, T c1, T x, T c2, T hi) {
  return c0 * lo + c1 * x + c2 * hi;
}

template <typename Kernel>
hipError_t allow_lds(Kernel kernel, size_t bytes) {
  if (bytes <= 65536)
    return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(bytes));
}

// ---------------------------------------------------------------------------
// lds_line_sweep: one line per workgroup (grid-stride over lines).

// Every level and pass of the sweep on one line held in LDS; `base` points at
// the line's first point, consecutive points are `B` elements apart.
template <typename T>
__device__ __forceinline__ void sweep_line(SweepParams const &p, T *base,
                                           int64_t B, T *line) {
  int const tid = threadIdx.x;
  int const m = p.m;
  int const n = 1 << p.L;
  for (int r = tid; r < n; r += kThreads)
    line[r] = base[r * B];
  __syncthreads();
  for (int l = 0; l < p.nlevels; ++l) {
    int const s = 1 << p.shift[l];
    for (int f = 0; f < p.npass; ++f) {
      int const odd = p.offset[f];
      T const c0 = static_cast<T>(p.coef[f][0]);
      T const c1 = static_cast<T>(p.coef[f][1]);
      T const c2 = static_cast<T>(p.coef[f][2]);
      int const count = (n - odd * s + 2 * s - 1) / (2 * s);
      for (int k = tid; k < count; k += kThreads) {
        int const x = odd * s + 2 * s * k;
        int const lo = x == 0 ? n - s : x - s;
        int const hi = x + s == n ? 0 : x + s;
        line[x] = lift(c0, line[lo], c1, line[x], c2, line[hi]);
      }
      __syncthreads();
    }
  }
  for (int r = tid; r < n; r += kThreads)
    base[r * B] = line[r];
  if (m != n && tid == 0)
    base[static_cast<int64_t>(n) * B] = line[0];
  __syncthreads();
}

template <typename T>
__global__ __launch_bounds__(kThreads) void lds_line_sweep(SweepParams p) {
  extern __shared__ unsigned char smem_raw[];
  T *const line = reinterpret_cast<T *>(smem_raw);
  T *const g = static_cast<T *>(p.data);
  int64_t const lines = p.A * p.B;
  // consecutive lines (neighbouring columns of a strided axis, which share
  // every cache line) run on one XCD and meet in its L2
  for (int64_t ln = fused::xcd_tile(blockIdx.x, gridDim.x); ln < lines; ln += gridDim.x) {
    int64_t const a = ln / p.B;
    int64_t const b = ln - a * p.B;
    sweep_line(p, g + a * p.m * p.B + b, p.B, line);
  }
}

// The lines of many grids with equal period and stored extent along the axis
// in one launch (a batched plan's long-line groups): line t of the launch is
// line t - tile_base of grid tile_grid[t].
template <typename T>
__global__ __launch_bounds__(kThreads) void lds_line_batched(SweepParams p,
                                                             ManyGrids grids,
                                                             int64_t lines) {
  extern __shared__ unsigned char smem_raw[];
  T *const line = reinterpret_cast<T *>(smem_raw);
  for (int64_t t = fused::xcd_tile(blockIdx.x, gridDim.x); t < lines; t += gridDim.x) {
    TileRef const ref = grids(t);
    int64_t const a = ref.local / ref.B;
    int64_t const b = ref.local - a * ref.B;
    sweep_line(p, static_cast<T *>(ref.data) + a * p.m * ref.B + b, ref.B, line);
  }
}

// ---------------------------------------------------------------------------
// per_pass_sweep: the reference's schedule, one launch per (level, pass).

template <typename T>
__global__ __launch_bounds__(kThreads) void per_pass_sweep(
    T *g, int64_t A, int64_t B, int64_t m, int64_t n, int64_t s, int odd,
    int64_t count, T c0, T c1, T c2) {
  int64_t const total = A * count * B;
  for (int64_t idx = blockIdx.x * int64_t(kThreads) + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * kThreads) {
    int64_t const b = idx % B;
    int64_t const t = idx / B;
    int64_t const k = t % count;
    int64_t const a = t / count;
    int64_t const x = odd * s + 2 * s * k;
    int64_t const lo = x == 0 ? n - s : x - s;
    int64_t const hi = x + s == n ? 0 : x + s;
    T *const blk = g + a * m * B + b;
    blk[x * B] = lift(c0, blk[lo * B], c1, blk[x * B], c2, blk[hi * B]);
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void copy_image_plane(T *g, int64_t A,
                                                             int64_t B,
                                                             int64_t m,
                                                             int64_t n) {
  int64_t const total = A * B;
  for (int64_t idx = blockIdx.x * int64_t(kThreads) + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * kThreads) {
    int64_t const a = idx / B;
    int64_t const b = idx - a * B;
    T *const blk = g + a * m * B + b;
    blk[n * B] = blk[0];
  }
}

uint32_t grid_for(int64_t work) {
  int64_t const blocks = (work + kThreads - 1) / kThreads;
  return static_cast<uint32_t>(
      std::max<int64_t>(1, std::min<int64_t>(blocks, 1 << 20)));
}

template <typename T>
hipError_t launch_per_pass_typed(SweepParams const &p, hipStream_t stream) {
  T *const g = static_cast<T *>(p.data);
  int64_t const n = int64_t(1) << p.L;
  for (int l = 0; l < p.nlevels; ++l) {
    int64_t const s = int64_t(1) << p.shift[l];
    for (int f = 0; f < p.npass; ++f) {
      int const odd = p.offset[f];
      int64_t const count = (n - odd * s + 2 * s - 1) / (2 * s);
      hipLaunchKernelGGL(per_pass_sweep<T>, dim3(grid_for(p.A * count * p.B)),
                         dim3(kThreads), 0, stream, g, p.A, p.B, int64_t(p.m),
                         n, s, odd, count, static_cast<T>(p.coef[f][0]),
                         static_cast<T>(p.coef[f][1]),
                         static_cast<T>(p.coef[f][2]));
      hipError_t const err = hipGetLastError();
      if (err != hipSuccess)
        return err;
    }
  }
  if (has_image(p)) {
    hipLaunchKernelGGL(copy_image_plane<T>, dim3(grid_for(p.A * p.B)),
                       dim3(kThreads), 0, stream, g, p.A, p.B, int64_t(p.m),
                       n);
    return hipGetLastError();
  }
  return hipSuccess;
}

template <typename T>
hipError_t launch_lds_line_typed(SweepParams const &p, hipStream_t stream) {
  size_t const bytes = sizeof(T) * (size_t(1) << p.L);
  auto kernel = lds_line_sweep<T>;
  hipError_t err = allow_lds(kernel, bytes);
  if (err != hipSuccess)
    return err;
  int64_t const lines = p.A * p.B;
  uint32_t const grid =
      static_cast<uint32_t>(std::min<int64_t>(lines, 1 << 20));
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads), bytes, stream, p);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_lds_line_batched_typed(SweepParams const &p, ManyGrids grids,
                                         int64_t lines, hipStream_t stream) {
  size_t const bytes = sizeof(T) * (size_t(1) << p.L);
  auto kernel = lds_line_batched<T>;
  hipError_t err = allow_lds(kernel, bytes);
  if (err != hipSuccess)
    return err;
  uint32_t const grid =
      static_cast<uint32_t>(std::min<int64_t>(lines, 1 << 20));
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads), bytes, stream, p,
                     grids, lines);
  return hipGetLastError();
}

} // namespace

void classify(SweepParams &p) {
  p.schedule = kScheduleGeneric;
  p.pattern = kPatNone;
  p.lo = p.hi = 0;
  if (p.npass == 1 && p.offset[0] == 1)
    p.pattern = kPatOdd;
  else if (p.npass == 2 && p.offset[0] == 1 && p.offset[1] == 0)
    p.pattern = kPatOddEven;
  else if (p.npass == 2 && p.offset[0] == 0 && p.offset[1] == 1)
    p.pattern = kPatEvenOdd;
  if (p.pattern == kPatNone || p.nlevels < 1)
    return;
  bool asc = true, desc = true;
  for (int i = 1; i < p.nlevels; ++i) {
    asc = asc && p.shift[i] == p.shift[i - 1] + 1;
    desc = desc && p.shift[i] == p.shift[i - 1] - 1;
  }
  if (asc) {
    p.schedule = kScheduleLadderAsc;
    p.lo = p.shift[0];
    p.hi = p.shift[p.nlevels - 1];
  } else if (desc) {
    p.schedule = kScheduleLadderDesc;
    p.lo = p.shift[p.nlevels - 1];
    p.hi = p.shift[0];
  }
}

namespace fused {

int resident_workgroups(const void *kernel, size_t lds_bytes, int threads) {
  static std::mutex mutex;
  static std::map<std::pair<int, std::pair<const void *, size_t>>, int> cache;
  int device = 0;
  if (hipGetDevice(&device) != hipSuccess)
    return 1024;
  auto const key = std::make_pair(device, std::make_pair(kernel, lds_bytes));
  // a kernel is always launched with the same block size
  {
    std::lock_guard<std::mutex> lock(mutex);
    auto it = cache.find(key);
    if (it != cache.end())
      return it->second;
  }
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads,
                                                   lds_bytes) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                            device) != hipSuccess ||
      cus < 1)
    cus = 256;
  int const resident = per_cu * cus;
  std::lock_guard<std::mutex> lock(mutex);
  cache[key] = resident;
  return resident;
}

} // namespace fused

int fused_max_period(int dtype) {
  return dtype == PALIWA_FLOAT64 ? 2048 : 4096; // the instantiated periods
}

int lds_line_max_period(int dtype) {
  return dtype == PALIWA_FLOAT64 ? 16384 : 32768;
}

hipError_t launch_fused(int dtype, SweepParams const &p, hipStream_t stream) {
  return dtype == PALIWA_FLOAT64 ? launch_fused_f64(p, stream)
                                 : launch_fused_f32(p, stream);
}

hipError_t launch_fused_batched(int dtype, SweepParams const &p, bool strided,
                                ManyGrids grids, int64_t tiles,
                                hipStream_t stream) {
  return dtype == PALIWA_FLOAT64
             ? launch_fused_batched_f64(p, strided, grids, tiles, stream)
             : launch_fused_batched_f32(p, strided, grids, tiles, stream);
}

hipError_t launch_lds_line(int dtype, SweepParams const &p,
                           hipStream_t stream) {
  return dtype == PALIWA_FLOAT64 ? launch_lds_line_typed<double>(p, stream)
                                 : launch_lds_line_typed<float>(p, stream);
}

hipError_t launch_per_pass(int dtype, SweepParams const &p,
                           hipStream_t stream) {
  return dtype == PALIWA_FLOAT64 ? launch_per_pass_typed<double>(p, stream)
                                 : launch_per_pass_typed<float>(p, stream);
}

hipError_t launch_lds_line_batched(int dtype, SweepParams const &p,
                                   ManyGrids grids, int64_t lines,
                                   hipStream_t stream) {
  return dtype == PALIWA_FLOAT64
             ? launch_lds_line_batched_typed<double>(p, grids, lines, stream)
             : launch_lds_line_batched_typed<float>(p, grids, lines, stream);
}

} // namespace paliwa_hip
