// Internal interface between the C ABI (capi.cpp) and the HIP kernels
// (lifting_kernels.hip).  Not installed; the public interface is
// include/paliwa_hip.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/paliwa_hip.h"

namespace paliwa_hip {

// One axis sweep: every line along the axis, every level of the level list in
// order, every pass of the filter list in order.  The array is viewed as
// [A][m][B]: A outer blocks, m stored points along the axis, B inner points.
struct SweepParams {
  void *data;
  int64_t A;
  int64_t B;
  int32_t m;       // stored extent along the axis: n or n+1 (periodic image)
  int32_t L;       // log2 of the lattice period n
  int32_t nlevels; // entries of shift[]
  int32_t npass;   // entries of offset[] / coef[]
  int8_t shift[PALIWA_MAX_LEVELS];  // log2 storage stride of each level
  int8_t offset[PALIWA_MAX_FILTERS]; // 0 even, 1 odd
  double coef[PALIWA_MAX_FILTERS][3];
  // Set by classify(): when the strides are 2^lo, 2^(lo+1), ..., 2^hi in
  // increasing or decreasing order and the pass parities form one of the
  // reference's table patterns, the fused kernel runs a compile-time ladder.
  int32_t schedule; // kSchedule*
  int32_t pattern;  // kPat* (ladder schedules only)
  int32_t lo, hi;   // log2 stride range (ladder schedules only)
  // int32 sweeps only: set to 1 when a result left the int32 range (or null)
  int32_t *overflow;
  // batched fused launches: workgroup size of the group's tiles chosen by the
  // plan from its line count (ladder_threads_for), 0 for the compiled default
  int32_t threads;
};

enum Schedule : int32_t {
  kScheduleGeneric = 0,
  kScheduleLadderAsc = 1,  // strides increase: hierarchize_in's level order
  kScheduleLadderDesc = 2, // strides decrease: dehierarchize_in's level order
};

// Parity sequences of the passes in the reference's filter tables
// (paliwa/paliwa_wavelets.hpp:15-28).
enum Pattern : int32_t {
  kPatOdd = 0,     // {odd}: hat, both directions
  kPatOddEven = 1, // {odd, even}: biorthogonal forward, fullweighting inverse
  kPatEvenOdd = 2, // {even, odd}: fullweighting forward, biorthogonal inverse
  kPatNone = 3,
};

// Where a tile lives: its grid viewed as [A][m][B] and the tile's index
// within that grid.
struct TileRef {
  void *data;
  int64_t A, B;
  int64_t local;
};

// Tiles of one grid (the single-grid entry points).
struct OneGrid {
  void *data;
  int64_t A, B;
  __device__ TileRef operator()(int64_t t) const { return {data, A, B, t}; }
};

// Tiles of many grids (the batched entry points): tile t belongs to grid
// tile_grid[t], whose descriptor gives its storage, its [A][m][B] view and
// the index of its first tile.
struct BatchGrid {
  void *data;
  int64_t A, B;
  int64_t tile_base;
};

struct ManyGrids {
  BatchGrid const *grids;
  int32_t const *tile_grid;
  __device__ TileRef operator()(int64_t t) const {
    BatchGrid const d = grids[tile_grid[t]];
    return {d.data, d.A, d.B, t - d.tile_base};
  }
};

// Fills schedule/pattern/lo/hi from shift[], offset[].
void classify(SweepParams &p);

inline bool has_image(SweepParams const &p) {
  return p.m == (int32_t(1) << p.L) + 1;
}

// Largest lattice period the fused (register) kernel handles per dtype.
int fused_max_period(int dtype);
// Largest lattice period the LDS-line kernel handles per dtype.
int lds_line_max_period(int dtype);

// Launchers; return a hipError_t.  The caller has validated p.
hipError_t launch_fused(int dtype, SweepParams const &p, hipStream_t stream);
hipError_t launch_fused_f64(SweepParams const &p, hipStream_t stream);
hipError_t launch_fused_f32(SweepParams const &p, hipStream_t stream);
// int32 ladders (the learned transform's interpolet sweeps): coefficient slots
// carry {sign, bias, shift} of x + sign * floor((lo + hi + bias) / 2^shift);
// line periods 2^1 .. 2^11, ladder schedules only.
hipError_t launch_fused_i32(SweepParams const &p, hipStream_t stream);

// Points per lane (chunk length K) of the fused kernel for a line of period
// 2^L; the kernel's tile holds 256*K / 2^L lines.
constexpr int kChunkBytes = 128; // register chunk per lane
constexpr int fused_chunk_of(int elem_bytes, int L) {
  int const target = kChunkBytes / elem_bytes;
  int const n = 1 << L;
  return n <= target ? n : (n / target <= 64 ? target : n / 64);
}
inline int fused_chunk(int dtype, int L) {
  return fused_chunk_of(dtype == PALIWA_FLOAT64 ? 8 : 4, L);
}
// Workgroup size of the fused kernel; wider workgroups along strided axes
// give each tile row a longer run of consecutive points.  A strided float
// sweep whose 512-thread tile rows would be shorter than 256 bytes takes 1024
// threads when the doubled tile still fits in LDS (one workgroup per CU):
// float64 lines of 512 points, float32 lines of 512 and 1024.  With
// 512 threads the strided sweeps of a 513^3 grid took 572 instead of 440 us
// (float64) and 249 instead of 230 us (float32), those of 1025-point float32
// lines 307 instead of 274 us.  Longer rows do not pay where they already
// are 256 bytes (257^3 float64: 53 us against 50.5 us), and the int32 ladder
// keeps 512 (it would spill at 1024).  Lines of 64 chunks (G = 64) keep 512
// here: a wide tile halves the tile count, and the 1025^2 float64 grid's
// strided sweep (65 wide tiles for 256 CUs) took 11.5 instead of 8.2 us.
// This is the compiled default; strided float ladders choose their width at
// launch from the line count (ladder_threads_for below: the single-grid launch
// from the grid's lines, a batched plan per group).
constexpr int kStridedThreads = 512;      // strided tiles of 4 or more lanes per line
constexpr int kContigThreads = 256;       // contiguous tiles
constexpr long kWideTileLds = 136 * 1024; // LDS a 1024-thread tile may use
constexpr int fused_threads(bool strided, int K, int G, int elem_bytes,
                            bool integral = false) {
  if (!strided)
    return kContigThreads;
  int const base = G >= 4 ? kStridedThreads : 256;
  if (integral || base != 512 || G > 32)
    return base;
  long const rows = 512 / G;                                   // tile width at 512
  long const wide = 2 * (long(K) * G + 1) * rows * elem_bytes; // LDS at 1024
  return rows * elem_bytes < 256 && wide <= kWideTileLds ? 1024 : base;
}
// Workgroup size of a strided float ladder over `lines` lines whose 512-thread
// tile rows are shorter than 256 bytes (4 <= G <= 64); 0 for the others (the
// compiled default).  1024 threads (rows twice as long) over at least 8192
// lines, two rounds of wide tiles on 256 CUs, where that tile fits in LDS; 256
// where even 512-thread tiles would be fewer than the 256 CUs (a small,
// cache-resident grid); 512 otherwise.
constexpr int ladder_threads_for(int K, int G, int elem_bytes, long lines) {
  if (G < 4 || G > 64 || (512 / G) * elem_bytes >= 256)
    return 0;
  bool const wide = 2 * (long(K) * G + 1) * (512 / G) * elem_bytes <= kWideTileLds;
  if (wide && lines >= 8192)
    return 1024;
  return lines * G < 256L * 512 ? 256 : 512;
}
// Lines per tile of the fused kernel.
inline int fused_tile_lines(int dtype, int L, bool strided) {
  int const K = fused_chunk(dtype, L);
  int const G = (1 << L) / K;
  return fused_threads(strided, K, G, dtype == PALIWA_FLOAT64 ? 8 : 4) / G;
}

// Batched fused sweep: `tiles` tiles of the grids in `grids`, all with line
// period 2^p.L and stored extent p.m along the swept axis; p.data/A/B unused.
hipError_t launch_fused_batched(int dtype, SweepParams const &p, bool strided,
                                ManyGrids grids, int64_t tiles,
                                hipStream_t stream);
hipError_t launch_fused_batched_f64(SweepParams const &p, bool strided,
                                    ManyGrids grids, int64_t tiles,
                                    hipStream_t stream);
hipError_t launch_fused_batched_f32(SweepParams const &p, bool strided,
                                    ManyGrids grids, int64_t tiles,
                                    hipStream_t stream);
hipError_t launch_lds_line(int dtype, SweepParams const &p,
                           hipStream_t stream);
// Batched LDS-line sweep: `lines` lines of the grids in `grids` (one "tile"
// per line), all with line period 2^p.L and stored extent p.m along the axis.
hipError_t launch_lds_line_batched(int dtype, SweepParams const &p,
                                   ManyGrids grids, int64_t lines,
                                   hipStream_t stream);
hipError_t launch_per_pass(int dtype, SweepParams const &p,
                           hipStream_t stream);

// Long lines (long_ladder.hip): periods above fused_max_period up to
// long_ladder_max_period, ladder schedules only; tiles of 256 or 1024 threads
// (long_threads_for: 256 where 1024-thread tiles would be fewer than the CUs
// and a line fits 256 threads), long_tile_lines(dtype, L, threads) lines each.
// Also the column tiles: strided float64 lines of 2^8 points in chunks of 16
// (64 columns per 1024-thread tile, two tiles per CU).
bool long_ladder_fits(int dtype, SweepParams const &p);
// the automatic schedule's choice: strided axes, and contiguous lines longer
// than the LDS-line kernel holds
bool long_ladder_preferred(int dtype, SweepParams const &p);
// the automatic schedule's choice of the column tiles over the fused kernel:
// at least 16384 lines (1024-thread tiles); for the two-pass tables, either
// at least 4096 lines between consecutive points (an outer axis) or a grid of
// at most 160 MiB (Infinity-Cache-resident when swept repeatedly); for the
// hat table, an outer axis of a grid above 256 MiB.  Measured against the
// fused kernel, biorthogonal: 257^3 axes 0 and 1 -4 to -6%, 257 x 1025^2
// axis 0 -4%, 257^2 x 513 axis 1 (271 MiB) +6%; hat: 257 x 1025^2 axis 0
// -9%, 257^3 axis 0 +1.5%, axis 1 -4%, 257^2 x 129 +5 to +6%
bool long_columns_preferred(int dtype, SweepParams const &p);
// contiguous lines of period 2^L that go through the long-line kernel's LDS
// staging buffer (a few lines per round)
bool long_contiguous_staged(int dtype, int L);
int long_ladder_max_period(int dtype);
int long_threads_for(int dtype, int L, int64_t lines);
int64_t long_tile_lines(int dtype, int L, int threads);
hipError_t launch_long_ladder(int dtype, SweepParams const &p, hipStream_t stream);
// Batched: `tiles` tiles of p.threads threads (256 or 1024) of the grids in
// `grids`.
hipError_t launch_long_ladder_batched(int dtype, SweepParams const &p, bool strided,
                                      ManyGrids grids, int64_t tiles, hipStream_t stream);

// Boxes (box.hip): consecutive axes a..b of a grid swept together in LDS.  A
// box is the sub-array of one index of the axes before a (the outer index),
// the full extents of axes a..b, and one tile of `width` consecutive indices
// of the flattened axes after b (the inner index); every axis of the box runs
// all its levels there, in axis order, so the box costs one HBM read and one
// write for its axes together.
constexpr int kBoxMaxAxes = 8;
struct BoxGrid {
  void *data;
  int64_t outer_stride; // points per outer index: product of the extents a..D-1
  int64_t inner;        // points per index of axis b: product of the extents after b
  int64_t outer;        // outer indices: product of the extents before a
  int64_t block_base;   // index of the grid's first workgroup in the launch
  int32_t tiles;        // tiles of the inner range per outer index
  int32_t width;        // inner points per tile: all of them, or a power of two
  int32_t rows;         // product of the box axes' extents
  int32_t batch;        // consecutive outer indices per workgroup (whole-inner tiles)
  int32_t naxes;
  int32_t ext[kBoxMaxAxes]; // stored extents of the box axes
  int8_t L[kBoxMaxAxes];    // log2 line periods
  int8_t nlev[kBoxMaxAxes]; // levels swept per axis (level - minimum level)
};
// One launch: `blocks` boxes of the grids in `grids`, box -> grid through
// `block_grid`; every axis runs the named-transform levels of its direction
// with the passes of `filters`.
struct BoxLaunch {
  BoxGrid const *grids;
  int32_t const *block_grid;
  int64_t blocks;
  int32_t forward;
  int32_t npass;     // 1 or 2
  int32_t offset[2]; // 32-bit: read with scalar loads
  double coef[2][3];
};
// LDS of a box workgroup (bytes): 40, 80 or 160 KiB, four, two or one per CU;
// the box's axis table takes kBoxTableBytes of it, its points the rest.
constexpr int kBoxCapacities[3] = {40 * 1024, 80 * 1024, 160 * 1024};
constexpr int kBoxTableBytes = 128;
constexpr int64_t box_points(int capacity_bytes, int elem_bytes) {
  return (capacity_bytes - kBoxTableBytes) / elem_bytes;
}
hipError_t launch_box_batched(int dtype, BoxLaunch const &launch, int capacity_bytes,
                              int threads, hipStream_t stream);

} // namespace paliwa_hip
