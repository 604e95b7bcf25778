// Batched transform of many grids (a combination-technique set): the plan
// groups the axis sweeps of all grids by (axis, line period, stored extent),
// so that one persistent fused launch covers the tiles of every grid in a
// group, and replays the whole launch sequence of a direction as one HIP
// graph.  The groups of one axis touch disjoint grids, so they run
// concurrently on a few internal streams (forked from and joined back into the
// caller's stream around each axis): the many small groups of a set overlap
// instead of each leaving most of the GPU idle for its launch.  Replaces the reference's per-grid loop
//   for each grid: paliwa::hierarchize(grid, level, minimum_level,
//                                      maximum_level, wavelet, instance)
// (test/test_paliwa_combination_technique.cpp:164-174).
#include "common.hpp"
#include "lifting.hpp"

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

using namespace paliwa_hip;

// Axes a..z of one grid as a box (box.hip): `outer` indices of the axes
// before a, the full extents of a..z (`rows` points), tiles of `width` of the
// `inner` points after z.  A box whose tile holds all inner points is one
// contiguous run of the grid; `batch` consecutive outer indices of such boxes
// share a workgroup.
struct BoxShape {
  int32_t a = 0, z = 0;
  int64_t outer = 1, rows = 1, inner = 1;
  int64_t width = 1, tiles = 1, batch = 1;
};

struct paliwa_batch {
  int device = 0;
  int32_t dtype = 0, ndim = 0, ngrids = 0;
  std::vector<int32_t> level;   // ngrids * ndim
  std::vector<int64_t> extents; // ngrids * ndim
  std::vector<int32_t> lmin, lmax;
  std::vector<void *> data;

  struct Group {
    int32_t phase = 0; // groups of one phase touch disjoint grids
    int32_t dim = 0, L = 0, m = 0;
    bool box = false;  // the boxes of this phase: one box launch
    std::vector<std::pair<int32_t, BoxShape>> boxes; // (grid, box)
    int64_t box_blocks = 0;
    BoxGrid *box_grids_dev = nullptr;
    int32_t *box_block_grid_dev = nullptr;
    bool strided = false;
    bool fused = false;
    bool long_ladder = false;        // longer lines: one batched long-line launch
    bool lds_lines = false;          // lines beyond that: one batched LDS-line launch
    int64_t tiles = 0;               // fused groups: tiles; LDS-line groups: lines
    int32_t threads = 0;             // fused strided groups: tile width (0: default)
    std::vector<int32_t> widths;     // tile widths the tuner may choose from
    int64_t lines = 0;               // lines of all members along the axis
    BatchGrid *grids_dev = nullptr;  // fused groups: descriptors on the device
    int32_t *tile_grid_dev = nullptr;
    std::vector<int32_t> members;    // grid indices
  };
  std::vector<Group> groups; // in phase order
  void *dev = nullptr;       // one allocation holding every descriptor table
  size_t dev_bytes = 0;

  // internal streams (default priority) and fork / join events
  int32_t nlanes = 4;      // streams for the groups of one axis
  int32_t long_weight = 4; // launch-cost multiple of a long-line group
  std::vector<hipStream_t> lanes;
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr;
  std::vector<int32_t> lane_of; // per group: its internal stream

  bool use_graphs = true;
  std::mutex mutex;
  std::map<std::pair<int, int>, hipGraphExec_t> graphs; // (wavelet, forward)

  // the last paliwa_batch_tune call
  double rule_ms = -1.0, tuned_ms = -1.0;
  int32_t evaluations = 0;

  // Box schedule (box.hip) when box_capacity > 0: each grid's axes split
  // into segments swept in turn, boxes of consecutive axes that fit the LDS
  // capacity or single axes through the groups' kernels; phase p holds the
  // p-th segment of every grid.  Without boxes phase p is axis p.
  int32_t box_capacity = 0; // bytes of LDS per box (kBoxCapacities), 0: no boxes
  int32_t box_width = 0;    // inner points per tile of a tiled box (power of two)
  int32_t box_threads = 256;
  void *box_dev = nullptr;  // the box groups' descriptor tables
  size_t box_dev_bytes = 0;

  int32_t const *level_of(int g) const { return &level[size_t(g) * ndim]; }
  int64_t const *extents_of(int g) const { return &extents[size_t(g) * ndim]; }
};

namespace {

int launch_group(paliwa_batch &b, paliwa_batch::Group const &grp,
                 paliwa_filter const *filters, int nfilters, bool forward,
                 hipStream_t stream) {
  if (grp.box) {
    if (nfilters < 1 || nfilters > 2)
      return fail(PALIWA_ERR_UNSUPPORTED, "boxes run filter tables of one or two passes");
    BoxLaunch launch{};
    launch.grids = grp.box_grids_dev;
    launch.block_grid = grp.box_block_grid_dev;
    launch.blocks = grp.box_blocks;
    launch.forward = forward;
    launch.npass = nfilters;
    for (int f = 0; f < nfilters; ++f) {
      launch.offset[f] = filters[f].offset;
      for (int c = 0; c < 3; ++c)
        launch.coef[f][c] = filters[f].coefficients[c];
    }
    hipError_t const err =
        launch_box_batched(b.dtype, launch, b.box_capacity, b.box_threads, stream);
    if (err != hipSuccess)
      return fail(PALIWA_ERR_HIP, std::string("HIP: ") + hipGetErrorString(err));
    return PALIWA_OK;
  }
  if (grp.fused) {
    SweepParams p{};
    p.m = grp.m;
    p.L = grp.L;
    p.threads = grp.threads;
    set_filters(p, filters, nfilters);
    set_named_levels(p, grp.L, b.lmin[grp.dim], forward);
    classify(p);
    hipError_t const err = launch_fused_batched(
        b.dtype, p, grp.strided, ManyGrids{grp.grids_dev, grp.tile_grid_dev},
        grp.tiles, stream);
    if (err != hipSuccess)
      return fail(PALIWA_ERR_HIP, std::string("HIP: ") + hipGetErrorString(err));
    return PALIWA_OK;
  }
  if (grp.long_ladder) {
    SweepParams p{};
    p.m = grp.m;
    p.L = grp.L;
    set_filters(p, filters, nfilters);
    set_named_levels(p, grp.L, b.lmin[grp.dim], forward);
    classify(p);
    p.threads = grp.threads;
    hipError_t const err = launch_long_ladder_batched(
        b.dtype, p, grp.strided, ManyGrids{grp.grids_dev, grp.tile_grid_dev}, grp.tiles, stream);
    if (err != hipSuccess)
      return fail(PALIWA_ERR_HIP, std::string("HIP: ") + hipGetErrorString(err));
    return PALIWA_OK;
  }
  if (grp.lds_lines) {
    SweepParams p{};
    p.m = grp.m;
    p.L = grp.L;
    set_filters(p, filters, nfilters);
    set_named_levels(p, grp.L, b.lmin[grp.dim], forward);
    classify(p);
    hipError_t const err = launch_lds_line_batched(
        b.dtype, p, ManyGrids{grp.grids_dev, grp.tile_grid_dev}, grp.tiles, stream);
    if (err != hipSuccess)
      return fail(PALIWA_ERR_HIP, std::string("HIP: ") + hipGetErrorString(err));
    return PALIWA_OK;
  }
  for (int32_t g : grp.members) {
    SweepParams p = make_params(b.data[g], b.ndim, b.extents_of(g),
                                b.level_of(g), grp.dim);
    set_filters(p, filters, nfilters);
    set_named_levels(p, grp.L, b.lmin[grp.dim], forward);
    int const rc = sweep(b.dtype, p, PALIWA_PATH_AUTO, stream);
    if (rc != PALIWA_OK)
      return rc;
  }
  return PALIWA_OK;
}

int launch_sequence(paliwa_batch &b, WaveletEntry const &w, bool forward,
                    hipStream_t stream) {
  paliwa_filter const *filters = forward ? w.forward : w.reconstruct;
  int const nfilters = forward ? w.nforward : w.nreconstruct;
  auto hip = [](hipError_t e) {
    return e == hipSuccess ? PALIWA_OK
                           : fail(PALIWA_ERR_HIP, std::string("HIP: ") + hipGetErrorString(e));
  };
  size_t i = 0;
  while (i < b.groups.size()) {
    // the groups of one phase
    size_t j = i;
    while (j < b.groups.size() && b.groups[j].phase == b.groups[i].phase)
      ++j;
    if (j - i == 1 || b.lanes.empty()) {
      for (size_t k = i; k < j; ++k) {
        int const rc = launch_group(b, b.groups[k], filters, nfilters, forward, stream);
        if (rc != PALIWA_OK)
          return rc;
      }
    } else {
      int rc = hip(hipEventRecord(b.fork, stream));
      std::vector<bool> used(b.lanes.size(), false);
      for (size_t k = i; k < j && rc == PALIWA_OK; ++k)
        used[b.lane_of[k]] = true;
      for (size_t t = 0; t < b.lanes.size() && rc == PALIWA_OK; ++t)
        if (used[t])
          rc = hip(hipStreamWaitEvent(b.lanes[t], b.fork, 0));
      // the box group first (the phase's largest launch, which should not
      // queue behind others), then long-line groups (few workgroups each,
      // latency-bound, so they run beside the large groups instead of after
      // them), then the rest
      auto order = [&](paliwa_batch::Group const &g) {
        return g.box ? 0 : (g.lds_lines || g.long_ladder) ? 1 : 2;
      };
      for (int pass = 0; pass < 3 && rc == PALIWA_OK; ++pass)
        for (size_t k = i; k < j && rc == PALIWA_OK; ++k)
          if (order(b.groups[k]) == pass)
            rc = launch_group(b, b.groups[k], filters, nfilters, forward,
                              b.lanes[b.lane_of[k]]);
      for (size_t t = 0; t < b.lanes.size() && rc == PALIWA_OK; ++t)
        if (used[t]) {
          rc = hip(hipEventRecord(b.joins[t], b.lanes[t]));
          if (rc == PALIWA_OK)
            rc = hip(hipStreamWaitEvent(stream, b.joins[t], 0));
        }
      if (rc != PALIWA_OK)
        return rc;
    }
    i = j;
  }
  return PALIWA_OK;
}

// Capture the launch sequence of one direction into a graph (once per plan,
// wavelet and direction).  Returns nullptr if capture is not possible.
hipGraphExec_t capture(paliwa_batch &b, WaveletEntry const &w, bool forward) {
  hipStream_t cs = nullptr;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
    return nullptr;
  hipGraphExec_t exec = nullptr;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) == hipSuccess) {
    int const rc = launch_sequence(b, w, forward, cs);
    hipGraph_t graph = nullptr;
    hipError_t const end = hipStreamEndCapture(cs, &graph);
    if (rc == PALIWA_OK && end == hipSuccess && graph != nullptr) {
      if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
        exec = nullptr;
    }
    if (graph != nullptr)
      (void)hipGraphDestroy(graph);
  }
  (void)hipStreamDestroy(cs);
  (void)hipGetLastError();
  return exec;
}

int execute(paliwa_batch *b, char const *wavelet_name, bool forward,
            void *stream) {
  if (b == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  WaveletEntry const *w = find_wavelet(wavelet_name);
  if (w == nullptr)
    return fail(PALIWA_ERR_UNKNOWN_WAVELET, "map::at");
  int current = 0;
  if (hipGetDevice(&current) != hipSuccess || current != b->device)
    return fail(PALIWA_ERR_INVALID_ARGUMENT,
                "the plan belongs to device " + std::to_string(b->device));
  hipStream_t const s = static_cast<hipStream_t>(stream);
  TraceRange const range(forward ? "paliwa batch hierarchize" : "paliwa batch dehierarchize");
  if (b->use_graphs) {
    hipGraphExec_t exec = nullptr;
    {
      std::lock_guard<std::mutex> lock(b->mutex);
      auto const key = std::make_pair(int(w - find_wavelet("hat")), int(forward));
      auto it = b->graphs.find(key);
      if (it == b->graphs.end()) {
        exec = capture(*b, *w, forward);
        if (exec != nullptr)
          b->graphs[key] = exec;
      } else {
        exec = it->second;
      }
    }
    if (exec != nullptr) {
      hipError_t const err = hipGraphLaunch(exec, s);
      if (err != hipSuccess)
        return fail(PALIWA_ERR_HIP, std::string("HIP: ") + hipGetErrorString(err));
      last_error().clear();
      return PALIWA_OK;
    }
  }
  int const rc = launch_sequence(*b, *w, forward, s);
  if (rc == PALIWA_OK)
    last_error().clear();
  return rc;
}

void drop_graphs(paliwa_batch *b) {
  for (auto &kv : b->graphs)
    (void)hipGraphExecDestroy(kv.second);
  b->graphs.clear();
}

// Lines per tile of a group at its current width.
int64_t tile_lines(paliwa_batch const &b, paliwa_batch::Group const &grp) {
  if (grp.long_ladder)
    return long_tile_lines(b.dtype, grp.L, grp.threads);
  if (!grp.fused)
    return 1;
  int const G = (1 << grp.L) / fused_chunk(b.dtype, grp.L);
  return grp.threads ? grp.threads / G : fused_tile_lines(b.dtype, grp.L, grp.strided);
}

// (Re)builds every group's tile table at its current width: one device
// allocation holding all of them.
int build_tables(paliwa_batch *b) {
  std::vector<std::vector<BatchGrid>> host_grids(b->groups.size());
  std::vector<std::vector<int32_t>> host_tiles(b->groups.size());
  for (size_t i = 0; i < b->groups.size(); ++i) {
    auto &grp = b->groups[i];
    grp.tiles = 0;
    if (!grp.fused && !grp.long_ladder && !grp.lds_lines)
      continue;
    int64_t const tb = tile_lines(*b, grp);
    for (size_t k = 0; k < grp.members.size(); ++k) {
      int32_t const g = grp.members[k];
      SweepParams const p = make_params(b->data[g], b->ndim, b->extents_of(g), b->level_of(g),
                                        grp.dim);
      int64_t const tiles = (p.A * p.B + tb - 1) / tb;
      host_grids[i].push_back(BatchGrid{b->data[g], p.A, p.B, grp.tiles});
      host_tiles[i].insert(host_tiles[i].end(), size_t(tiles), int32_t(k));
      grp.tiles += tiles;
    }
    if (grp.tiles > 0x7fffffffLL)
      return fail(PALIWA_ERR_UNSUPPORTED, "batch too large");
  }
  std::vector<size_t> grid_off(b->groups.size()), tile_off(b->groups.size());
  size_t bytes = 0;
  for (size_t i = 0; i < b->groups.size(); ++i) {
    grid_off[i] = bytes;
    bytes += host_grids[i].size() * sizeof(BatchGrid);
    bytes = (bytes + 255) & ~size_t(255);
    tile_off[i] = bytes;
    bytes += host_tiles[i].size() * sizeof(int32_t);
    bytes = (bytes + 255) & ~size_t(255);
  }
  drop_graphs(b);
  if (b->dev != nullptr && bytes != b->dev_bytes) {
    (void)hipFree(b->dev);
    b->dev = nullptr;
    b->dev_bytes = 0;
  }
  if (bytes == 0)
    return PALIWA_OK;
  std::vector<unsigned char> staging(bytes, 0);
  for (size_t i = 0; i < b->groups.size(); ++i) {
    std::memcpy(staging.data() + grid_off[i], host_grids[i].data(),
                host_grids[i].size() * sizeof(BatchGrid));
    std::memcpy(staging.data() + tile_off[i], host_tiles[i].data(),
                host_tiles[i].size() * sizeof(int32_t));
  }
  if ((b->dev == nullptr && hipMalloc(&b->dev, bytes) != hipSuccess) ||
      hipMemcpy(b->dev, staging.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
    return fail(PALIWA_ERR_HIP, "could not place the batch descriptors");
  b->dev_bytes = bytes;
  for (size_t i = 0; i < b->groups.size(); ++i) {
    auto *base = static_cast<unsigned char *>(b->dev);
    b->groups[i].grids_dev = reinterpret_cast<BatchGrid *>(base + grid_off[i]);
    b->groups[i].tile_grid_dev = reinterpret_cast<int32_t *>(base + tile_off[i]);
  }
  return PALIWA_OK;
}

// ---- box schedule --------------------------------------------------------

// The box of axes a..z of grid g within `cap` points, or false: the tile holds
// all inner points when they fit (several consecutive outer indices per
// workgroup then, up to the capacity), else `width` of them.
bool make_box(paliwa_batch const &b, int32_t g, int32_t a, int32_t z, int64_t cap,
              int64_t width, BoxShape *out) {
  int64_t const *e = b.extents_of(g);
  BoxShape x;
  x.a = a;
  x.z = z;
  for (int32_t d = 0; d < a; ++d)
    x.outer *= e[d];
  for (int32_t d = a; d <= z; ++d)
    x.rows *= e[d];
  for (int32_t d = z + 1; d < b.ndim; ++d)
    x.inner *= e[d];
  // a box's span in its grid stays below 2^28 points: 32-bit offsets
  if (z - a + 1 > kBoxMaxAxes || x.rows > cap || x.rows * x.inner >= (int64_t(1) << 28))
    return false;
  if (x.rows * x.inner <= cap) {
    x.width = x.inner;
    x.batch = std::min(x.outer, cap / (x.rows * x.inner));
  } else if (x.rows * width <= cap && width < x.inner) {
    x.width = width;
  } else {
    return false;
  }
  x.tiles = (x.inner + x.width - 1) / x.width;
  *out = x;
  return true;
}

// One sweep of a grid's schedule: a box of axes a..z, or the single axis a
// through the groups' kernels.
struct Segment {
  bool box = false;
  BoxShape shape;
};

// The segments of grid g in sweep order.  Without boxes: every axis with
// levels to sweep.  With boxes: the fewest segments; among as many, the
// fewest single-axis boxes (a single axis runs in the groups' register
// kernels), then the fewest points in the empty part of partial tiles.
// Segments with no levels to sweep are left out.
std::vector<Segment> segments_of(paliwa_batch const &b, int32_t g) {
  int32_t const D = b.ndim;
  int32_t const *lv = b.level_of(g);
  auto swept = [&](int32_t a, int32_t z) {
    for (int32_t d = a; d <= z; ++d)
      if (lv[d] > b.lmin[d])
        return true;
    return false;
  };
  std::vector<Segment> out;
  if (b.box_capacity == 0) {
    for (int32_t d = 0; d < D; ++d)
      if (swept(d, d)) {
        Segment s;
        s.shape.a = s.shape.z = d;
        out.push_back(s);
      }
    return out;
  }
  int64_t const cap = box_points(b.box_capacity, b.dtype == PALIWA_FLOAT64 ? 8 : 4);
  struct Best {
    bool ok = false;
    int32_t segments = 0, single_boxes = 0;
    int64_t waste = 0;
    int32_t from = 0;
    Segment last;
  };
  std::vector<Best> best(size_t(D) + 1);
  best[0].ok = true;
  for (int32_t z = 0; z < D; ++z)
    for (int32_t a = 0; a <= z; ++a) {
      if (!best[a].ok)
        continue;
      auto offer = [&](Segment const &seg, int64_t waste) {
        Best const &f = best[a];
        Best c{true, f.segments + 1, f.single_boxes + (seg.box && a == z ? 1 : 0),
               f.waste + waste, a, seg};
        Best &t = best[size_t(z) + 1];
        if (!t.ok || std::tie(c.segments, c.single_boxes, c.waste) <
                         std::tie(t.segments, t.single_boxes, t.waste))
          t = c;
      };
      Segment seg;
      seg.box = true;
      if (make_box(b, g, a, z, cap, b.box_width, &seg.shape))
        offer(seg, seg.shape.outer * seg.shape.rows *
                       (seg.shape.tiles * seg.shape.width - seg.shape.inner));
      if (a == z) {
        Segment one;
        one.shape.a = one.shape.z = a;
        offer(one, 0);
      }
    }
  for (int32_t z = D; z > 0; z = best[z].from)
    out.push_back(best[z].last);
  std::reverse(out.begin(), out.end());
  out.erase(std::remove_if(out.begin(), out.end(),
                           [&](Segment const &x) { return !swept(x.shape.a, x.shape.z); }),
            out.end());
  return out;
}

void drop_box_tables(paliwa_batch *b) {
  if (b->box_dev != nullptr)
    (void)hipFree(b->box_dev);
  b->box_dev = nullptr;
  b->box_dev_bytes = 0;
  for (auto &grp : b->groups) {
    grp.box_grids_dev = nullptr;
    grp.box_block_grid_dev = nullptr;
  }
}

// Descriptor tables of the box groups: one device allocation.
int build_box_tables(paliwa_batch *b) {
  drop_box_tables(b);
  std::vector<std::vector<BoxGrid>> grids(b->groups.size());
  std::vector<std::vector<int32_t>> block_grid(b->groups.size());
  for (size_t i = 0; i < b->groups.size(); ++i) {
    auto &grp = b->groups[i];
    if (!grp.box)
      continue;
    grp.box_blocks = 0;
    for (auto const &[g, x] : grp.boxes) {
      BoxGrid d{};
      d.data = b->data[g];
      d.outer_stride = x.rows * x.inner;
      d.inner = x.inner;
      d.outer = x.outer;
      d.block_base = grp.box_blocks;
      d.tiles = int32_t(x.tiles);
      d.width = int32_t(x.width);
      d.rows = int32_t(x.rows);
      d.batch = int32_t(x.batch);
      d.naxes = x.z - x.a + 1;
      for (int32_t k = 0; k < d.naxes; ++k) {
        int32_t const dim = x.a + k;
        d.ext[k] = int32_t(b->extents_of(g)[dim]);
        d.L[k] = int8_t(b->level_of(g)[dim]);
        d.nlev[k] = int8_t(b->level_of(g)[dim] - b->lmin[dim]);
      }
      int64_t const n = (x.outer + x.batch - 1) / x.batch * x.tiles;
      if (grp.box_blocks + n > 0x7fffff00LL)
        return fail(PALIWA_ERR_UNSUPPORTED, "batch too large");
      block_grid[i].insert(block_grid[i].end(), size_t(n), int32_t(grids[i].size()));
      grids[i].push_back(d);
      grp.box_blocks += n;
    }
  }
  std::vector<size_t> grid_off(b->groups.size()), block_off(b->groups.size());
  size_t bytes = 0;
  for (size_t i = 0; i < b->groups.size(); ++i) {
    grid_off[i] = bytes;
    bytes = (bytes + grids[i].size() * sizeof(BoxGrid) + 255) & ~size_t(255);
    block_off[i] = bytes;
    bytes = (bytes + block_grid[i].size() * sizeof(int32_t) + 255) & ~size_t(255);
  }
  if (bytes == 0)
    return PALIWA_OK;
  std::vector<unsigned char> staging(bytes, 0);
  for (size_t i = 0; i < b->groups.size(); ++i) {
    std::memcpy(staging.data() + grid_off[i], grids[i].data(), grids[i].size() * sizeof(BoxGrid));
    std::memcpy(staging.data() + block_off[i], block_grid[i].data(),
                block_grid[i].size() * sizeof(int32_t));
  }
  if (hipMalloc(&b->box_dev, bytes) != hipSuccess ||
      hipMemcpy(b->box_dev, staging.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    drop_box_tables(b);
    return fail(PALIWA_ERR_HIP, "could not place the box descriptors");
  }
  b->box_dev_bytes = bytes;
  auto *base = static_cast<unsigned char *>(b->box_dev);
  for (size_t i = 0; i < b->groups.size(); ++i)
    if (b->groups[i].box) {
      b->groups[i].box_grids_dev = reinterpret_cast<BoxGrid *>(base + grid_off[i]);
      b->groups[i].box_block_grid_dev = reinterpret_cast<int32_t *>(base + block_off[i]);
    }
  return PALIWA_OK;
}

// Points read per direction by the plan's current segments (boxes with the
// empty part of their partial tiles).
int64_t traffic(paliwa_batch const &b) {
  int64_t total = 0;
  for (int32_t g = 0; g < b.ngrids; ++g) {
    int64_t e = 1;
    for (int32_t d = 0; d < b.ndim; ++d)
      e *= b.extents_of(g)[d];
    for (Segment const &seg : segments_of(b, g))
      total += seg.box ? seg.shape.outer * seg.shape.rows * seg.shape.tiles * seg.shape.width : e;
  }
  return total;
}

constexpr int kMaxLanes = 8;

// Whether some phase has more than one group (then its groups run on the
// internal streams).
bool uses_lanes(paliwa_batch const &b) {
  for (size_t i = 1; i < b.groups.size(); ++i)
    if (b.groups[i].phase == b.groups[i - 1].phase)
      return true;
  return false;
}
// the box schedule a new plan takes when it pays: four boxes per CU, tile
// rows of 512 bytes, 256 threads (float64; float32 boxes take 512 threads,
// whose register rows then hold the whole box).  On the 1325-grid 4-D set
// (|l|_1 in 13..16) this was the fastest of the tuner's 19 settings: 0.89 ms
// per forward direction against 1.32 ms without boxes, 1.03 ms for 80 KiB
// boxes and 1.26 ms or more for 160 KiB ones.
constexpr int32_t kBoxRuleCapacity = 40 * 1024;
constexpr int64_t kBoxRuleRowBytes = 512;

// Lanes for the groups of each axis: largest group first onto the least
// loaded of b->nlanes streams.  Work estimate in swept points: the group's
// points plus a fixed cost per launch (about 10 us of launch, fill and drain,
// the time the sweep needs for ~3.75 M float64 points), long_weight times that
// for a long-line group, whose few workgroups make it latency-bound.  Creates
// the streams when the lane count changed.
int setup_lanes(paliwa_batch *b) {
  drop_graphs(b); // captured graphs hold the old assignment
  b->lane_of.assign(b->groups.size(), 0);
  bool concurrent = false;
  for (size_t i = 0; i < b->groups.size();) {
    size_t j = i;
    while (j < b->groups.size() && b->groups[j].phase == b->groups[i].phase)
      ++j;
    if (j - i > 1) {
      concurrent = true;
      std::vector<std::pair<int64_t, size_t>> work;
      for (size_t k = i; k < j; ++k) {
        int64_t pts = 0;
        for (int32_t g : b->groups[k].members) {
          int64_t e = 1;
          for (int32_t d = 0; d < b->ndim; ++d)
            e *= b->extents_of(g)[d];
          pts += e;
        }
        for (auto const &[g, x] : b->groups[k].boxes)
          pts += x.outer * x.rows * x.tiles * x.width;
        int64_t const launch = b->dtype == PALIWA_FLOAT64 ? 3750000 : 7500000;
        bool const long_lines = b->groups[k].lds_lines || b->groups[k].long_ladder;
        pts += long_lines ? b->long_weight * launch : launch;
        work.push_back({pts, k});
      }
      std::sort(work.begin(), work.end(), [](auto const &x, auto const &y) {
        return x.first != y.first ? x.first > y.first : x.second < y.second;
      });
      std::vector<int64_t> load(size_t(b->nlanes), 0);
      for (auto const &[pts, k] : work) {
        int best = 0;
        for (int t = 1; t < b->nlanes; ++t)
          if (load[t] < load[best])
            best = t;
        load[best] += pts;
        b->lane_of[k] = best;
      }
    }
    i = j;
  }
  // the streams are created once, kMaxLanes of them; a schedule uses the
  // first nlanes
  if (!concurrent || !b->lanes.empty())
    return PALIWA_OK;
  b->lanes.assign(kMaxLanes, nullptr);
  b->joins.assign(kMaxLanes, nullptr);
  bool ok = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
  for (int t = 0; t < kMaxLanes && ok; ++t)
    ok = hipStreamCreateWithFlags(&b->lanes[t], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&b->joins[t], hipEventDisableTiming) == hipSuccess;
  if (!ok)
    return fail(PALIWA_ERR_HIP, "could not create the plan's streams");
  return PALIWA_OK;
}

// A group of single-axis sweeps: grids of one phase whose segment is axis d
// with line period 2^L and stored extent m.
paliwa_batch::Group single_group(paliwa_batch const &b, int32_t phase, int32_t d, int32_t L,
                                 int32_t m, std::vector<int32_t> members) {
  int32_t const dtype = b.dtype;
  paliwa_batch::Group grp;
  grp.phase = phase;
  grp.dim = d;
  grp.L = L;
  grp.m = m;
  grp.strided = d + 1 < b.ndim;
  grp.members = std::move(members);
  grp.fused = (int64_t(1) << grp.L) <= fused_max_period(dtype);
  for (int32_t g : grp.members) {
    SweepParams const q = make_params(b.data[g], b.ndim, b.extents_of(g), b.level_of(g), d);
    grp.lines += q.A * q.B;
  }
  // long lines: the long-line kernel on strided axes and on contiguous lines
  // its staging buffer holds, the LDS-line kernel on longer contiguous lines
  // as far as it holds them (as the single-grid schedule chooses), the
  // long-line kernel beyond
  int64_t const n = int64_t(1) << grp.L;
  grp.long_ladder = !grp.fused && n <= long_ladder_max_period(dtype) &&
                    (grp.strided || long_contiguous_staged(dtype, grp.L) ||
                     n > lds_line_max_period(dtype));
  grp.lds_lines = !grp.fused && !grp.long_ladder && n <= lds_line_max_period(dtype);
  if (grp.long_ladder) {
    grp.threads = long_threads_for(dtype, grp.L, grp.lines);
    int const G = int(n / (dtype == PALIWA_FLOAT64 ? 32 : 64));
    if (G <= 256)
      grp.widths = {256, 1024};
  }
  if (grp.fused && grp.strided) {
    // the tile width of a strided float ladder follows the group's lines
    int const K = fused_chunk(dtype, grp.L);
    int const G = (1 << grp.L) / K;
    int const eb = dtype == PALIWA_FLOAT64 ? 8 : 4;
    grp.threads = ladder_threads_for(K, G, eb, grp.lines);
    if (grp.threads != 0) {
      grp.widths = {256, 512};
      if (2 * (int64_t(K) * G + 1) * (512 / G) * eb <= kWideTileLds)
        grp.widths.push_back(1024);
    }
  }
  return grp;
}

// The plan's schedule from its box settings: the segments of every grid, the
// groups of each phase (one box group, single-axis groups by axis, line
// period and stored extent), their tables and streams.
int build_structure(paliwa_batch *b) {
  drop_graphs(b);
  drop_box_tables(b);
  b->groups.clear();
  std::map<std::tuple<int32_t, int32_t, int32_t, int32_t>, std::vector<int32_t>> singles;
  std::map<int32_t, std::vector<std::pair<int32_t, BoxShape>>> boxes;
  for (int32_t g = 0; g < b->ngrids; ++g) {
    std::vector<Segment> const segs = segments_of(*b, g);
    for (size_t p = 0; p < segs.size(); ++p) {
      Segment const &seg = segs[p];
      int32_t const phase = b->box_capacity == 0 ? seg.shape.a : int32_t(p);
      if (seg.box) {
        boxes[phase].push_back({g, seg.shape});
      } else {
        int32_t const d = seg.shape.a;
        singles[{phase, d, b->level_of(g)[d], int32_t(b->extents_of(g)[d])}].push_back(g);
      }
    }
  }
  for (int32_t phase = 0; phase < b->ndim; ++phase) {
    auto const it = boxes.find(phase);
    if (it != boxes.end()) {
      paliwa_batch::Group grp;
      grp.phase = phase;
      grp.box = true;
      grp.boxes = it->second;
      b->groups.push_back(std::move(grp));
    }
    for (auto &[key, members] : singles)
      if (std::get<0>(key) == phase)
        b->groups.push_back(single_group(*b, phase, std::get<1>(key), std::get<2>(key),
                                         std::get<3>(key), members));
  }
  int rc = build_tables(b);
  if (rc == PALIWA_OK)
    rc = build_box_tables(b);
  if (rc == PALIWA_OK)
    rc = setup_lanes(b);
  return rc;
}

void release(paliwa_batch *b) {
  for (auto &kv : b->graphs)
    (void)hipGraphExecDestroy(kv.second);
  for (hipStream_t s : b->lanes)
    if (s != nullptr)
      (void)hipStreamDestroy(s);
  for (hipEvent_t e : b->joins)
    if (e != nullptr)
      (void)hipEventDestroy(e);
  if (b->fork != nullptr)
    (void)hipEventDestroy(b->fork);
  if (b->dev != nullptr)
    (void)hipFree(b->dev);
  drop_box_tables(b);
  delete b;
}


// ---- autotuning -------------------------------------------------------------
// The rules above (tile width from a group's line count, four streams, a
// launch-cost estimate for the lane balance) are a starting point; the tuner
// times alternatives on the plan's own grids and keeps the fastest.  What it
// chooses: the tile width of every strided fused group whose width is chosen
// at launch (256 / 512 / 1024 threads where the tile fits), and the stream
// schedule (number of streams, launch-cost weight of long-line groups).  The
// objective is the whole direction's time as the plan executes it (one graph
// replay), so interactions between concurrent groups count.

uint64_t fnv(uint64_t h, int64_t v) {
  for (int i = 0; i < 8; ++i) {
    h ^= uint64_t(v >> (8 * i)) & 0xff;
    h *= 1099511628211ull;
  }
  return h;
}

std::string direction_tag(char const *wavelet, bool forward) {
  return std::string(wavelet) + (forward ? " forward" : " inverse");
}

std::string width_key(paliwa_batch const &b, paliwa_batch::Group const &grp, char const *wavelet,
                      bool forward) {
  int bucket = 0;
  while ((int64_t(2) << bucket) <= grp.lines)
    ++bucket;
  std::ostringstream k;
  k << "width " << (b.dtype == PALIWA_FLOAT64 ? "f64" : "f32") << " "
    << (grp.strided ? "strided" : "contiguous") << " L=" << grp.L << " m=" << grp.m
    << " lines=2^" << bucket << " " << direction_tag(wavelet, forward);
  return k.str();
}

std::string schedule_key(paliwa_batch const &b, char const *wavelet, bool forward) {
  uint64_t h = 1469598103934665603ull;
  h = fnv(fnv(h, b.dtype), b.ndim);
  for (auto const &grp : b.groups) {
    h = fnv(fnv(fnv(h, grp.dim), grp.L), grp.m);
    h = fnv(fnv(h, grp.phase), grp.box ? int64_t(grp.boxes.size()) : -1);
    h = fnv(fnv(h, grp.lines), int64_t(grp.members.size()));
  }
  std::ostringstream k;
  k << "schedule " << std::hex << h << std::dec << " " << direction_tag(wavelet, forward);
  return k.str();
}

using TuneCache = std::map<std::string, std::string>;

TuneCache read_cache(char const *path) {
  TuneCache c;
  if (path == nullptr)
    return c;
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#')
      continue;
    auto const eq = line.find(" = ");
    if (eq != std::string::npos)
      c[line.substr(0, eq)] = line.substr(eq + 3);
  }
  return c;
}

int write_cache(char const *path, TuneCache const &c) {
  std::ofstream out(path, std::ios::trunc);
  out << "# paliwa batched-transform tuning cache (paliwa_batch_tune), one choice per line:\n"
         "#   width <dtype> <strided|contiguous> L=<log2 period> m=<stored extent> "
         "lines=2^<floor log2 of the group's lines> <wavelet> <direction> = <threads>\n"
         "#   schedule <plan hash> <wavelet> <direction> = <streams> <long-line weight>\n"
         "#   boxes <grid-set hash> <wavelet> <direction> = <box bytes> <tile width> <threads>\n";
  for (auto const &[k, v] : c)
    out << k << " = " << v << "\n";
  return out.good() ? PALIWA_OK : fail(PALIWA_ERR_IO, std::string("cannot write ") + path);
}

// Median of `reps` timed replays of one direction (ms); each is followed by
// an untimed replay of the other direction, so the values stay bounded.
int time_direction(paliwa_batch *b, char const *wavelet, bool forward, hipStream_t s, int reps,
                   double *ms) {
  std::vector<hipEvent_t> ev(2 * size_t(reps), nullptr);
  int rc = PALIWA_OK;
  for (auto &e : ev)
    if (hipEventCreate(&e) != hipSuccess)
      rc = fail(PALIWA_ERR_HIP, "could not create timing events");
  for (int r = 0; r < 2 && rc == PALIWA_OK; ++r) // warm-up (and graph capture)
    if ((rc = execute(b, wavelet, forward, s)) == PALIWA_OK)
      rc = execute(b, wavelet, !forward, s);
  for (int r = 0; r < reps && rc == PALIWA_OK; ++r) {
    (void)hipEventRecord(ev[2 * r], s);
    rc = execute(b, wavelet, forward, s);
    (void)hipEventRecord(ev[2 * r + 1], s);
    if (rc == PALIWA_OK)
      rc = execute(b, wavelet, !forward, s);
  }
  std::vector<double> t;
  if (rc == PALIWA_OK && hipStreamSynchronize(s) != hipSuccess)
    rc = fail(PALIWA_ERR_HIP, "timing run failed");
  for (int r = 0; r < reps && rc == PALIWA_OK; ++r) {
    float x = 0.0f;
    if (hipEventElapsedTime(&x, ev[2 * r], ev[2 * r + 1]) == hipSuccess)
      t.push_back(x);
  }
  for (auto e : ev)
    if (e != nullptr)
      (void)hipEventDestroy(e);
  if (rc != PALIWA_OK)
    return rc;
  if (t.empty())
    return fail(PALIWA_ERR_HIP, "no timing");
  std::sort(t.begin(), t.end());
  *ms = t[t.size() / 2];
  return PALIWA_OK;
}

std::string boxes_key(paliwa_batch const &b, char const *wavelet, bool forward) {
  uint64_t h = 1469598103934665603ull;
  h = fnv(fnv(fnv(h, b.dtype), b.ndim), b.ngrids);
  for (int32_t d = 0; d < b.ndim; ++d)
    h = fnv(h, b.lmin[d]);
  for (int32_t g = 0; g < b.ngrids; ++g)
    for (int32_t d = 0; d < b.ndim; ++d)
      h = fnv(fnv(h, b.level_of(g)[d]), b.extents_of(g)[d]);
  std::ostringstream k;
  k << "boxes " << std::hex << h << std::dec << " " << direction_tag(wavelet, forward);
  return k.str();
}

bool valid_boxes(int32_t cap, int32_t width, int32_t threads) {
  if (cap == 0)
    return true;
  bool const known = cap == kBoxCapacities[0] || cap == kBoxCapacities[1] || cap == kBoxCapacities[2];
  return known && width >= 1 && width <= 4096 && (width & (width - 1)) == 0 &&
         (threads == 256 || threads == 512);
}

int tune(paliwa_batch *b, char const *wavelet_name, bool forward, char const *cache_path,
         int32_t flags, hipStream_t s) {
  WaveletEntry const *w = find_wavelet(wavelet_name);
  if (w == nullptr)
    return fail(PALIWA_ERR_UNKNOWN_WAVELET, "map::at");
  if (b->dtype != PALIWA_FLOAT64)
    return fail(PALIWA_ERR_UNSUPPORTED,
                "tuning of float32 plans is switched off (a float32 tuning run ended in "
                "a host segmentation fault whose cause is not found); the rules apply");
  char const *wname = w->name;
  bool const force = flags & PALIWA_TUNE_FORCE;
  bool const measure = flags & PALIWA_TUNE_MEASURE;
  TuneCache cache = read_cache(cache_path);
  int32_t const rule_box[3] = {b->box_capacity, b->box_width, b->box_threads};
  int32_t const rule_lanes = b->nlanes, rule_weight = b->long_weight;
  auto set_boxes = [&](int32_t cap, int32_t width, int32_t threads) {
    b->box_capacity = cap;
    b->box_width = cap > 0 ? width : 0;
    b->box_threads = cap > 0 ? threads : 256;
    return build_structure(b);
  };

  // the box schedule the cache holds (applied first: the groups' keys depend
  // on it)
  std::string const bkey = boxes_key(*b, wname, forward);
  int32_t cached_box[3] = {0, 0, 0};
  bool search_boxes = true;
  if (flags & PALIWA_TUNE_KEEP_BOXES) {
    cached_box[0] = rule_box[0];
    cached_box[1] = rule_box[1];
    cached_box[2] = rule_box[2];
    search_boxes = false;
  } else if (!force) {
    auto const it = cache.find(bkey);
    std::istringstream v(it == cache.end() ? std::string() : it->second);
    if (v >> cached_box[0] >> cached_box[1] >> cached_box[2] &&
        valid_boxes(cached_box[0], cached_box[1], cached_box[2]))
      search_boxes = false;
  }
  int rc = PALIWA_OK;
  if (!search_boxes)
    rc = set_boxes(cached_box[0], cached_box[1], cached_box[2]);

  // choices of the groups the cache answers, and what is left to search
  auto group_search = [&]() {
    std::vector<size_t> search;
    for (size_t i = 0; i < b->groups.size(); ++i) {
      auto const &grp = b->groups[i];
      if (grp.widths.empty())
        continue;
      auto const it = force ? cache.end() : cache.find(width_key(*b, grp, wname, forward));
      int32_t const cached = it != cache.end() ? std::atoi(it->second.c_str()) : 0;
      if (std::find(grp.widths.begin(), grp.widths.end(), cached) == grp.widths.end())
        search.push_back(i);
    }
    return search;
  };
  std::vector<size_t> search = group_search();
  std::string skey = schedule_key(*b, wname, forward);
  bool search_schedule = uses_lanes(*b) && (force || cache.find(skey) == cache.end());
  b->evaluations = 0;
  b->rule_ms = b->tuned_ms = -1.0;
  bool const timing = measure || search_boxes || !search.empty() || search_schedule;

  void *backup = nullptr;
  size_t const es = b->dtype == PALIWA_FLOAT64 ? 8 : 4;
  std::vector<size_t> bytes(b->ngrids);
  size_t total = 0;
  for (int32_t g = 0; g < b->ngrids; ++g) {
    int64_t e = 1;
    for (int32_t d = 0; d < b->ndim; ++d)
      e *= b->extents_of(g)[d];
    bytes[g] = size_t(e) * es;
    total += bytes[g];
  }
  auto copy_all = [&](bool save) {
    size_t off = 0;
    for (int32_t g = 0; g < b->ngrids; ++g) {
      auto *bk = static_cast<unsigned char *>(backup) + off;
      hipError_t const e = save ? hipMemcpyAsync(bk, b->data[g], bytes[g],
                                                 hipMemcpyDeviceToDevice, s)
                                : hipMemcpyAsync(b->data[g], bk, bytes[g],
                                                 hipMemcpyDeviceToDevice, s);
      if (e != hipSuccess)
        return fail(PALIWA_ERR_HIP, std::string("HIP: ") + hipGetErrorString(e));
      off += bytes[g];
    }
    return hipStreamSynchronize(s) == hipSuccess ? PALIWA_OK
                                                 : fail(PALIWA_ERR_HIP, "copy failed");
  };
  // candidates are timed as the plan executes them: as graph replays (each
  // candidate's graph captured in its warm-up) unless graphs are switched off
  if (rc == PALIWA_OK && timing) {
    if (hipMalloc(&backup, total) != hipSuccess) {
      backup = nullptr;
      (void)hipGetLastError();
      rc = fail(PALIWA_ERR_HIP, "no device memory for the tuning copy of the grids");
    } else {
      rc = copy_all(true);
    }
  }
  constexpr int kReps = 15;
  auto evaluate = [&](double *ms) {
    ++b->evaluations;
    return time_direction(b, wname, forward, s, kReps, ms);
  };
  // the plan as the rules made it
  if (rc == PALIWA_OK && timing) {
    rc = set_boxes(rule_box[0], rule_box[1], rule_box[2]);
    if (rc == PALIWA_OK)
      rc = evaluate(&b->rule_ms);
  }
  double best = b->rule_ms;

  // the box schedule: none, or capacity x tile row (128 / 256 / 512 bytes) x
  // threads, the whole direction timed for each
  if (rc == PALIWA_OK && search_boxes) {
    int32_t keep[3] = {rule_box[0], rule_box[1], rule_box[2]};
    std::vector<std::array<int32_t, 3>> candidates = {{0, 0, 256}};
    for (int32_t cap : kBoxCapacities)
      for (int32_t row : {128, 256, 512})
        for (int32_t threads : {256, 512})
          candidates.push_back({cap, int32_t(row / int32_t(es)), threads});
    for (auto const &c : candidates) {
      if (rc != PALIWA_OK)
        break;
      if (c[0] == rule_box[0] && (c[0] == 0 || (c[1] == rule_box[1] && c[2] == rule_box[2])))
        continue;
      double ms = 0.0;
      if ((rc = set_boxes(c[0], c[1], c[2])) == PALIWA_OK && (rc = evaluate(&ms)) == PALIWA_OK &&
          ms < 0.99 * best) {
        best = ms;
        keep[0] = c[0];
        keep[1] = c[1];
        keep[2] = c[2];
      }
    }
    cached_box[0] = keep[0];
    cached_box[1] = keep[1];
    cached_box[2] = keep[2];
  }
  if (rc == PALIWA_OK)
    rc = set_boxes(cached_box[0], cached_box[1], cached_box[2]);
  if (rc == PALIWA_OK && search_boxes) { // the groups of the chosen schedule
    search = group_search();
    skey = schedule_key(*b, wname, forward);
    search_schedule = uses_lanes(*b) && (force || cache.find(skey) == cache.end());
  }
  bool const concurrent = uses_lanes(*b);

  // apply the cached choices of the groups
  if (rc == PALIWA_OK) {
    for (auto &grp : b->groups)
      if (!grp.widths.empty() && !force) {
        auto const it = cache.find(width_key(*b, grp, wname, forward));
        if (it != cache.end()) {
          int32_t const t = std::atoi(it->second.c_str());
          if (std::find(grp.widths.begin(), grp.widths.end(), t) != grp.widths.end())
            grp.threads = t;
        }
      }
    if (!search_schedule && concurrent && !force) {
      std::istringstream v(cache[skey]);
      int32_t nl = 0, lw = 0;
      if (v >> nl >> lw && nl >= 1 && nl <= kMaxLanes && lw >= 0) {
        b->nlanes = nl;
        b->long_weight = lw;
      }
    }
    rc = build_tables(b);
    if (rc == PALIWA_OK)
      rc = setup_lanes(b);
  }
  if (rc == PALIWA_OK && timing)
    rc = evaluate(&best);
  // coordinate descent: one choice at a time, kept when at least 1% faster.
  // `search` indexes the groups of the chosen box schedule; after a failed
  // step the plan may hold another candidate's groups, so it is not touched.
  if (rc == PALIWA_OK)
    std::sort(search.begin(), search.end(), [&](size_t x, size_t y) {
      return b->groups[x].lines * (int64_t(1) << b->groups[x].L) >
             b->groups[y].lines * (int64_t(1) << b->groups[y].L);
    });
  for (int round = 0; round < 2 && rc == PALIWA_OK; ++round) {
    for (size_t i : search) {
      auto &grp = b->groups[i];
      int32_t const keep = grp.threads;
      int32_t chosen = keep;
      for (int32_t t : grp.widths) {
        if (t == keep || rc != PALIWA_OK)
          continue;
        grp.threads = t;
        double ms = 0.0;
        if ((rc = build_tables(b)) == PALIWA_OK && (rc = evaluate(&ms)) == PALIWA_OK &&
            ms < 0.99 * best) {
          best = ms;
          chosen = t;
        }
      }
      grp.threads = chosen;
      if (rc == PALIWA_OK)
        rc = build_tables(b);
    }
    if (search_schedule && rc == PALIWA_OK) {
      int32_t keep_l = b->nlanes, keep_w = b->long_weight;
      for (int32_t nl : {2, 3, 4, 6, 8})
        for (int32_t lw : {1, 4, 16}) {
          if ((nl == keep_l && lw == keep_w) || rc != PALIWA_OK)
            continue;
          b->nlanes = nl;
          b->long_weight = lw;
          double ms = 0.0;
          if ((rc = setup_lanes(b)) == PALIWA_OK && (rc = evaluate(&ms)) == PALIWA_OK &&
              ms < 0.99 * best) {
            best = ms;
            keep_l = nl;
            keep_w = lw;
          }
        }
      b->nlanes = keep_l;
      b->long_weight = keep_w;
      if (rc == PALIWA_OK)
        rc = setup_lanes(b);
    }
    if (search.empty())
      break;
  }
  if (timing)
    b->tuned_ms = best;
  if (backup != nullptr) {
    int const restore = copy_all(false);
    (void)hipFree(backup);
    if (rc == PALIWA_OK)
      rc = restore;
  }
  if (rc != PALIWA_OK) { // leave the plan as the rules made it
    std::string const msg = last_error();
    b->nlanes = rule_lanes;
    b->long_weight = rule_weight;
    (void)set_boxes(rule_box[0], rule_box[1], rule_box[2]);
    return fail(rc, msg);
  }
  if (cache_path != nullptr) {
    if (!(flags & PALIWA_TUNE_KEEP_BOXES))
      cache[bkey] = std::to_string(b->box_capacity) + " " + std::to_string(b->box_width) + " " +
                    std::to_string(b->box_threads);
    for (auto const &grp : b->groups)
      if (!grp.widths.empty())
        cache[width_key(*b, grp, wname, forward)] = std::to_string(grp.threads);
    if (concurrent)
      cache[skey] = std::to_string(b->nlanes) + " " + std::to_string(b->long_weight);
    rc = write_cache(cache_path, cache);
  }
  return rc;
}

} // namespace

extern "C" {

int paliwa_batch_create(paliwa_batch **plan, int32_t dtype, int32_t ndim,
                        int32_t ngrids, void *const *data,
                        int64_t const *extents, int32_t const *levels,
                        int32_t const *minimum_level,
                        int32_t const *maximum_level) {
  if (plan == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  *plan = nullptr;
  if (ngrids < 1 || data == nullptr || extents == nullptr || levels == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT,
                "need ngrids >= 1 and data, extents, levels arrays");
  if (ndim < 1 || ndim > PALIWA_MAX_DIMS)
    return fail(PALIWA_ERR_UNSUPPORTED, "ndim must be in [1, 26]");
  for (int32_t g = 0; g < ngrids; ++g) {
    int64_t elements = 0;
    int rc = check_grid(data[g], dtype, ndim, extents + size_t(g) * ndim,
                        levels + size_t(g) * ndim, maximum_level, &elements);
    if (rc == PALIWA_OK)
      rc = check_level_range(ndim, levels + size_t(g) * ndim, minimum_level,
                             maximum_level);
    if (rc != PALIWA_OK)
      return fail(rc, "grid " + std::to_string(g) + ": " + last_error());
  }
  auto *b = new paliwa_batch;
  b->dtype = dtype;
  b->ndim = ndim;
  b->ngrids = ngrids;
  b->level.assign(levels, levels + size_t(ngrids) * ndim);
  b->extents.assign(extents, extents + size_t(ngrids) * ndim);
  b->lmin.assign(minimum_level, minimum_level + ndim);
  b->lmax.assign(maximum_level, maximum_level + ndim);
  b->data.assign(data, data + ngrids);
  if (hipGetDevice(&b->device) != hipSuccess) {
    release(b);
    return fail(PALIWA_ERR_HIP, "no HIP device");
  }

  // the box schedule where it reads at most three quarters of what the
  // per-axis sweeps read
  int64_t const per_axis = traffic(*b);
  b->box_capacity = kBoxRuleCapacity;
  b->box_width = int32_t(kBoxRuleRowBytes / (dtype == PALIWA_FLOAT64 ? 8 : 4));
  b->box_threads = dtype == PALIWA_FLOAT64 ? 256 : 512;
  if (4 * traffic(*b) > 3 * per_axis) {
    b->box_capacity = 0;
    b->box_width = 0;
  }
  int rc = build_structure(b);
  if (rc != PALIWA_OK) {
    std::string const msg = last_error();
    release(b);
    return fail(rc, msg);
  }
  *plan = b;
  last_error().clear();
  return PALIWA_OK;
}

int paliwa_batch_hierarchize(paliwa_batch *plan, char const *wavelet_name,
                             void *stream) {
  return execute(plan, wavelet_name, true, stream);
}

int paliwa_batch_dehierarchize(paliwa_batch *plan, char const *wavelet_name,
                               void *stream) {
  return execute(plan, wavelet_name, false, stream);
}

int paliwa_batch_set_graphs(paliwa_batch *plan, int32_t enable) {
  if (plan == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  plan->use_graphs = enable != 0;
  return PALIWA_OK;
}

int paliwa_batch_set_boxes(paliwa_batch *plan, int32_t capacity_bytes, int32_t width,
                           int32_t threads) {
  if (plan == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  bool const known = capacity_bytes == 0 || capacity_bytes == kBoxCapacities[0] ||
                     capacity_bytes == kBoxCapacities[1] || capacity_bytes == kBoxCapacities[2];
  if (!known)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "box capacity must be 0, 40960, 81920 or 163840");
  if (capacity_bytes > 0 &&
      (width < 1 || width > 4096 || (width & (width - 1)) != 0 || (threads != 256 && threads != 512)))
    return fail(PALIWA_ERR_INVALID_ARGUMENT,
                "box width must be a power of two in [1, 4096] and threads 256 or 512");
  int32_t const keep[3] = {plan->box_capacity, plan->box_width, plan->box_threads};
  plan->box_capacity = capacity_bytes;
  plan->box_width = capacity_bytes > 0 ? width : 0;
  plan->box_threads = capacity_bytes > 0 ? threads : 256;
  int rc = build_structure(plan);
  if (rc != PALIWA_OK) {
    std::string const msg = last_error();
    plan->box_capacity = keep[0];
    plan->box_width = keep[1];
    plan->box_threads = keep[2];
    (void)build_structure(plan);
    return fail(rc, msg);
  }
  last_error().clear();
  return PALIWA_OK;
}

int paliwa_batch_get_boxes(paliwa_batch const *plan, int32_t *capacity_bytes, int32_t *width,
                           int32_t *threads) {
  if (plan == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (capacity_bytes != nullptr)
    *capacity_bytes = plan->box_capacity;
  if (width != nullptr)
    *width = plan->box_width;
  if (threads != nullptr)
    *threads = plan->box_threads;
  return PALIWA_OK;
}

int paliwa_batch_traffic(paliwa_batch const *plan, int64_t *points_read) {
  if (plan == nullptr || points_read == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan and points_read are required");
  *points_read = traffic(*plan);
  return PALIWA_OK;
}

int paliwa_batch_info(paliwa_batch const *plan, int32_t *launches,
                      int64_t *points, int64_t *descriptor_bytes) {
  if (plan == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  int32_t n = 0;
  for (auto const &grp : plan->groups)
    n += grp.box || grp.fused || grp.long_ladder || grp.lds_lines ? 1
                                                                   : int32_t(grp.members.size());
  int64_t total = 0;
  for (int32_t g = 0; g < plan->ngrids; ++g) {
    int64_t e = 1;
    for (int32_t d = 0; d < plan->ndim; ++d)
      e *= plan->extents_of(g)[d];
    total += e;
  }
  if (launches != nullptr)
    *launches = n;
  if (points != nullptr)
    *points = total;
  if (descriptor_bytes != nullptr)
    *descriptor_bytes = int64_t(plan->dev_bytes + plan->box_dev_bytes);
  return PALIWA_OK;
}

int paliwa_batch_tune(paliwa_batch *plan, char const *wavelet_name, int32_t forward,
                      char const *cache_path, int32_t flags, void *stream) {
  if (plan == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  int current = 0;
  if (hipGetDevice(&current) != hipSuccess || current != plan->device)
    return fail(PALIWA_ERR_INVALID_ARGUMENT,
                "the plan belongs to device " + std::to_string(plan->device));
  int const rc = tune(plan, wavelet_name, forward != 0, cache_path, flags,
                      static_cast<hipStream_t>(stream));
  if (rc == PALIWA_OK)
    last_error().clear();
  return rc;
}

int paliwa_batch_tune_info(paliwa_batch const *plan, double *rule_ms, double *tuned_ms,
                           int32_t *evaluations, int32_t *streams) {
  if (plan == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (rule_ms != nullptr)
    *rule_ms = plan->rule_ms;
  if (tuned_ms != nullptr)
    *tuned_ms = plan->tuned_ms;
  if (evaluations != nullptr)
    *evaluations = plan->evaluations;
  if (streams != nullptr)
    *streams = uses_lanes(*plan) ? plan->nlanes : 0;
  return PALIWA_OK;
}

int paliwa_batch_destroy(paliwa_batch *plan) {
  if (plan != nullptr)
    release(plan);
  return PALIWA_OK;
}

} // extern "C"
