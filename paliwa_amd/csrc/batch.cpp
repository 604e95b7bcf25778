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
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

using namespace paliwa_hip;

struct paliwa_batch {
  int device = 0;
  int32_t dtype = 0, ndim = 0, ngrids = 0;
  std::vector<int32_t> level;   // ngrids * ndim
  std::vector<int64_t> extents; // ngrids * ndim
  std::vector<int32_t> lmin, lmax;
  std::vector<void *> data;

  struct Group {
    int32_t dim = 0, L = 0, m = 0;
    bool strided = false;
    bool fused = false;
    bool lds_lines = false;          // longer lines: one batched LDS-line launch
    int64_t tiles = 0;               // fused groups: tiles; LDS-line groups: lines
    int32_t threads = 0;             // fused strided groups: tile width (0: default)
    BatchGrid *grids_dev = nullptr;  // fused groups: descriptors on the device
    int32_t *tile_grid_dev = nullptr;
    std::vector<int32_t> members;    // grid indices
  };
  std::vector<Group> groups; // in axis order
  void *dev = nullptr;       // one allocation holding every descriptor table
  size_t dev_bytes = 0;

  // internal streams (default priority) and fork / join events
  std::vector<hipStream_t> lanes;
  std::vector<hipEvent_t> joins;
  hipEvent_t fork = nullptr;
  std::vector<int32_t> lane_of; // per group: its internal stream

  bool use_graphs = true;
  std::mutex mutex;
  std::map<std::pair<int, int>, hipGraphExec_t> graphs; // (wavelet, forward)

  int32_t const *level_of(int g) const { return &level[size_t(g) * ndim]; }
  int64_t const *extents_of(int g) const { return &extents[size_t(g) * ndim]; }
};

namespace {

int launch_group(paliwa_batch &b, paliwa_batch::Group const &grp,
                 paliwa_filter const *filters, int nfilters, bool forward,
                 hipStream_t stream) {
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
    // the groups of one axis
    size_t j = i;
    while (j < b.groups.size() && b.groups[j].dim == b.groups[i].dim)
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
      // long-line groups first: few workgroups each, latency-bound, so they
      // run beside the large groups instead of after them
      for (int pass = 0; pass < 2 && rc == PALIWA_OK; ++pass)
        for (size_t k = i; k < j && rc == PALIWA_OK; ++k)
          if (b.groups[k].lds_lines == (pass == 0))
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
  delete b;
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

  // Group the sweeps: per axis, grids with equal (line period, stored extent).
  std::vector<std::vector<BatchGrid>> host_grids;
  std::vector<std::vector<int32_t>> host_tiles;
  for (int32_t d = 0; d < ndim; ++d) {
    std::map<std::pair<int32_t, int32_t>, std::vector<int32_t>> by_shape;
    for (int32_t g = 0; g < ngrids; ++g)
      if (b->level_of(g)[d] > b->lmin[d])
        by_shape[{b->level_of(g)[d], int32_t(b->extents_of(g)[d])}].push_back(g);
    for (auto &[shape, members] : by_shape) {
      paliwa_batch::Group grp;
      grp.dim = d;
      grp.L = shape.first;
      grp.m = shape.second;
      grp.strided = d + 1 < ndim;
      grp.members = members;
      grp.fused = (int64_t(1) << grp.L) <= fused_max_period(dtype);
      grp.lds_lines = !grp.fused && (int64_t(1) << grp.L) <= lds_line_max_period(dtype);
      std::vector<BatchGrid> gd;
      std::vector<int32_t> tg;
      if (grp.fused && grp.strided) {
        // the tile width of a strided float ladder follows the group's lines
        int64_t lines = 0;
        for (int32_t g : members) {
          SweepParams const q = make_params(b->data[g], ndim, b->extents_of(g), b->level_of(g), d);
          lines += q.A * q.B;
        }
        int const K = fused_chunk(dtype, grp.L);
        grp.threads = ladder_threads_for(K, (1 << grp.L) / K, dtype == PALIWA_FLOAT64 ? 8 : 4, lines);
      }
      if (grp.fused || grp.lds_lines) {
        // fused: tiles of fused_tile_lines lines (or threads / G); LDS-line:
        // one line per tile
        int64_t const tb = !grp.fused ? 1
                           : grp.threads ? grp.threads / ((1 << grp.L) / fused_chunk(dtype, grp.L))
                                         : fused_tile_lines(dtype, grp.L, grp.strided);
        for (size_t k = 0; k < members.size(); ++k) {
          int32_t const g = members[k];
          SweepParams const p =
              make_params(b->data[g], ndim, b->extents_of(g), b->level_of(g), d);
          int64_t const tiles = (p.A * p.B + tb - 1) / tb;
          gd.push_back(BatchGrid{b->data[g], p.A, p.B, grp.tiles});
          tg.insert(tg.end(), size_t(tiles), int32_t(k));
          grp.tiles += tiles;
        }
        if (grp.tiles > 0x7fffffffLL) {
          release(b);
          return fail(PALIWA_ERR_UNSUPPORTED, "batch too large");
        }
      }
      b->groups.push_back(std::move(grp));
      host_grids.push_back(std::move(gd));
      host_tiles.push_back(std::move(tg));
    }
  }

  // One device allocation for all descriptor tables.
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
  if (bytes > 0) {
    std::vector<unsigned char> staging(bytes, 0);
    for (size_t i = 0; i < b->groups.size(); ++i) {
      std::memcpy(staging.data() + grid_off[i], host_grids[i].data(),
                  host_grids[i].size() * sizeof(BatchGrid));
      std::memcpy(staging.data() + tile_off[i], host_tiles[i].data(),
                  host_tiles[i].size() * sizeof(int32_t));
    }
    if (hipMalloc(&b->dev, bytes) != hipSuccess ||
        hipMemcpy(b->dev, staging.data(), bytes, hipMemcpyHostToDevice) !=
            hipSuccess) {
      release(b);
      return fail(PALIWA_ERR_HIP, "could not place the batch descriptors");
    }
    b->dev_bytes = bytes;
    for (size_t i = 0; i < b->groups.size(); ++i) {
      auto *base = static_cast<unsigned char *>(b->dev);
      b->groups[i].grids_dev = reinterpret_cast<BatchGrid *>(base + grid_off[i]);
      b->groups[i].tile_grid_dev = reinterpret_cast<int32_t *>(base + tile_off[i]);
    }
  }
  // Lanes for the groups of each axis: largest group first onto the least
  // loaded lane.  Work estimate in swept points: the group's points plus a
  // fixed cost per launch (about 10 us of launch, fill and drain, the time the
  // sweep needs for ~3.75 M float64 points), four times that for a long-line
  // group, whose few workgroups make it latency-bound.
  constexpr int kLanes = 4; // 3: 1.57 ms, 8: 1.61 ms, 4: 1.43 ms on the 1325-grid set
  b->lane_of.assign(b->groups.size(), 0);
  bool concurrent = false;
  for (size_t i = 0; i < b->groups.size();) {
    size_t j = i;
    while (j < b->groups.size() && b->groups[j].dim == b->groups[i].dim)
      ++j;
    if (j - i > 1) {
      concurrent = true;
      std::vector<std::pair<int64_t, size_t>> work;
      for (size_t k = i; k < j; ++k) {
        int64_t pts = 0;
        for (int32_t g : b->groups[k].members) {
          int64_t e = 1;
          for (int32_t d = 0; d < ndim; ++d)
            e *= b->extents_of(g)[d];
          pts += e;
        }
        int64_t const launch = dtype == PALIWA_FLOAT64 ? 3750000 : 7500000;
        pts += b->groups[k].lds_lines ? 4 * launch : launch;
        work.push_back({pts, k});
      }
      std::sort(work.begin(), work.end(), [](auto const &x, auto const &y) {
        return x.first != y.first ? x.first > y.first : x.second < y.second;
      });
      int64_t load[kLanes] = {};
      for (auto const &[pts, k] : work) {
        int best = 0;
        for (int t = 1; t < kLanes; ++t)
          if (load[t] < load[best])
            best = t;
        load[best] += pts;
        b->lane_of[k] = best;
      }
    }
    i = j;
  }
  if (concurrent) {
    b->lanes.assign(kLanes, nullptr);
    b->joins.assign(kLanes, nullptr);
    bool ok = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess;
    for (int t = 0; t < kLanes && ok; ++t)
      ok = hipStreamCreateWithFlags(&b->lanes[t], hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&b->joins[t], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
      release(b);
      return fail(PALIWA_ERR_HIP, "could not create the plan's streams");
    }
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

int paliwa_batch_info(paliwa_batch const *plan, int32_t *launches,
                      int64_t *points, int64_t *descriptor_bytes) {
  if (plan == nullptr)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "plan is NULL");
  int32_t n = 0;
  for (auto const &grp : plan->groups)
    n += grp.fused ? 1 : int32_t(grp.members.size());
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
    *descriptor_bytes = int64_t(plan->dev_bytes);
  return PALIWA_OK;
}

int paliwa_batch_destroy(paliwa_batch *plan) {
  if (plan != nullptr)
    release(plan);
  return PALIWA_OK;
}

} // extern "C"
