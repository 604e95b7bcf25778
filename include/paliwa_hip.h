/* paliwa_hip.h -- C ABI of the MI355X (gfx950) lifting-wavelet transform.
 *
 * Every entry point takes plain pointers and sizes, returns an int status
 * (PALIWA_OK or a negative PALIWA_ERR_* code with a message available from
 * paliwa_last_error()), is stream-ordered and asynchronous on `stream` (a
 * hipStream_t; NULL is the default stream), never synchronizes the device and
 * allocates no device memory (the batched plans' set-up calls alone allocate:
 * paliwa_batch_create and paliwa_batch_set_boxes the descriptor tables of a
 * plan, paliwa_batch_tune a copy of its grids for the duration of the call).
 * Arguments are validated on the host before
 * any HIP call, so a rejected call has no effect on the device.
 *
 * Data layout: a dense row-major array (last axis contiguous, as DDC's
 * layout_right; paliwa/paliwa_distributed_transform.hpp:184), in device
 * memory, of float64 or float32.  Along axis d the extent is 2^level[d] (the
 * reference's periodic lattice, paliwa/paliwa_domains.hpp:33-49) or
 * 2^level[d]+1 (the lattice plus a periodic image at index 2^level[d], which
 * each sweep along d rewrites as a copy of index 0; its input value is not
 * read).  The transform is in place.
 *
 * The reference is header-only C++ templates over DDC spans; these functions
 * are what its templates reduce to once the span is a pointer plus extents.
 * include/paliwa/paliwa_transform.hpp wraps them in the reference's own
 * names, argument order, defaults and exception types.
 */
#ifndef PALIWA_HIP_H
#define PALIWA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; the C++ header maps each to the exception type the reference
 * raises in the same situation. */
#define PALIWA_OK 0
#define PALIWA_ERR_INVALID_ARGUMENT (-1) /* std::invalid_argument */
#define PALIWA_ERR_UNKNOWN_WAVELET (-2)  /* std::out_of_range, map::at
                                            (paliwa/paliwa_transform.hpp:370,397) */
#define PALIWA_ERR_FILTER_OFFSET (-3)    /* std::runtime_error("Filter offset not
                                            supported") (paliwa_transform.hpp:239) */
#define PALIWA_ERR_LEVEL (-4)            /* level ranges: the reference asserts
                                            (paliwa_transform.hpp:232, 359-362) */
#define PALIWA_ERR_SHAPE (-5)            /* extent is neither 2^l nor 2^l+1 */
#define PALIWA_ERR_DTYPE (-6)
#define PALIWA_ERR_UNSUPPORTED (-7)      /* size limits below */
#define PALIWA_ERR_HIP (-8)              /* a HIP runtime call failed */
#define PALIWA_ERR_IO (-9)               /* a file could not be opened, read or
                                            written */

#define PALIWA_FLOAT64 0
#define PALIWA_FLOAT32 1

#define PALIWA_MAX_DIMS 26     /* the reference defines 26 axis tags,
                                  paliwa/paliwa_dimensions.hpp:10-69 */
#define PALIWA_MAX_FILTERS 16  /* passes per filter list */
#define PALIWA_MAX_LEVELS 64   /* entries of a level list */
#define PALIWA_MAX_LEVEL 40    /* maximum_level of an axis; sizes the
                                  learned transform's statistics */
#define PALIWA_MAX_AXIS_LEVEL 30 /* stored level of an axis: an extent is at
                                    most 2^30 + 1 points (lines are indexed
                                    with 32-bit integers); above it a call
                                    returns PALIWA_ERR_UNSUPPORTED */

/* Kernel schedules for paliwa_transform_in_path. */
#define PALIWA_PATH_AUTO 0     /* the long-line kernel's column tiles for
                                  strided float64 lines of 256 points on large
                                  grids (paliwa_auto_path reports the choice),
                                  else fused where it holds the line, else the
                                  long-line kernel (strided axes, contiguous
                                  lines its staging buffer holds), else the
                                  LDS-line kernel, else per pass */
#define PALIWA_PATH_FUSED 1    /* one launch per axis: lines staged in LDS,
                                  every level and pass in registers */
#define PALIWA_PATH_LDS_LINE 2 /* one launch per axis: one line per workgroup
                                  in LDS, a barrier between passes */
#define PALIWA_PATH_PER_PASS 3 /* one launch per (level, pass) straight on HBM:
                                  the launch structure of the reference's
                                  detail::lifting_pass */
#define PALIWA_PATH_LONG_LADDER 4 /* lines longer than the fused kernel holds
                                     (float64 2^12..2^15, float32 2^13..2^16
                                     points), and strided float64 lines of
                                     2^8 points (column tiles of 16-point
                                     chunks), named transforms' level lists:
                                     one launch per axis, a tile of lines
                                     across a workgroup, fine levels in
                                     registers, coarse ones in LDS */

/* One lifting pass: points of parity `offset` (0 even, 1 odd) at the operating
 * level get  g = c0*g[lower] + c1*g + c2*g[upper]  (evaluated left to right).
 * Mirrors std::pair<int, std::array<double,3>> of paliwa/paliwa_wavelets.hpp. */
typedef struct paliwa_filter {
  int32_t offset;
  int32_t reserved; /* must be 0 */
  double coefficients[3];
} paliwa_filter;

/* Message of the last failed call on this thread ("" if none). */
const char *paliwa_last_error(void);

/* Library version string. */
const char *paliwa_version(void);

/* Filter table lookup: paliwa/paliwa_wavelets.hpp:15-28.  `reconstruct` = 0
 * selects lifting_wavelet_filter_offsets_and_coefficients, 1 selects
 * lifting_wavelet_reconstruct_offsets_and_coefficients.  Writes up to
 * `capacity` passes to `out` and the table's length to *count.
 * PALIWA_ERR_UNKNOWN_WAVELET for a name not in the table. */
int paliwa_wavelet_filters(const char *wavelet_name, int32_t reconstruct,
                           paliwa_filter *out, int32_t capacity,
                           int32_t *count);

/* Replaces paliwa::transform_in<Dim>(grid, level, maximum_level, levels,
 * filters, instance) for a complete grid (paliwa/paliwa_transform.hpp:215-251):
 * for each entry l of levels[0..nlevels) in order (1 <= l <= level[dim]) and
 * each filter in order, one lifting pass along axis `dim` at storage stride
 * 2^(level[dim]-l), period 2^level[dim].  maximum_level may be NULL (it only
 * scales the reference's coordinates; it is checked against level when
 * given). */
int paliwa_transform_in(void *data, int32_t dtype, int32_t ndim,
                        const int64_t *extents, const int32_t *level,
                        const int32_t *maximum_level, int32_t dim,
                        const int32_t *levels, int32_t nlevels,
                        const paliwa_filter *filters, int32_t nfilters,
                        void *stream);

/* paliwa_transform_in with an explicit kernel schedule (PALIWA_PATH_*).
 * PALIWA_ERR_UNSUPPORTED when the schedule cannot hold the line. */
int paliwa_transform_in_path(void *data, int32_t dtype, int32_t ndim,
                             const int64_t *extents, const int32_t *level,
                             const int32_t *maximum_level, int32_t dim,
                             const int32_t *levels, int32_t nlevels,
                             const paliwa_filter *filters, int32_t nfilters,
                             int32_t path, void *stream);

/* The kernel schedule (PALIWA_PATH_*) that PALIWA_PATH_AUTO takes for the
 * named transform of `wavelet_name` along `dim` (forward: hierarchize_in's
 * level list, else dehierarchize_in's), or PALIWA_PATH_AUTO itself when the
 * axis has no level to sweep (minimum_level[dim] == level[dim]).  No
 * counterpart in the reference: introspection of this library's schedule
 * rules, for callers that tune around them.  Launches nothing, needs no
 * device, and reads no grid data. */
int paliwa_auto_path(int32_t dtype, int32_t ndim, const int64_t *extents,
                     const int32_t *level, const int32_t *minimum_level, int32_t dim,
                     const char *wavelet_name, int32_t forward, int32_t *path);

/* Replaces paliwa::hierarchize_in<Dim> (paliwa/paliwa_transform.hpp:346-372):
 * levels level[dim], level[dim]-1, ..., minimum_level[dim]+1 with the forward
 * table of `wavelet_name` ("hat", "biorthogonal", "fullweighting"). */
int paliwa_hierarchize_in(void *data, int32_t dtype, int32_t ndim,
                          const int64_t *extents, const int32_t *level,
                          const int32_t *minimum_level,
                          const int32_t *maximum_level, int32_t dim,
                          const char *wavelet_name, void *stream);

/* Replaces paliwa::dehierarchize_in<Dim> (paliwa/paliwa_transform.hpp:374-399):
 * levels minimum_level[dim]+1, ..., level[dim] with the reconstruct table. */
int paliwa_dehierarchize_in(void *data, int32_t dtype, int32_t ndim,
                            const int64_t *extents, const int32_t *level,
                            const int32_t *minimum_level,
                            const int32_t *maximum_level, int32_t dim,
                            const char *wavelet_name, void *stream);

/* Replaces paliwa::hierarchize (paliwa/paliwa_transform.hpp:401-415):
 * hierarchize_in along axes 0, 1, ..., ndim-1. */
int paliwa_hierarchize(void *data, int32_t dtype, int32_t ndim,
                       const int64_t *extents, const int32_t *level,
                       const int32_t *minimum_level,
                       const int32_t *maximum_level, const char *wavelet_name,
                       void *stream);

/* Replaces paliwa::dehierarchize (paliwa/paliwa_transform.hpp:417-431):
 * dehierarchize_in along axes 0, 1, ..., ndim-1 (the same axis order as the
 * forward transform, as in the reference). */
int paliwa_dehierarchize(void *data, int32_t dtype, int32_t ndim,
                         const int64_t *extents, const int32_t *level,
                         const int32_t *minimum_level,
                         const int32_t *maximum_level,
                         const char *wavelet_name, void *stream);

/* ---- Partial buffers ---------------------------------------------------- *
 * The reference's transform_in takes a checked branch for every buffer that is
 * not the complete strided lattice (paliwa/paliwa_transform.hpp:224-228,
 * 241-248): detail::lifting_pass<CheckBounds = true> (:24-51) updates a point
 * of the pass's half lattice only when it and both neighbours are stored
 * (:31-32, 42-45).  A partial buffer is a dense row-major array over the
 * positions of each axis; paliwa_axis_storage gives the finest-level
 * coordinate (units of 2^-maximum_level[d]) of every position of axis d:
 *   affine:  coordinates == NULL, coordinate(p) = front + p * step -- DDC's
 *            DiscreteDomain (step 1) and StridedDiscreteDomain;
 *   list:    coordinates = a device array of `count` ascending, distinct
 *            coordinates -- DDC's SparseDiscreteDomain, e.g. local points
 *            plus the ghosts of get_required_transform_domains
 *            (paliwa/paliwa_domains.hpp:590-673).  Not read on the host.
 * Every coordinate lies in [0, 2^maximum_level[d]); there are no periodic
 * image planes.  A pass along `dim` at level l visits coordinates of parity
 * offset * 2^(M-l) mod 2^(M-l+1) along dim (M = maximum_level[dim]) and
 * multiples of 2^(maximum_level[d] - level[d]) along the other axes; other
 * stored points keep their values.  A buffer that is the complete lattice
 * of every axis (front 0, step 2^(maximum_level - level), count 2^level) runs
 * the unchecked kernels, as the reference does.  Results are bit-identical to
 * the reference's checked branch. */
typedef struct paliwa_axis_storage {
  int64_t count;              /* positions along the axis, at most 2^30 + 1 */
  int64_t front;              /* affine: coordinate of position 0 (>= 0) */
  int64_t step;               /* affine: spacing of the coordinates (>= 1) */
  const int64_t *coordinates; /* list: device array, or NULL for affine */
} paliwa_axis_storage;

/* Replaces paliwa::transform_in<Dim> on a partial buffer
 * (paliwa/paliwa_transform.hpp:215-251, CheckBounds branch). */
int paliwa_transform_in_checked(void *data, int32_t dtype, int32_t ndim,
                                const paliwa_axis_storage *storage,
                                const int32_t *level,
                                const int32_t *maximum_level, int32_t dim,
                                const int32_t *levels, int32_t nlevels,
                                const paliwa_filter *filters, int32_t nfilters,
                                void *stream);

/* Replace paliwa::hierarchize_in<Dim> / dehierarchize_in<Dim> on a partial
 * buffer (paliwa/paliwa_transform.hpp:346-399). */
int paliwa_hierarchize_in_checked(void *data, int32_t dtype, int32_t ndim,
                                  const paliwa_axis_storage *storage,
                                  const int32_t *level,
                                  const int32_t *minimum_level,
                                  const int32_t *maximum_level, int32_t dim,
                                  const char *wavelet_name, void *stream);
int paliwa_dehierarchize_in_checked(void *data, int32_t dtype, int32_t ndim,
                                    const paliwa_axis_storage *storage,
                                    const int32_t *level,
                                    const int32_t *minimum_level,
                                    const int32_t *maximum_level, int32_t dim,
                                    const char *wavelet_name, void *stream);

/* Replace paliwa::hierarchize / dehierarchize on a partial buffer
 * (paliwa/paliwa_transform.hpp:401-431): axes 0..ndim-1, as in the test of
 * test/test_paliwa_restrict.cpp:94-139. */
int paliwa_hierarchize_checked(void *data, int32_t dtype, int32_t ndim,
                               const paliwa_axis_storage *storage,
                               const int32_t *level,
                               const int32_t *minimum_level,
                               const int32_t *maximum_level,
                               const char *wavelet_name, void *stream);
int paliwa_dehierarchize_checked(void *data, int32_t dtype, int32_t ndim,
                                 const paliwa_axis_storage *storage,
                                 const int32_t *level,
                                 const int32_t *minimum_level,
                                 const int32_t *maximum_level,
                                 const char *wavelet_name, void *stream);

/* ---- Batched transform of many grids (combination-technique sets) ------ */

/* A plan over `ngrids` grids of one dtype and rank: grid g is data[g] with
 * extents extents[g*ndim .. g*ndim+ndim) and level levels[g*ndim ..]; all
 * grids share minimum_level and maximum_level.  The plan groups the axis
 * sweeps of all grids by (axis, line period, stored extent) so that one
 * launch covers every grid of a group, and replays a whole direction as one
 * HIP graph.  Creating a plan validates every grid and allocates its
 * descriptor tables on the current device (the only allocation; executing a
 * plan allocates nothing).  The grids' memory must stay valid while the plan
 * is used.  One plan must not execute concurrently on two streams.
 * Replaces the reference's per-grid loop of paliwa::hierarchize calls over a
 * combination-technique set (test/test_paliwa_combination_technique.cpp:
 * 164-174); the result of each grid is bit-identical to hierarchize on it. */
typedef struct paliwa_batch paliwa_batch;

int paliwa_batch_create(paliwa_batch **plan, int32_t dtype, int32_t ndim,
                        int32_t ngrids, void *const *data,
                        const int64_t *extents, const int32_t *levels,
                        const int32_t *minimum_level,
                        const int32_t *maximum_level);

/* hierarchize (forward) / dehierarchize (inverse) of every grid of the plan,
 * asynchronous on `stream`. */
int paliwa_batch_hierarchize(paliwa_batch *plan, const char *wavelet_name,
                             void *stream);
int paliwa_batch_dehierarchize(paliwa_batch *plan, const char *wavelet_name,
                               void *stream);

/* 0: launch the grouped kernels directly instead of replaying a graph. */
int paliwa_batch_set_graphs(paliwa_batch *plan, int32_t enable);

/* Box schedule (no counterpart in the reference).  Instead of one launch per
 * group of axis sweeps, each grid's axes split into boxes of consecutive
 * axes a..b: the full extents of those axes for one index of the axes before
 * a and a tile of `width` consecutive indices of the flattened axes after b
 * (all of them when the box then fits).  A box is swept in LDS, every axis of
 * it with all its levels, so it costs one HBM read and one write for its axes
 * together; launch p sweeps the p-th box of every grid.  Each grid takes the
 * fewest boxes that fit `capacity_bytes` of LDS (40960, 81920 or 163840: four,
 * two or one box per CU).  width: a power of two in [1, 4096]; threads: 256
 * or 512 per box.  capacity_bytes 0: the per-axis groups.  Rebuilds the
 * plan's descriptor tables (allocation, synchronous).  A new plan takes
 * boxes of 40960 bytes with 512-byte tile rows (256 threads for float64, 512
 * for float32) when they read at most three quarters of what the groups
 * read.  Results are bit-identical
 * either way.  Fails, leaving the plan as it was, when an axis of some grid
 * does not fit a box alone. */
int paliwa_batch_set_boxes(paliwa_batch *plan, int32_t capacity_bytes,
                           int32_t width, int32_t threads);
int paliwa_batch_get_boxes(const paliwa_batch *plan, int32_t *capacity_bytes,
                           int32_t *width, int32_t *threads);

/* Points one direction of the plan reads from device memory (and writes
 * back): each grid once per segment of its schedule -- once per axis with
 * levels without boxes, once per box segment with them (with the empty part
 * of partial tiles) -- the model of the plan's memory traffic. */
int paliwa_batch_traffic(const paliwa_batch *plan, int64_t *points_read);

/* Kernel launches per direction, total stored points, descriptor bytes. */
int paliwa_batch_info(const paliwa_batch *plan, int32_t *launches,
                      int64_t *points, int64_t *descriptor_bytes);

int paliwa_batch_destroy(paliwa_batch *plan);

/* Autotuning of a plan (no counterpart in the reference, whose per-grid loop
 * has nothing to tune).  Times alternatives for one wavelet and direction on
 * the plan's own grids and keeps the fastest, by coordinate descent on the
 * whole direction's time (median of 15 replays, a choice kept when at least
 * 1% faster):
 *   - the box schedule (paliwa_batch_set_boxes): none, or each LDS capacity
 *     with tile rows of 128, 256 or 512 bytes and 256 or 512 threads (unless
 *     flags has PALIWA_TUNE_KEEP_BOXES: the plan's setting stays);
 *   - then, for the groups of that schedule, the tile width (256 / 512 / 1024
 *     threads, where the tile fits in LDS) of every strided group whose lines
 *     hold 512 points or more;
 *   - the stream schedule: 2, 3, 4, 6 or 8 streams for the groups of an axis,
 *     and the launch-cost weight (1, 4, 16) of long-line groups in the
 *     balance across them.
 * The grids are copied to a temporary device buffer first and restored
 * afterwards, bit for bit, so the call allocates memory the size of the set
 * for its duration and synchronizes `stream`; results of the plan do not
 * depend on any choice.  cache_path (may be NULL): a text file of earlier
 * choices, one per line ("<key> = <value>", keys naming dtype, layout, line
 * period, stored extent, line-count bucket, wavelet and direction for widths,
 * a hash of the plan's groups for the schedule, a hash of the grid shapes
 * for the box schedule); choices found there are
 * applied without timing and the file is rewritten with every choice of this
 * plan.  When the cache holds every choice nothing is timed or allocated
 * unless flags has PALIWA_TUNE_MEASURE. */
#define PALIWA_TUNE_FORCE 1   /* ignore the cache: time every choice */
#define PALIWA_TUNE_MEASURE 2 /* time the rule-based and final settings */
#define PALIWA_TUNE_KEEP_BOXES 4 /* keep the plan's box schedule */
int paliwa_batch_tune(paliwa_batch *plan, const char *wavelet_name,
                      int32_t forward, const char *cache_path, int32_t flags,
                      void *stream);
/* The last tuning: rule-based and tuned time of the direction in ms (-1 when
 * not timed), evaluations timed, streams per axis now in use. */
int paliwa_batch_tune_info(const paliwa_batch *plan, double *rule_ms,
                           double *tuned_ms, int32_t *evaluations,
                           int32_t *streams);

/* ---- Partial-output transform ------------------------------------------- *
 * The reference's transform_in with an `outputs` domain
 * (paliwa/paliwa_transform.hpp:253-344) and its dependency planner
 * (paliwa/paliwa_domains.hpp:315-487).
 *
 * An inclusive periodic interval of finest-level indices on a lattice: min >
 * max wraps around the period, {-1, -1} is empty (the reference's
 * PeriodicBounds). */
typedef struct paliwa_periodic_bounds {
  int64_t min;
  int64_t max;
} paliwa_periodic_bounds;

/* Replaces paliwa::transform_bounds (paliwa/paliwa_domains.hpp:337-444): the
 * exact writes of every (level, pass), writes[i * nfilters + p] on the
 * lattice of stride 2^(maximum_level - levels[i] + 1), and the required inputs
 * of every hierarchical level h = 0..level, required[h] (level 0: the point 0;
 * h >= 1: odd multiples of 2^(maximum_level - h)), for the output points
 * `outputs` (on the lattice of stride 2^(maximum_level - level)).  Levels must
 * be consecutive and monotone (PALIWA_ERR_INVALID_ARGUMENT, as the
 * reference's std::invalid_argument); filter coefficients are not read.
 * Host code: no device is touched.  writes and required may be NULL. */
int paliwa_transform_bounds(paliwa_periodic_bounds outputs, int32_t level,
                            int32_t maximum_level, const int32_t *levels,
                            int32_t nlevels, const paliwa_filter *filters,
                            int32_t nfilters, paliwa_periodic_bounds *writes,
                            paliwa_periodic_bounds *required);

/* Replaces paliwa::covering_periodic_bounds (paliwa/paliwa_domains.hpp:
 * 472-487): the smallest periodic interval holding the sorted, distinct
 * indices. */
int paliwa_covering_periodic_bounds(const int64_t *indices, int64_t count,
                                    int64_t period,
                                    paliwa_periodic_bounds *cover);

/* A partial-output transform along axis `dim`, stored at `level` (period
 * P = 2^level; storage indices).  The caller's block (dense row-major,
 * `extents`) holds the run [local_start, local_start + extents[dim]) mod P of
 * the axis; every line of the block is transformed; the outputs are the run
 * [output_start, output_start + output_count) mod P, inside the local run.
 * Levels and filters as paliwa_transform_in (consecutive, monotone levels).
 * The required inputs outside the local run are the ghosts
 * (paliwa_partial_ghosts: their storage indices, ascending); the caller passes
 * them as a second array shaped like the block with extents[dim] = their
 * count.  Executing writes the output points only: ghosts and the other local
 * points keep their values (the reference's packed-pole execution from local
 * and remote storage, paliwa/paliwa_transform.hpp:53-210; its direct mode
 * also leaves intermediate values in non-output points).  Outputs are
 * bit-identical to the full transform's.  Creating a plan allocates its
 * tables; a window of required inputs longer than 64 KiB per line is
 * PALIWA_ERR_UNSUPPORTED. */
typedef struct paliwa_partial paliwa_partial;

int paliwa_partial_create(paliwa_partial **plan, int32_t dtype, int32_t ndim,
                          const int64_t *extents, int32_t dim, int32_t level,
                          const int32_t *levels, int32_t nlevels,
                          const paliwa_filter *filters, int32_t nfilters,
                          int64_t local_start, int64_t output_start,
                          int64_t output_count);
/* *count = number of ghosts; indices (may be NULL) receives them. */
int paliwa_partial_ghosts(const paliwa_partial *plan, int64_t *count,
                          int64_t *indices);
int paliwa_partial_transform(const paliwa_partial *plan, void *local,
                             const void *ghosts, void *stream);
int paliwa_partial_destroy(paliwa_partial *plan);

/* ---- Raw dumps ---------------------------------------------------------- *
 * Replaces paliwa::dump_chunk_span_to_binary_file (paliwa/paliwa_io.hpp:
 * 13-23): the `bytes` bytes of a device array (a grid's elements, row-major)
 * to `filename`, headerless, native byte order -- the format the reference's
 * postprocess.py reads with numpy.fromfile.  Synchronizes `stream` (the copy
 * goes through two 16 MiB pinned host buffers, allocated on the first call
 * and kept until the process ends; calls are serialized). */
int paliwa_dump_to_binary_file(const void *data, int64_t bytes,
                               const char *filename, void *stream);
/* The inverse: a file of exactly `bytes` bytes into device memory
 * (PALIWA_ERR_SHAPE when the size differs). */
int paliwa_load_from_binary_file(void *data, int64_t bytes,
                                 const char *filename, void *stream);

/* ---- Combination-technique combine -------------------------------------- *
 * The steps after the batched hierarchize in the reference's combination
 * technique (test/test_paliwa_combination_technique.cpp:176-313): subspace
 * enumeration (paliwa/paliwa_utils.hpp:9-25), collection of the component
 * grids onto sparse-grid storage, scatter onto the full grid at the maximum
 * level.
 *
 * Grid g (dense row-major, extent 2^l or 2^l+1 per axis, dtype of the plan)
 * stores, at index j along axis d, the point j * 2^(maximum_level[d] - l_d)
 * of the full grid (paliwa/paliwa_domains.hpp:33-49).  Subspace k (0 <= k_d
 * <= l_d) holds the points (2m+1) 2^(M_d - k_d), m < 2^(k_d - 1), along each
 * axis with k_d >= 1 and the point 0 where k_d = 0
 * (paliwa/paliwa_domains.hpp:51-83).  The plan keeps the subspaces contained
 * in at least `min_count` grids (the reference keeps count > 1, i.e.
 * min_count = 2; 1 keeps every subspace), in lexicographic order of k, each
 * stored row-major over its points.  Sparse value at a point: the fold over
 * the grids containing its subspace, in grid order, of
 * s += coefficients[g] * grid_g (product then sum, each rounded; for float32
 * grids both in double and the sum stored as float).
 * Creating a plan validates the grids and allocates its tables (the only
 * allocation); the grids' memory must stay valid while the plan is used. */
#define PALIWA_COMBINE_MAX_DIMS 8
typedef struct paliwa_combine paliwa_combine;

int paliwa_combine_create(paliwa_combine **plan, int32_t dtype, int32_t ndim,
                          int32_t ngrids, void *const *data,
                          const int64_t *extents, const int32_t *levels,
                          const double *coefficients,
                          const int32_t *maximum_level, int32_t min_count);
/* Kept subspaces, sparse points, multiply-adds of a gather, table bytes. */
int paliwa_combine_info(const paliwa_combine *plan, int32_t *subspaces,
                        int64_t *sparse_points, int64_t *terms,
                        int64_t *table_bytes);
/* Host copies: levels [subspaces][ndim], offsets [subspaces + 1], counts
 * [subspaces] (grids containing each subspace); any may be NULL. */
int paliwa_combine_subspaces(const paliwa_combine *plan, int32_t *levels,
                             int64_t *offsets, int32_t *counts);
/* sparse (device, sparse_points elements) := the combined subspaces. */
int paliwa_combine_gather(paliwa_combine *plan, void *sparse, void *stream);
/* full (device, extent 2^maximum_level or 2^maximum_level+1 per axis; an
 * image plane repeats index 0) := 0 + sparse value at the points of kept
 * subspaces, 0 elsewhere. */
int paliwa_combine_scatter(const paliwa_combine *plan, const void *sparse,
                           void *full, const int64_t *full_extents,
                           void *stream);
int paliwa_combine_destroy(paliwa_combine *plan);

/* ---- Learned lifting (no counterpart in the reference) ------------------ *
 * An extension on top of the plain method: an integer lifting transform of an
 * int32 fixed-point grid whose predict step at levels >= 3 adds to the
 * interpolet (hat) prediction  round(clamp(unit * mlp(u), +-bound * unit)),
 * from c = the 4 coarse neighbours on each side along the lifted axis:
 *   features 0 (raw):       u = c / max|c|, unit = max|c|, bound 4;
 *   features 1 (detrended): D_k = 2 c_k - (c_3 + c_4) - (c_4 - c_3)(2k - 7),
 *                           the deviation from the interpolet line (times 2),
 *                           u = D / max|D|, unit = max|D| / 2, bound 1;
 * mlp = 8 -> 32 -> 32 -> 1 with GELU, bf16 weights and activations and fp32
 * accumulation on MFMA.  Integer lifting makes every step exactly invertible
 * (the inverse recomputes the bit-identical prediction), so the only round-trip
 * error is the quantization paliwa_quantize_f32 applies.  `update` = 1 adds the
 * integer CDF 5/3 update after each predict.  mlp_mask[d]: bit l set = level l
 * of axis d uses the MLP (else the interpolet, the fallback).  Forward axes run
 * 0..ndim-1, inverse ndim-1..0.  Weights: paliwa_learned_param_count() fp32 in
 * device memory, packed W1[32][8], b1[32], W2[32][32], b2[32], w3[32], b3.
 * Lines of at most 2^11 points.  `overflow` (device int32, may be NULL) is set
 * to 1 when an intermediate leaves the int32 range (the result is then not
 * invertible). */
int paliwa_learned_param_count(void);

/* out[i] = round(in[i] / scale) (int32); in[i] = out[i] * scale. */
int paliwa_quantize_f32(const float *in, int32_t *out, int64_t count,
                        float scale, void *stream);
int paliwa_dequantize_f32(const int32_t *in, float *out, int64_t count,
                          float scale, void *stream);

/* Forward transform along `axis` (or every axis, 0..ndim-1, for axis = -1).
 * features: 0 = the neighbours divided by their max-abs; 1 = the neighbours'
 * deviations from the interpolet line through the two nearest ones, divided
 * by their max-abs.  `stats` (device double[ndim][2*(PALIWA_MAX_LEVEL+1)], may
 * be NULL) accumulates, per axis and MLP level, sum |detail| with the MLP
 * prediction ([2l]) and with the interpolet ([2l+1]). */
int paliwa_learned_hierarchize(int32_t *data, int32_t ndim,
                               const int64_t *extents, const int32_t *level,
                               const int32_t *minimum_level, int32_t axis,
                               const float *weights, const uint32_t *mlp_mask,
                               int32_t update, int32_t features, double *stats,
                               int32_t *overflow, void *stream);
/* Inverse along `axis` (or every axis, ndim-1..0, for axis = -1). */
int paliwa_learned_dehierarchize(int32_t *data, int32_t ndim,
                                 const int64_t *extents, const int32_t *level,
                                 const int32_t *minimum_level, int32_t axis,
                                 const float *weights,
                                 const uint32_t *mlp_mask, int32_t update,
                                 int32_t features, int32_t *overflow,
                                 void *stream);

/* Training pairs from the interpolet transform of `data` (transformed in
 * place): at every axis and level >= 3, a sampled subset (hash below
 * sample_threshold, out of 2^32) of the predicted points writes
 * {8 network inputs, target (x - interpolet) / unit clamped to +-bound, key}
 * (10 floats) to pairs[*pair_count], up to `capacity` pairs. */
int paliwa_learned_extract_pairs(int32_t *data, int32_t ndim,
                                 const int64_t *extents, const int32_t *level,
                                 const int32_t *minimum_level, int32_t update,
                                 int32_t features, float *pairs,
                                 uint32_t *pair_count, uint32_t capacity,
                                 uint32_t sample_threshold, uint32_t salt,
                                 void *stream);

/* The network on inputs[count][8] through the MFMA path of the transform. */
int paliwa_learned_mlp_eval(const float *inputs, const float *weights,
                            float *out, int32_t count, void *stream);

/* Adds to grad the gradient of mean_i |mlp(u_i) - t_i| over rows
 * index[0..count) (or 0..count) of pairs (row stride `stride` floats: 8
 * inputs then the target), and the loss to *loss_sum.  `workspace`: device
 * floats, paliwa_learned_gradient_workspace(count) of them, for per-workgroup
 * partial sums that are added in a fixed order: the result is deterministic. */
int64_t paliwa_learned_gradient_workspace(int32_t count);
int paliwa_learned_gradient(const float *pairs, int32_t stride,
                            const int32_t *index, int32_t count,
                            const float *weights, float *grad,
                            float *loss_sum, float *workspace, void *stream);

/* paliwa_learned_gradient without a workspace argument.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * per-workgroup partial sums live in a static device array of the library
 * (1.4 MiB, one copy per GPU, held until the process ends), so count is at
 * most PALIWA_LEARNED_STATIC_MAX_ROWS (else PALIWA_ERR_UNSUPPORTED).  Nothing
 * is allocated; the first call on a device looks the array's address up
 * (hipGetSymbolAddress) and keeps it.  Same kernels and result as
 * paliwa_learned_gradient.  Calls that overlap on one GPU share the array and
 * race: order them on one stream or with events. */
#define PALIWA_LEARNED_STATIC_MAX_ROWS 16384
int paliwa_learned_gradient_static(const float *pairs, int32_t stride,
                                   const int32_t *index, int32_t count,
                                   const float *weights, float *grad,
                                   float *loss_sum, void *stream);

/* One Adam step (step = 1, 2, ...) on the fp32 weights; clears grad. */
int paliwa_learned_adam(float *weights, float *grad, float *m1, float *m2,
                        float lr, float beta1, float beta2, float eps,
                        int32_t step, void *stream);

/* p[0..count) = value. */
int paliwa_learned_fill(float *p, int32_t count, float value, void *stream);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* PALIWA_HIP_H */
