/* paliwa_hip.h -- C ABI of the MI355X (gfx950) lifting-wavelet transform.
 *
 * Every entry point takes plain pointers and sizes, returns an int status
 * (PALIWA_OK or a negative PALIWA_ERR_* code with a message available from
 * paliwa_last_error()), is stream-ordered and asynchronous on `stream` (a
 * hipStream_t; NULL is the default stream), never synchronizes the device and
 * allocates no device memory (paliwa_batch_create alone allocates, once, the
 * descriptor tables of its plan).  Arguments are validated on the host before
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

#define PALIWA_FLOAT64 0
#define PALIWA_FLOAT32 1

#define PALIWA_MAX_DIMS 26     /* the reference defines 26 axis tags,
                                  paliwa/paliwa_dimensions.hpp:10-69 */
#define PALIWA_MAX_FILTERS 16  /* passes per filter list */
#define PALIWA_MAX_LEVELS 64   /* entries of a level list */
#define PALIWA_MAX_LEVEL 40    /* per-axis level */

/* Kernel schedules for paliwa_transform_in_path. */
#define PALIWA_PATH_AUTO 0     /* best available (fused when it fits) */
#define PALIWA_PATH_FUSED 1    /* one launch per axis: lines staged in LDS,
                                  every level and pass in registers */
#define PALIWA_PATH_LDS_LINE 2 /* one launch per axis: one line per workgroup
                                  in LDS, a barrier between passes */
#define PALIWA_PATH_PER_PASS 3 /* one launch per (level, pass) straight on HBM:
                                  the launch structure of the reference's
                                  detail::lifting_pass */

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

/* Kernel launches per direction, total stored points, descriptor bytes. */
int paliwa_batch_info(const paliwa_batch *plan, int32_t *launches,
                      int64_t *points, int64_t *descriptor_bytes);

int paliwa_batch_destroy(paliwa_batch *plan);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* PALIWA_HIP_H */
