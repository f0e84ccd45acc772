/*
 * gaussnufft_batch.h -- many NUFFT instances of uneven sizes in one set of launches,
 * optionally replayed from captured HIP graphs, and a self-written tile autotuner.
 * Same library and conventions as gaussnufft.h.
 *
 * Reference interfaces replaced (paths in the mripy repository): none directly.
 * The reference processes such workloads one operator call at a time:
 * NUFFT3d_cuda.forward / backward (pics/operators_cuda_class.py:107-120) per
 * instance, each call launching its own numba kernels and cuFFT plans
 * (fft/nufft_func_cuda.py:1463-1559).  gn_batch_execute computes, for every
 * instance i, exactly what gn_execute on a plan of instance i computes.
 *
 * A batch owns one plan per instance (records binned at creation) and one
 * oversampled-grid buffer in which instances of equal grid size are adjacent.
 * Each stage (gather, spread, split-tile reduce, and the pruned FFT passes
 * fused with pad / crop) is one kernel launch over the work items of all
 * instances (a few launches when instances differ in grid length, coil-count
 * class or tile shape).  When a grid length is not a product of 2, 3 and 5 up
 * to 2048, pad, one batched rocFFT execution per grid size and crop are used
 * instead of the pruned passes.
 */
#ifndef GAUSSNUFFT_BATCH_H
#define GAUSSNUFFT_BATCH_H

#include <stdint.h>

#include "gaussnufft.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gn_batch_s gn_batch;

/* one instance: image modes, coils, samples and its trajectory (device memory;
   copied at creation, so the caller may free it afterwards) */
typedef struct {
    int64_t modes[3];     /* first dim entries used */
    int64_t ncoils;
    int64_t M;
    const void *x, *y, *z; /* radians, float (coord_f64 = 0) or double (coord_f64 = 1); z unused in 2-D */
    const float *dcf;      /* M density weights or NULL */
} gn_instance;

/* All instances share dim, oversampling rat, nspread, iflag, layout, coordinate
   precision and df.  Work runs on stream. */
int gn_batch_create(int dim, int n, const gn_instance *inst, int rat, int nspread, int iflag, int layout,
                    int coord_f64, double df, void *stream, gn_batch **batch);

/* in[i] / out[i]: host arrays of n device pointers (layouts as gn_execute).
   kind: GN_TYPE1, GN_TYPE2, GN_ADJOINT or GN_TYPE21. */
int gn_batch_execute(gn_batch *batch, int kind, const void *const *in, void *const *out);

/* enable != 0: gn_batch_execute replays a HIP graph captured on its first call for
   each kind (on a stream of the batch's own, ordered after and before the caller's
   stream with events); pointer changes only update a device argument table. */
int gn_batch_set_graph(gn_batch *batch, int enable);

/* spread tiles tile[0..2], gather tiles tile[3..5] (0 = default), for every
   instance; re-bins all instances */
int gn_batch_set_tiles(gn_batch *batch, const int *tile);

/* Times `reps` executions of `kind` for each candidate tile configuration and
   keeps the fastest.  Candidates (those that fit the LDS): 3-D spread tiles
   4x4x8, 4x8x8, 8x8x8, 4x4x16, 8x8x16 (2 to 16 waves of 64-cell blocks per tile)
   times gather tiles 10x8x12, 10x10x16, 12x12x12, 8x8x16, 8x8x8; 2-D spread tiles
   8x16, 16x8, 16x16, 8x32 times gather tiles 24x24, 16x32, 32x32, 32x64, 48x48.
   tile (6 ints) and best_ms receive the choice (may be NULL). */
int gn_batch_autotune(gn_batch *batch, int kind, const void *const *in, void *const *out, int reps, int *tile,
                      double *best_ms);

/* n instances, current tiles (6), kernel launches per gn_batch_execute of each kind
   [type2, type1] */
int gn_batch_info(const gn_batch *batch, int *n, int *tile, int *launches);

int gn_batch_destroy(gn_batch *batch);

/* Single-plan autotuner: as gn_batch_autotune for one plan; the coordinates are
   passed again because every candidate re-bins them (gn_set_points arguments). */
int gn_plan_autotune(gn_plan *plan, int kind, int64_t M, const void *x, const void *y, const void *z, int coord_f64,
                     double df, const float *dcf, const void *in, void *out, int reps, int *tile, double *best_ms);

#ifdef __cplusplus
}
#endif

#endif /* GAUSSNUFFT_BATCH_H */
