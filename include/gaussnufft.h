/*
 * gaussnufft.h -- C ABI of the MI355X (gfx950) Gaussian-gridding NUFFT.
 *
 * Plain C: opaque plan handle, device pointers, sizes, int status codes.
 * Complex data are interleaved complex64 (float re, float im).  All device
 * pointers must be HIP device memory of the process's HIP runtime; the plan
 * runs every kernel and FFT on the HIP stream it is bound to.
 *
 * Reference interfaces replaced (paths in the mripy repository):
 *   gn_grid_params   <- _compute_2d_grid_params / _compute_3d_grid_params
 *                       fft/nufft_func.py:215-226, 234-249
 *   gn_plan_create   <- NUFFT3d_cuda.__init__ pics/operators_cuda_class.py:65-83 and the
 *                       per-call parameter set-up of nufft3d*_gaussker_cuda
 *                       fft/nufft_func_cuda.py:1463-1470, 1512-1518
 *   gn_set_points    <- NUFFT3d_cuda.set_ktraj pics/operators_cuda_class.py:92-96 (x,y,z,dcf)
 *   gn_execute       <- GN_TYPE2  : nufft3d2_gaussker_cuda fft/nufft_func_cuda.py:1512-1559,
 *                                   nufft2d2_gaussker_cuda :609-636, NUFFT3d_cuda.forward
 *                                   pics/operators_cuda_class.py:107-111
 *                       GN_TYPE1  : density_weighting + nufft3d1_gaussker_cuda
 *                                   fft/nufft_func_cuda.py:1463-1509, nufft2d1_gaussker_cuda
 *                                   :581-606, NUFFT3d_cuda.backward operators_cuda_class.py:113-120
 *                       GN_TYPE21 : nufft3d21_gaussker_cuda fft/nufft_func_cuda.py:1562-1637,
 *                                   nufft2d21_gaussker_cuda :639-679,
 *                                   NUFFT3d_cuda.forward_backward operators_cuda_class.py:122-127
 *                       GN_ADJOINT: exact adjoint of GN_TYPE2 (no counterpart; M * type 1
 *                                   without density weighting, for autograd and CG)
 *   gn_plan_destroy  <- (Python garbage collection of the operator object)
 *   gn_last_error    <- the Python exceptions the reference raises (ValueError for eps,
 *                       fft/nufft_func.py:200-202)
 */
#ifndef GAUSSNUFFT_H
#define GAUSSNUFFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GN_VERSION 1

/* status codes */
#define GN_OK 0
#define GN_EINVAL 1  /* invalid argument (sizes, eps range, null pointer, kind) */
#define GN_ENOMEM 2  /* device allocation failed */
#define GN_EHIP 3    /* HIP runtime error */
#define GN_EROCFFT 4 /* rocFFT error */
#define GN_ESTATE 5  /* call out of order (e.g. execute before set_points) */

/* data layouts of images and samples at the boundary */
#define GN_COIL_LAST 0  /* reference layout: image (m1,m2[,m3],nc), samples (M,nc), C order */
#define GN_COIL_MAJOR 1 /* image (nc,m1,m2[,m3]), samples (nc,M) */

/* transform kinds for gn_execute */
#define GN_TYPE1 1   /* image = (1/M) (pi/tau)^(d/2) e^{tau|k|^2} crop(ifft(spread(dcf*y))) */
#define GN_TYPE2 2   /* y = (1/G) gather(fft(pad((pi/tau)^(d/2) e^{tau|k|^2} image)))       */
#define GN_ADJOINT 3 /* exact adjoint of GN_TYPE2: M * GN_TYPE1 without dcf                  */
#define GN_TYPE21 21 /* GN_TYPE1(GN_TYPE2(image)) with dcf                                   */
/* SENSE kinds (maps set by gn_plan_set_sense): the image is one coil-less image x of
 * modes[0..dim-1] points; the samples keep their nc coils.                             */
#define GN_SENSE_TYPE2 32    /* y = GN_TYPE2(S x): coil c's image is x * S_c                  */
#define GN_SENSE_TYPE1 31    /* x = sum_c conj(S_c) GN_TYPE1(y)_c (with dcf)                  */
#define GN_SENSE_ADJOINT 33  /* exact adjoint of GN_SENSE_TYPE2 (no dcf, no 1/M)              */
#define GN_SENSE_TYPE21 321  /* GN_SENSE_TYPE1(GN_SENSE_TYPE2(x)) with dcf                    */
/* The gridding step alone (the reference's build_grid_{2d,3d}{1,2}_fast,
 * fft/nufft_func.py:550-639, 863-1055), on a caller grid laid out coil-major
 * (nc, nf[0], .., nf[d-1]) complex64:                                                  */
#define GN_GRID_SPREAD 41    /* grid[g] = sum_j w_j(g) y_j (no dcf, no scale)               */
#define GN_GRID_GATHER 42    /* y_j = (1/G) sum_g w_j(g) grid[g]                            */

/* FFT engines for gn_plan_set_fft_engine */
#define GN_FFT_AUTO 0   /* pruned in-place passes fused with pad/crop when every nf is a product of
                           2, 3 and 5 (<= 2048); rocFFT otherwise */
#define GN_FFT_ROCFFT 1 /* separate pad, rocFFT and crop */

typedef struct gn_plan_s gn_plan;

/* Library version (GN_VERSION). */
int gn_version(void);

/* Last error message of the calling thread ("" if none). */
const char *gn_last_error(void);

/* Grid parameters from eps, the reference rule (fft/nufft_func.py:198-249):
 * rat = 2 if eps > 1e-11 else 3; nspread = int(-ln(eps)/(pi(rat-1)/(rat-0.5)) + 0.5),
 * clamped to [3, max_nspread] when max_nspread > 0 (the 3-D drivers pass 6);
 * nf[a] = max(rat*modes[a], 2*nspread); tau = pi*nspread/(rat(rat-0.5))/modes[0]^2.
 * Returns GN_EINVAL unless 1e-33 < eps < 1e-1. */
int gn_grid_params(int dim, const int64_t *modes, double eps, int max_nspread, int *rat, int *nspread,
                   int64_t *nf, double *tau);

/* Create a plan for a dim-D (2 or 3) transform of an image of size modes[0..dim-1]
 * with ncoils coils, oversampling rat and half-width nspread (the window spans
 * 2*nspread taps per axis).  Supported nspread: 2..8 in 3-D (the reference drivers
 * clamp it to 3..6), 2..16 in 2-D (eps down to about 1e-18 at rat 3; smaller eps
 * gives GN_EINVAL).  The default gather tile shrinks for wide windows so that tile
 * plus halo fits the LDS.  iflag >= 0: type 2 uses e^{-ik.x}; iflag < 0 flips both
 * signs.  stream: hipStream_t (NULL = default stream). */
int gn_plan_create(int dim, const int64_t *modes, int rat, int nspread, int iflag, int64_t ncoils, int layout,
                   void *stream, gn_plan **plan);

/* Rebind the plan to another HIP stream. */
int gn_plan_set_stream(gn_plan *plan, void *stream);

/* Set the M sample coordinates in radians (device pointers; z ignored for dim 2),
 * float32 when coord_f64 == 0 else float64, each multiplied by df.  dcf: optional
 * float32 density weights (device pointer, M entries) applied by GN_TYPE1/GN_TYPE21.
 * Bins and sorts the samples; the caller may free its buffers afterwards. */
int gn_set_points(gn_plan *plan, int64_t M, const void *x, const void *y, const void *z, int coord_f64, double df,
                  const float *dcf);

/* Run one transform (kind = GN_TYPE1/GN_TYPE2/GN_ADJOINT/GN_TYPE21 or a GN_SENSE_*
 * kind) on the plan's stream.  in/out: device complex64 buffers in the plan's layout.
 * Asynchronous. */
int gn_execute(gn_plan *plan, int kind, const void *in, void *out);

/* SENSE coil sensitivities for the GN_SENSE_* kinds (replaces the composition
 * joint2operators(espirit(S), NUFFT3d_cuda), pics/operators_class.py:437-472,496-514:
 * forward = NUFFT.forward(x * S), backward = sum_c conj(S_c) NUFFT.backward(y)_c with
 * keepdims, forward_backward = espirit.backward(NUFFT.forward_backward(espirit.forward(x)))).
 * maps: device complex64, the plan's image layout with nc coils (coil-last
 * (m0,m1[,m2],nc) or coil-major); copied into the plan.  NULL detaches them.
 * The expansion x * S_c is fused into the first FFT pass and the combine
 * sum_c conj(S_c) into the last one (one block holds all coils of an image line when
 * they fit the LDS; otherwise the coil images go through a plan buffer). */
int gn_plan_set_sense(gn_plan *plan, const void *maps);

/* GN_TYPE21 with the intermediate samples in the library's static device scratch
 * instead of plan memory.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation); the call makes no hipMalloc / hipFree and no allocating library
 * call.  Limits: M * ncoils <= 4,194,304 (a 32 MiB array per GPU, held until the
 * process ends; larger problems give GN_EINVAL), the pruned FFT passes (grid lengths
 * with prime factors 2, 3, 5 and the default engine, else GN_ESTATE) and stage timers
 * off.  Every call on one GPU shares that GPU's scratch: calls that overlap on one GPU
 * (several streams or threads) race; serialize them.  Output equals gn_execute(plan,
 * GN_TYPE21, in, out) bit for bit. */
int gn_execute_type21_noalloc(gn_plan *plan, const void *in, void *out);

/* Gradient of the forward map with respect to the sample coordinates
 * (no counterpart in the reference; for autograd): for image f and complex
 * per-sample weights w (device, sample layout of the plan),
 *   grad[a*M + j] = sum_c Re( conj(w[j,c]) * d(GN_TYPE2 f)[j,c] / dx_a[j] ),
 * from the analytic derivative of the Gaussian kernel, in the original sample
 * order (float32 device buffer of dim*M).  With w = dL/dy this is dL/dx for
 * y = A(x) f; with f = dL/d(image) and w = y it is dL/dx for image = A(x)^H y. */
int gn_traj_grad(gn_plan *plan, const void *image, const void *weights, float *grad);

/* gn_traj_grad with flags (no counterpart in the reference; for autograd, whose
 * backward of y = A(x) f follows the forward's type 2 of the same f).
 * GN_TRAJ_REUSE_GRID: when the plan's last operation was GN_TYPE2 of this same
 * image pointer, or a trajectory gradient of it, the plan's grid still holds that
 * image's type-2 grid and the FFT passes are skipped; the caller guarantees that the
 * image memory was not changed since.  Otherwise, and without the flag, the grid is
 * recomputed as in gn_traj_grad.  Results are identical either way. */
#define GN_TRAJ_REUSE_GRID 1
int gn_traj_grad_ex(gn_plan *plan, const void *image, const void *weights, float *grad, int flags);

/* Query: nf[3] oversampled grid, nspread, tau, number of samples, coils, and
 * tile[6] = spread (output) tile shape followed by the gather (home) tile shape. */
int gn_plan_info(const gn_plan *plan, int64_t *nf, int *nspread, double *tau, int64_t *M, int64_t *ncoils,
                 int *tile);

/* Override the tile shapes, tile[6] = spread tile (3 axes) then gather tile (3 axes),
 * in cells; 0 keeps the default.  Invalidates the binned points: call gn_set_points
 * again before gn_execute.  Spread tiles are at most 32 cells per axis; a gather tile
 * plus its halo (tile + 2 nspread - 1 cells per axis, 16 bytes per cell: a coil pair)
 * must fit 160 KiB of LDS, else GN_EINVAL. */
int gn_plan_set_tiles(gn_plan *plan, const int *tile);

/* Stage timers (replaces the wall-clock utc.timing prints around the gridding and FFT
 * stages, utilities/utilities_class.py:30-60, fft/nufft_func_cuda.py:1472-1496).
 * When enabled, gn_execute records a HIP event pair around every stage on the plan's
 * stream.  gn_plan_stage_times synchronizes the stream and returns, per stage, the
 * accumulated milliseconds and launch counts (GN_STAGE_COUNT entries each). */
#define GN_STAGE_PAD 0     /* deapodize + zero-pad + transpose (type 2)  */
#define GN_STAGE_FFT2 1    /* oversampled FFT of type 2                  */
#define GN_STAGE_GATHER 2  /* interpolation onto the samples (type 2)    */
#define GN_STAGE_SPREAD 3  /* gridding of the samples (type 1)           */
#define GN_STAGE_FFT1 4    /* oversampled FFT of type 1                  */
#define GN_STAGE_CROP 5    /* crop + deapodize + scale + transpose       */
#define GN_STAGE_COUNT 6
int gn_plan_set_timing(gn_plan *plan, int enable);
int gn_plan_stage_times(gn_plan *plan, double *ms, int64_t *counts, int reset);

/* Copy the sample order chosen by gn_set_points (for each sorted position the
 * original sample index; M ints) to host memory.  No counterpart in the
 * reference: it exposes the stable counting-sort result to tests and tools. */
int gn_plan_sample_order(const gn_plan *plan, int *order_host);

/* select the FFT engine (GN_FFT_AUTO or GN_FFT_ROCFFT); results agree to FP32 roundoff */
int gn_plan_set_fft_engine(gn_plan *plan, int engine);

int gn_plan_destroy(gn_plan *plan);

/* ---- Cartesian FFTs -------------------------------------------------------------
 * Batched complex-to-complex FFT over the leading `rank` (1..3) axes of a C-order
 * complex64 array of shape (shape[0], .., shape[rank-1], batch) whose batch elements
 * (coils) are interleaved: element (i, b) at i * batch + b.  Replaces the cuFFT
 * wrappers fftnc2c_cuda / ifftnc2c_cuda / fft2c2c_cuda / ifft2c2c_cuda
 * (fft/cufft.py:27-97: forward unnormalized, inverse divided by prod(shape)) and, with
 * the shift flags and the mask, the transforms of FFTnd_cuda_kmask / FFT2d_cuda_kmask
 * (pics/operators_cuda_class.py:7-58: forward = ifftshift(FFT(fftshift(im))) * mask,
 * backward = ifftshift(IFFT(fftshift(ksp))), shifts over the transformed axes).
 * mask: null or float32 of prod(shape) entries (broadcast over the batch, mask_batch
 * = 1) or prod(shape) * batch entries (mask_batch = batch); it multiplies the output.
 * in and out may be the same buffer.  Host-synchronous on `stream` (the shifted copy
 * uses one workspace per device): calls from several threads are serialized. */
#define GN_FFT_FORWARD 1
#define GN_FFT_INVERSE -1
#define GN_FFT_SHIFT_IN 1  /* fftshift the input over the transformed axes  */
#define GN_FFT_SHIFT_OUT 2 /* ifftshift the output over the transformed axes */
int gn_fft_c2c(int rank, const int64_t *shape, int64_t batch, int direction, int flags, const float *mask,
               int64_t mask_batch, const void *in, void *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GAUSSNUFFT_H */
