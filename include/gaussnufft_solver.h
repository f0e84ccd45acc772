/*
 * gaussnufft_solver.h -- device primitives of the iterative reconstructions that
 * call the NUFFT operator (conjugate gradient, backtracking line search, ADMM with
 * a total-variation prox, IST/FISTA).  Same library and conventions as
 * gaussnufft.h: plain C, device pointers, int status codes, gn_last_error().
 *
 * Reference interfaces replaced (paths in the mripy repository):
 *   gn_dot_f64     <- np.dot(p.flatten(), np.conj(df(x)).flatten()) in
 *                     BacktrackingLineSearch pics/opt_alg.py:20
 *   gn_sqdist_f64  <- np.linalg.norm(Afunc(xi)-b)**2 and np.linalg.norm(dx) in
 *                     prox_l2_Afxnb_CGD / prox_l2_Afxnb_CGD2 pics/proximal_func.py:287,
 *                     296, 303, 319, 328, 335; np.linalg.norm(x-z) in
 *                     ADMM_l2Afxnb_tvx pics/CS_MRI_solvers_func.py:354
 *   gn_tv_prox     <- prox_tv2d_r pics/proximal_func.py:114-133 with the TV2d_r operator
 *                     (grad, Div, amp) pics/tvop_class.py:59-129 (ndim 2), and prox_tv3d /
 *                     prox_tv3d_r pics/proximal_func.py:136-177 with TV3d / TV3d_r
 *                     pics/tvop_class.py:131-253 (ndim 3)
 *
 * Reductions accumulate in FP64 in a fixed order that depends only on n, so a
 * result is bitwise reproducible run to run.  They need a caller-provided device
 * workspace of gn_reduce_work_bytes() bytes and write their result to DEVICE
 * memory (stream ordered; the caller synchronizes before reading it on the host).
 */
#ifndef GAUSSNUFFT_SOLVER_H
#define GAUSSNUFFT_SOLVER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types of solver vectors */
#define GN_F32 1  /* float                          */
#define GN_C64 2  /* complex64: interleaved floats  */
#define GN_C128 3 /* complex128: interleaved doubles */

/* TV amplitude rules for gn_tv_prox */
#define GN_TV_AMP_SQUARE 0  /* sqrt(sum_a G_a^2) with complex squares and numpy's lexicographic
                               complex maximum against 1, exactly as TV2d_r.amp + np.maximum
                               (pics/tvop_class.py:116-124, pics/proximal_func.py:128) */
#define GN_TV_AMP_MODULUS 1 /* sqrt(sum_a |G_a|^2) compared with 1 (textbook isotropic TV) */

/* bytes of device workspace the reductions need */
int64_t gn_reduce_work_bytes(void);

/* out[0] + i*out[1] = sum_i conj(a_i) * b_i over n elements of type dtype, in FP64.
   For GN_F32 out[1] = 0.  out: 2 doubles of device memory. */
int gn_dot_f64(int dtype, int64_t n, const void *a, const void *b, void *work, double *out, void *stream);

/* out[0] = sum_i |a_i - b_i|^2 in FP64 (b may be NULL: sum_i |a_i|^2).  out: 1 double, device. */
int gn_sqdist_f64(int dtype, int64_t n, const void *a, const void *b, void *work, double *out, void *stream);

/* bytes of device workspace gn_tv_prox needs for an array of ndim image axes
   shape[0..ndim-1] followed by nt trailing elements */
int64_t gn_tv_work_bytes(int ndim, const int64_t *shape, int64_t nt);

/* f = prox of lambda * TV over the first ndim (2 or 3) axes of a C-ordered array y
   of shape (shape[0], .., shape[ndim-1], nt), nt = product of the trailing dims, by
   niter projected dual gradient steps of size step (Chambolle), computed in FP64:
     G = 0; repeat niter: G = G - step*grad(Div(G) - y/lambda); G = G / max(amp(G), 1)
     f = y - lambda * Div(G)
   grad = forward differences with a zero last slice per axis, Div = its adjoint.
   y and f have type dtype (GN_C64 or GN_C128) and may not alias; every image axis
   has at least 2 samples.  shape is a host array. */
int gn_tv_prox(int dtype, int ndim, const int64_t *shape, int64_t nt, const void *y, double lambda, double step,
               int niter, int amp_rule, void *work, void *f, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GAUSSNUFFT_SOLVER_H */
