"""FP64 NumPy restatement of mripy's reconstruction callers of the NUFFT -- TEST INFRASTRUCTURE ONLY.

The iterative solvers that drive the NUFFT operator, restated from the reference's
Python 2 sources in complex128, call for call (no caching of repeated
evaluations), so that the HIP-path solvers in ``mripy_amd.solvers`` can be checked
against "the same algorithm in FP64".  Only ``tests/`` import this module.

Followed line by line (paths relative to the mripy repository root):

* ``BacktrackingLineSearch``            pics/opt_alg.py:8-33
* ``prox_l2_Afxnb_CGD``                 pics/proximal_func.py:282-311
* ``prox_l2_Afxnb_CGD2``                pics/proximal_func.py:314-343
* ``prox_l1_soft_thresh(2)``            pics/proximal_func.py:12-23
* ``prox_l1_Tf_soft_thresh2``           pics/proximal_func.py:44-52
* ``TV2d_r`` (grad, Div, amp)           pics/tvop_class.py:59-129
* ``prox_tv2d_r``                       pics/proximal_func.py:114-133
* ``TV3d``, ``prox_tv3d``               pics/tvop_class.py:131-188, pics/proximal_func.py:136-155
* ``ADMM_l2Afxnb_tvx``                  pics/CS_MRI_solvers_func.py:336-355
* ``IST_2/22/3/32``, ``IST_wrap``       pics/CS_MRI_solvers_func.py:30-110
* ``FIST_2/22/3/32``, ``FIST_wrap``     pics/CS_MRI_solvers_func.py:151-247
* ``optscaling``                        utilities/utilities_func.py:381-383
* ``espirit``, ``joint2operators``      pics/operators_class.py:437-472, 496-514
* operator ``NUFFT2d`` (forward = type 2, backward = type 1 of dcf*y, forward_backward
  = type 21)                            pics/operators_class.py:302-364

The per-iteration ``print`` calls of the reference are dropped.  Pins (tests/
test_solvers_oracle.py): CG and IST/FISTA on small dense operators against the
least-squares solution, the proximal CG against its normal equations, line-search
acceptance, ``Div`` as the exact adjoint of ``grad``, the TV prox of a constant and
of a two-level step against the closed-form TV-denoising solution, and the
reference's complex-amplitude rule against an independent element-wise evaluation.
"""
from __future__ import annotations

import numpy as np

from . import coracle as C
from . import nufft_oracle as O


# ----------------------------------------------------------------- line search / CG

def BacktrackingLineSearch(f, df, x, p, c=0.0001, rho=0.2, ls_Nite=10):
    """Armijo backtracking (pics/opt_alg.py:8-33); returns (alpha, steps)."""
    derphi = np.real(np.dot(p.flatten(), np.conj(df(x)).flatten()))
    f0 = f(x)
    alphak = 1.0
    f_try = f(x + alphak * p)
    i = 0
    while i < ls_Nite and f_try - f0 > c * alphak * derphi and f_try > f0:
        alphak = alphak * rho
        f_try = f(x + alphak * p)
        i += 1
    return alphak, i


def prox_l2_Afxnb_CGD(Afunc, invAfunc, b, x0, rho, Nite, ls_Nite=10):
    """Nonlinear CG on ||A x - b||^2 + (rho/2)||x - x0||^2 from x0 (pics/proximal_func.py:282-311)."""
    eps = 0.001
    i = 0

    def f(xi):
        return np.linalg.norm(Afunc(xi) - b) ** 2 + (rho / 2) * np.linalg.norm(xi - x0) ** 2

    def df(xi):
        return 2 * invAfunc(Afunc(xi) - b) + rho * (xi - x0)

    dx = -df(x0)
    alpha, nstp = BacktrackingLineSearch(f, df, x0, dx, ls_Nite=ls_Nite)
    x = x0 + alpha * dx
    s = dx
    delta0 = np.linalg.norm(dx)
    deltanew = delta0
    while i < Nite and deltanew > eps * delta0 and nstp < ls_Nite:
        dx = -df(x)
        deltaold = deltanew
        deltanew = np.linalg.norm(dx)
        beta = float(deltanew / float(deltaold))
        s = dx + beta * s
        alpha, nstp = BacktrackingLineSearch(f, df, x, s, ls_Nite=ls_Nite)
        x = x + alpha * s
        i = i + 1
    return x


def prox_l2_Afxnb_CGD2(Afunc, invAfunc, b, Nite, ls_Nite=10):
    """Nonlinear CG on ||A x - b||^2 from x = invA(b) (pics/proximal_func.py:314-343)."""
    x = invAfunc(b)
    eps = 0.001
    i = 0

    def f(xi):
        return np.linalg.norm(Afunc(xi) - b) ** 2

    def df(xi):
        return 2 * invAfunc(Afunc(xi) - b)

    dx = -df(x)
    alpha, nstp = BacktrackingLineSearch(f, df, x, dx, ls_Nite=ls_Nite)
    x = x + alpha * dx
    s = dx
    delta0 = np.linalg.norm(dx)
    deltanew = delta0
    while i < Nite and deltanew > eps * delta0 and nstp < ls_Nite:
        dx = -df(x)
        deltaold = deltanew
        deltanew = np.linalg.norm(dx)
        beta = float(deltanew / float(deltaold))
        s = dx + beta * s
        alpha, nstp = BacktrackingLineSearch(f, df, x, s, ls_Nite=ls_Nite)
        x = x + alpha * s
        i = i + 1
    return x


# ----------------------------------------------------------------- l1 proxes

def prox_l1_soft_thresh(x0, th):
    """exp(i angle(x)) * max(|x| - th, 0) (pics/proximal_func.py:12-16)."""
    a_th = np.abs(x0) - th
    a_th[a_th < 0] = 0
    return np.multiply(np.exp(1j * np.angle(x0)), a_th)


def prox_l1_soft_thresh2(x0, th):
    """x/(|x| + 1e-6) * max(|x| - th, 0) (pics/proximal_func.py:19-23)."""
    a_th = np.abs(x0) - th
    a_th[a_th < 0] = 0
    a_dir = np.divide(x0, np.abs(x0) + 1e-6)
    return np.multiply(a_dir, a_th)


def prox_l1_Tf_soft_thresh2(Tfunc, invTfunc, x0, th):
    """invT(soft2(T(x))) (pics/proximal_func.py:44-52)."""
    Tfx0 = Tfunc(x0)
    a_th = np.abs(Tfx0) - th
    a_th[a_th < 0] = 0
    a_dir = np.divide(Tfx0, np.abs(Tfx0) + 1e-6)
    return invTfunc(np.multiply(a_dir, a_th))


# ----------------------------------------------------------------- TV

class TV2d_r:
    """2-D finite differences over the first two axes of an n-d array (pics/tvop_class.py:59-129)."""

    ndim = 2

    def grad(self, x):
        sx, sy = x.shape[0], x.shape[1]
        Dx = x[np.r_[1:sx, sx - 1], :] - x
        Dy = x[:, np.r_[1:sy, sy - 1]] - x
        res = np.zeros(x.shape + (self.ndim,), dtype=x.dtype)
        res[..., 0] = Dx
        res[..., 1] = Dy
        return res

    def adjDy(self, x):
        res = x[:, np.r_[0, 0:x.shape[1] - 1]] - x
        res[:, 0] = -x[:, 0]
        res[:, -1] = x[:, -2]
        return res

    def adjDx(self, x):
        res = x[np.r_[0, 0:x.shape[0] - 1], :] - x
        res[0, :] = -x[0, :]
        res[-1, :] = x[-2, :]
        return res

    def Div(self, y):
        return self.adjDx(y[..., 0]) + self.adjDy(y[..., 1])

    def amp(self, grad):
        amp = np.sqrt(np.sum(grad ** 2, axis=(len(grad.shape) - 1)))
        d = np.ones(amp.shape + (self.ndim,), dtype=amp.dtype)
        return np.multiply(amp.reshape(amp.shape + (1,)), d)

    def backward(self, x):
        return self.grad(x)

    def forward(self, y):
        return self.Div(y)


def prox_tv2d_r(y, lambda_tv, step=0.1, niter=40):
    """Chambolle projected dual gradient, 40 iterations (pics/proximal_func.py:114-133).

    ``niter`` exists for the closed-form pin (the reference hard-codes 40)."""
    sizeg = y.shape + (2,)
    G = np.zeros(sizeg)
    tvopt = TV2d_r()
    for _ in range(niter):
        dG = tvopt.grad(tvopt.Div(G) - y / lambda_tv)
        G = G - step * dG
        d = tvopt.amp(G)
        G = G / np.maximum(d, 1.0 * np.ones(sizeg))
    return y - lambda_tv * tvopt.Div(G)


class TV3d:
    """3-D finite differences (pics/tvop_class.py:131-188)."""

    ndim = 3

    def grad(self, x):
        sx, sy, sz = x.shape
        res = np.zeros((sx, sy, sz, 3), dtype=x.dtype)
        res[..., 0] = x[np.r_[1:sx, sx - 1], :, :] - x
        res[..., 1] = x[:, np.r_[1:sy, sy - 1], :] - x
        res[..., 2] = x[:, :, np.r_[1:sz, sz - 1]] - x
        return res

    def adjDx(self, x):
        res = x[np.r_[0, 0:x.shape[0] - 1], :, :] - x
        res[0] = -x[0]
        res[-1] = x[-2]
        return res

    def adjDy(self, x):
        res = x[:, np.r_[0, 0:x.shape[1] - 1], :] - x
        res[:, 0] = -x[:, 0]
        res[:, -1] = x[:, -2]
        return res

    def adjDz(self, x):
        res = x[:, :, np.r_[0, 0:x.shape[2] - 1]] - x
        res[:, :, 0] = -x[:, :, 0]
        res[:, :, -1] = x[:, :, -2]
        return res

    def Div(self, y):
        return self.adjDx(y[..., 0]) + self.adjDy(y[..., 1]) + self.adjDz(y[..., 2])

    def amp(self, grad):
        amp = np.sqrt(np.sum(grad ** 2, axis=3))
        return np.tile(amp[:, :, :, np.newaxis], (1, 1, 1, 3))

    def backward(self, x):
        return self.grad(x)

    def forward(self, y):
        return self.Div(y)


def prox_tv3d(y, lambda_tv, step=0.1, niter=40):
    """(pics/proximal_func.py:136-155), exactly 3-D input."""
    sizeg = y.shape + (y.ndim,)
    G = np.zeros(sizeg)
    tvopt = TV3d()
    for _ in range(niter):
        dG = tvopt.grad(tvopt.Div(G) - y / lambda_tv)
        G = G - step * dG
        d = tvopt.amp(G)
        G = G / np.maximum(d, 1.0 * np.ones(sizeg))
    return y - lambda_tv * tvopt.Div(G)


# ----------------------------------------------------------------- ADMM, IST, FISTA

def ADMM_l2Afxnb_tvx(Afunc, invAfunc, b, Nite, step, tv_r, rho, cgd_Nite=3, tvndim=2):
    """ADMM for ||A x - b||^2 + tv_r TV(x) (pics/CS_MRI_solvers_func.py:336-355)."""
    tvprox = {2: prox_tv2d_r, 3: prox_tv3d}[tvndim]
    z = invAfunc(b)
    u = np.zeros(z.shape)
    x = None
    for _ in range(Nite):
        x = prox_l2_Afxnb_CGD(Afunc, invAfunc, b, z - u, rho, cgd_Nite)
        z = tvprox(x + u, 2.0 * tv_r / rho)
        u = u + step * (x - z)
    return x


def IST_2(Afunc, invAfunc, b, Nite, step, th):
    """(pics/CS_MRI_solvers_func.py:30-40)"""
    x_acc = step * invAfunc(b)
    x = None
    for _ in range(Nite):
        x = prox_l1_soft_thresh2(x_acc, th)
        r = Afunc(x) - b
        x_acc = x_acc - step * invAfunc(r)
    return x


def IST_22(Afunc, invAAfunc, invAfunc, b, Nite, step, th):
    """(pics/CS_MRI_solvers_func.py:42-53)"""
    invAb = invAfunc(b)
    x_acc = step * invAb
    x = None
    for _ in range(Nite):
        x = prox_l1_soft_thresh2(x_acc, th)
        r = invAAfunc(x) - invAb
        x_acc = x_acc - step * r
    return x


def IST_3(Afunc, invAfunc, Tfunc, invTfunc, b, Nite, step, th):
    """(pics/CS_MRI_solvers_func.py:63-73)"""
    x_acc = step * invAfunc(b)
    x = None
    for _ in range(Nite):
        x = prox_l1_Tf_soft_thresh2(Tfunc, invTfunc, x_acc, th)
        r = Afunc(x) - b
        x_acc = x_acc - step * invAfunc(r)
    return x


def IST_32(Afunc, invAAfunc, invAfunc, Tfunc, invTfunc, b, Nite, step, th):
    """(pics/CS_MRI_solvers_func.py:74-86)"""
    invAb = invAfunc(b)
    x_acc = step * invAb
    x = None
    for _ in range(Nite):
        x = prox_l1_Tf_soft_thresh2(Tfunc, invTfunc, x_acc, th)
        r = invAAfunc(x) - invAb
        x_acc = x_acc - step * r
    return x


def IST_wrap(Aopt, Topt=None, b=None, Nite=None, step=None, th=None):
    """Dispatch on ``"forward_backward" in dir(Aopt)`` (pics/CS_MRI_solvers_func.py:89-110)."""
    if b is None:
        return None
    Nite = 20 if Nite is None else Nite
    step = 0.1 if step is None else step
    th = 0.04 if th is None else th
    if Topt is not None:
        if "forward_backward" in dir(Aopt):
            return IST_32(Aopt.forward, Aopt.forward_backward, Aopt.backward, Topt.backward, Topt.forward, b,
                          Nite, step, th)
        return IST_3(Aopt.forward, Aopt.backward, Topt.backward, Topt.forward, b, Nite, step, th)
    if "forward_backward" in dir(Aopt):
        return IST_22(Aopt.forward, Aopt.forward_backward, Aopt.backward, b, Nite, step, th)
    return IST_2(Aopt.forward, Aopt.backward, b, Nite, step, th)


def _fista(grad_step, prox, y0, Nite):
    # shared by the FIST_* variants: y_acc accumulates the gradient steps taken at y
    y = y0
    y_acc = np.zeros(y.shape, dtype=np.complex128)
    x_pre = y
    t = np.ones(Nite + 1)
    x = None
    for k in range(Nite):
        y_acc = y_acc - grad_step(y)
        x = prox(y_acc)
        t[k + 1] = (1 + (1 + 4 * t[k] ** 2) ** 0.5) / 2
        y = x + np.multiply((t[k] - 1) / (t[k + 1]), (x - x_pre))
        x_pre = x
    return x


def FIST_2(Afunc, invAfunc, b, Nite, step, th):
    """(pics/CS_MRI_solvers_func.py:151-167)"""
    return _fista(lambda y: step * invAfunc(Afunc(y) - b), lambda v: prox_l1_soft_thresh2(v, th),
                  step * invAfunc(b), Nite)


def FIST_22(Afunc, invAAfunc, invAfunc, b, Nite, step, th):
    """(pics/CS_MRI_solvers_func.py:169-186)"""
    invAb = invAfunc(b)
    return _fista(lambda y: step * (invAAfunc(y) - invAb), lambda v: prox_l1_soft_thresh2(v, th),
                  step * invAb, Nite)


def FIST_3(Afunc, invAfunc, Tfunc, invTfunc, b, Nite, step, th):
    """(pics/CS_MRI_solvers_func.py:189-205)"""
    return _fista(lambda y: step * invAfunc(Afunc(y) - b),
                  lambda v: prox_l1_Tf_soft_thresh2(Tfunc, invTfunc, v, th), step * invAfunc(b), Nite)


def FIST_32(Afunc, invAAfunc, invAfunc, Tfunc, invTfunc, b, Nite, step, th):
    """(pics/CS_MRI_solvers_func.py:207-225)"""
    invAb = invAfunc(b)
    return _fista(lambda y: step * (invAAfunc(y) - invAb),
                  lambda v: prox_l1_Tf_soft_thresh2(Tfunc, invTfunc, v, th), step * invAb, Nite)


def FIST_wrap(Aopt, Topt=None, b=None, Nite=None, step=None, th=None):
    """(pics/CS_MRI_solvers_func.py:227-247)"""
    if b is None:
        return None
    Nite = 20 if Nite is None else Nite
    step = 0.1 if step is None else step
    th = 0.04 if th is None else th
    if Topt is not None:
        if "forward_backward" in dir(Aopt):
            return FIST_32(Aopt.forward, Aopt.forward_backward, Aopt.backward, Topt.backward, Topt.forward, b,
                           Nite, step, th)
        return FIST_3(Aopt.forward, Aopt.backward, Topt.backward, Topt.forward, b, Nite, step, th)
    if "forward_backward" in dir(Aopt):
        return FIST_22(Aopt.forward, Aopt.forward_backward, Aopt.backward, b, Nite, step, th)
    return FIST_2(Aopt.forward, Aopt.backward, b, Nite, step, th)


def optscaling(FT, b):
    """max |FT.backward(b)| (utilities/utilities_func.py:381-383)."""
    return max(np.absolute(FT.backward(b)).flatten())


# ----------------------------------------------------------------- operators (FP64)

class NUFFT2d:
    """FP64 multi-coil 2-D NUFFT operator with the reference's methods
    (pics/operators_class.py:302-364): forward = type 2, backward = type 1 of
    dcf * kdata (1/M included), forward_backward = type 21 with dcf.  ``threads``
    splits the coil axis over threads (each coil is computed independently, so the
    result does not depend on it)."""

    def __init__(self, im_shape, dcf, x, y, p: O.Params, threads=1):
        self.im_shape = tuple(im_shape)
        self.coords = [np.asarray(x, np.float64), np.asarray(y, np.float64)]
        self.dcf = None if dcf is None else np.asarray(dcf, np.float64).reshape(-1)
        self.p = p
        self.threads = threads

    def _per_coil(self, fn, a):
        nc = a.shape[-1] if a.ndim > 1 else 1
        if self.threads <= 1 or nc < 2 or a.ndim < 2:
            return fn(a)
        from concurrent.futures import ThreadPoolExecutor
        parts = np.array_split(np.arange(nc), min(self.threads, nc))
        with ThreadPoolExecutor(len(parts)) as ex:
            outs = list(ex.map(lambda ix: fn(np.ascontiguousarray(a[..., ix])), parts))
        return np.concatenate(outs, axis=-1)

    def forward(self, im):
        return self._per_coil(lambda a: C.type2(self.coords, a, self.p, fft_workers=1), im)

    def backward(self, kdata):
        k = kdata if self.dcf is None else kdata * self.dcf.reshape((-1,) + (1,) * (kdata.ndim - 1))
        return self._per_coil(lambda a: C.type1(self.coords, a, self.p, fft_workers=1), k)

    def forward_backward(self, im):
        return self.backward(self.forward(im))


class espirit:
    """forward(im) = im * S, backward(x) = sum_c conj(S) x keepdims (pics/operators_class.py:437-472)."""

    def __init__(self, sensitivity):
        self.sens = np.asarray(sensitivity)

    def forward(self, im):
        return im.reshape(self.sens.shape[:-1] + (-1,)) * self.sens

    def backward(self, x):
        return np.sum(x * np.conj(self.sens), axis=-1, keepdims=True)


class joint2operators:
    """B(A(x)) (pics/operators_class.py:496-514)."""

    def __init__(self, Aopt, Bopt):
        self.Aopt, self.Bopt = Aopt, Bopt

    def forward(self, x):
        return self.Bopt.forward(self.Aopt.forward(x))

    def backward(self, x):
        return self.Aopt.backward(self.Bopt.backward(x))

    def forward_backward(self, x):
        return self.Aopt.backward(self.Bopt.forward_backward(self.Aopt.forward(x)))
