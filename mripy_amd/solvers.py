"""Iterative reconstructions that drive the NUFFT operator, on the GPU, with mripy's
names, arguments and iteration logic.

* ``BacktrackingLineSearch``              pics/opt_alg.py:8-33
* ``prox_l2_Afxnb_CGD`` / ``..._CGD2``      pics/proximal_func.py:282-343
* ``prox_l1_soft_thresh(2)``, ``prox_l1_Tf_soft_thresh2``  pics/proximal_func.py:12-52
* ``TV2d_r``, ``prox_tv2d_r``              pics/tvop_class.py:59-129, pics/proximal_func.py:114-133
* ``TV3d``, ``prox_tv3d``                  pics/tvop_class.py:131-253, pics/proximal_func.py:136-177
* ``ADMM_l2Afxnb_tvx``                    pics/CS_MRI_solvers_func.py:336-355
* ``IST_2/22/3/32``, ``IST_wrap``         pics/CS_MRI_solvers_func.py:30-110
* ``FIST_2/22/3/32``, ``FIST_wrap``       pics/CS_MRI_solvers_func.py:151-247
* ``optscaling``                          utilities/utilities_func.py:381-383

Vectors live on the GPU as torch tensors (complex64 images and k-space, as the NUFFT
computes in).  Every inner product and norm the iteration logic branches on is a
deterministic FP64 reduction on the device (``gn_dot_f64`` / ``gn_sqdist_f64``), and
the TV prox runs in FP64 in two HIP kernels per Chambolle step (``gn_tv_prox``).
Operator callables receive and return device tensors (the operator classes in
``mripy_amd.operators`` do); a numpy ``b`` is moved to the device and the result is
returned as numpy.

Differences from the reference that do not change results: the conjugate-gradient
line search reuses the gradient and cost the caller already evaluated at the same
point (the reference recomputes ``df(x)`` and ``f(x)`` inside
``BacktrackingLineSearch``; the operators are deterministic, so the values are
identical), and the per-iteration ``print`` calls are behind ``verbose``.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib

# ------------------------------------------------------------------ FP64 reductions

_DTYPES = {torch.float32: _lib.GN_F32, torch.complex64: _lib.GN_C64, torch.complex128: _lib.GN_C128}
_work = {}


def _workspace(device):
    key = (device.type, device.index)
    w = _work.get(key)
    if w is None:
        nbytes = int(_lib.load().gn_reduce_work_bytes())
        w = (torch.empty(nbytes, dtype=torch.uint8, device=device), torch.zeros(2, dtype=torch.float64,
                                                                                  device=device))
        _work[key] = w
    return w


def _code(t):
    try:
        return _DTYPES[t.dtype]
    except KeyError:
        raise TypeError("FP64 reductions take float32, complex64 or complex128 tensors, not %s" % t.dtype)


def _flat(t):
    if not (torch.is_tensor(t) and t.is_cuda):
        raise TypeError("expected a CUDA tensor")
    return t.contiguous().reshape(-1)


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def dot_f64(a, b):
    """sum conj(a) * b accumulated in FP64 on the device; returns a Python complex."""
    a, b = _flat(a), _flat(b)
    if a.dtype != b.dtype or a.numel() != b.numel():
        raise ValueError("dot_f64: operands differ in dtype or size")
    work, out = _workspace(a.device)
    _lib.check(_lib.load().gn_dot_f64(_code(a), a.numel(), a.data_ptr(), b.data_ptr(), work.data_ptr(),
                                      out.data_ptr(), _stream(a.device)))
    re, im = out.tolist()
    return complex(re, im)


def sqdist_f64(a, b=None):
    """sum |a - b|^2 (or sum |a|^2) accumulated in FP64 on the device; returns a float."""
    a = _flat(a)
    if b is not None:
        b = _flat(b)
        if a.dtype != b.dtype or a.numel() != b.numel():
            raise ValueError("sqdist_f64: operands differ in dtype or size")
    work, out = _workspace(a.device)
    _lib.check(_lib.load().gn_sqdist_f64(_code(a), a.numel(), a.data_ptr(), None if b is None else b.data_ptr(),
                                         work.data_ptr(), out.data_ptr(), _stream(a.device)))
    return float(out[0].item())


def norm_f64(a):
    """Euclidean norm with an FP64 accumulation (np.linalg.norm of the flattened array)."""
    return math.sqrt(sqdist_f64(a))


def _dev(x, like=None, keep_dtype=False):
    """Device tensor of x; numpy double data become single precision (the NUFFT's)
    unless keep_dtype."""
    if torch.is_tensor(x):
        return x if x.is_cuda else x.cuda()
    dev = like.device if torch.is_tensor(like) else torch.device("cuda", torch.cuda.current_device())
    a = np.asarray(x)
    if not keep_dtype and a.dtype == np.complex128:
        a = a.astype(np.complex64)
    elif not keep_dtype and a.dtype == np.float64:
        a = a.astype(np.float32)
    return torch.as_tensor(a).to(dev)


def _back(x, like):
    if torch.is_tensor(like):
        return x
    o = x.cpu().numpy()
    return o.astype(np.asarray(like).dtype) if np.iscomplexobj(o) == np.iscomplexobj(like) else o


# ------------------------------------------------------------------ line search and CG

def BacktrackingLineSearch(f, df, x, p, c=0.0001, rho=0.2, ls_Nite=10):
    """Armijo backtracking (pics/opt_alg.py:8-33): alpha = rho^i, i < ls_Nite, until
    f(x + alpha p) - f(x) <= c alpha Re<p, df(x)> or f(x + alpha p) <= f(x).
    ``f`` returns a float; ``df`` a device tensor.  Returns (alpha, i)."""
    derphi = dot_f64(df(x), p).real
    return _backtrack(f, x, p, f(x), derphi, c, rho, ls_Nite)


def _backtrack(f, x, p, f0, derphi, c=0.0001, rho=0.2, ls_Nite=10):
    alphak = 1.0
    f_try = f(x + alphak * p)
    i = 0
    while i < ls_Nite and f_try - f0 > c * alphak * derphi and f_try > f0:
        alphak = alphak * rho
        f_try = f(x + alphak * p)
        i += 1
    return alphak, i


def _nlcg(Afunc, invAfunc, b, x, x0, rho, Nite, ls_Nite):
    """Nonlinear CG with the reference's un-squared Fletcher-Reeves beta
    (pics/proximal_func.py:282-343); x0 is None for the plain least-squares cost."""
    eps = 0.001

    def cost(xi):
        v = sqdist_f64(Afunc(xi), b)
        if x0 is not None:
            v += (rho / 2) * sqdist_f64(xi, x0)
        return v

    def cost_and_grad(xi):
        r = Afunc(xi) - b
        v = sqdist_f64(r)
        g = 2 * invAfunc(r)
        if x0 is not None:
            v += (rho / 2) * sqdist_f64(xi, x0)
            g = g + rho * (xi - x0)
        return v, g

    def search(xk, p, g, fk):
        return _backtrack(cost, xk, p, fk, dot_f64(g, p).real, ls_Nite=ls_Nite)

    fx, g = cost_and_grad(x)
    dx = -g
    alpha, nstp = search(x, dx, g, fx)
    x = x + alpha * dx
    s = dx
    delta0 = norm_f64(dx)
    deltanew = delta0
    i = 0
    while i < Nite and deltanew > eps * delta0 and nstp < ls_Nite:
        fx, g = cost_and_grad(x)
        dx = -g
        deltaold = deltanew
        deltanew = norm_f64(dx)
        beta = float(deltanew / float(deltaold))
        s = dx + beta * s
        alpha, nstp = search(x, s, g, fx)
        x = x + alpha * s
        i = i + 1
    return x


def prox_l2_Afxnb_CGD(Afunc, invAfunc, b, x0, rho, Nite, ls_Nite=10):
    """argmin ||A x - b||^2 + (rho/2)||x - x0||^2 by nonlinear CG from x0
    (pics/proximal_func.py:282-311)."""
    bd = _dev(b)
    return _back(_nlcg(Afunc, invAfunc, bd, _dev(x0, bd), _dev(x0, bd), rho, Nite, ls_Nite), b)


def prox_l2_Afxnb_CGD2(Afunc, invAfunc, b, Nite, ls_Nite=10):
    """argmin ||A x - b||^2 by nonlinear CG from x = invA(b) (pics/proximal_func.py:314-343)."""
    bd = _dev(b)
    return _back(_nlcg(Afunc, invAfunc, bd, invAfunc(bd), None, 0.0, Nite, ls_Nite), b)


# ------------------------------------------------------------------ l1 proxes

def prox_l1_soft_thresh(x0, th):
    """exp(i angle(x)) max(|x| - th, 0) (pics/proximal_func.py:12-16)."""
    a = (x0.abs() - th).clamp_min(0)
    return torch.polar(a, x0.angle()) if x0.is_complex() else a * torch.sign(x0)


def prox_l1_soft_thresh2(x0, th):
    """x/(|x| + 1e-6) max(|x| - th, 0) (pics/proximal_func.py:19-23)."""
    m = x0.abs()
    return x0 / (m + 1e-6) * (m - th).clamp_min(0)


def prox_l1_Tf_soft_thresh2(Tfunc, invTfunc, x0, th):
    """invT(soft2(T(x))) (pics/proximal_func.py:44-52)."""
    return invTfunc(prox_l1_soft_thresh2(Tfunc(x0), th))


# ------------------------------------------------------------------ TV

class TV2d_r:
    """2-D finite differences over the first two axes of a device tensor
    (pics/tvop_class.py:59-129): grad = forward differences with a zero last
    row/column, Div = its adjoint, amp = sqrt(sum grad**2) over the last axis."""

    ndim = 2

    def grad(self, x):
        Dx = torch.cat([x[1:], x[-1:]], 0) - x
        Dy = torch.cat([x[:, 1:], x[:, -1:]], 1) - x
        return torch.stack([Dx, Dy], -1)

    def adjDx(self, x):
        res = torch.cat([x[:1], x[:-1]], 0) - x
        res[0] = -x[0]
        res[-1] = x[-2]
        return res

    def adjDy(self, x):
        res = torch.cat([x[:, :1], x[:, :-1]], 1) - x
        res[:, 0] = -x[:, 0]
        res[:, -1] = x[:, -2]
        return res

    def Div(self, y):
        return self.adjDx(y[..., 0]) + self.adjDy(y[..., 1])

    def amp(self, grad):
        a = torch.sqrt((grad ** 2).sum(-1, keepdim=True))
        return a.expand(grad.shape).contiguous()

    def backward(self, x):
        return self.grad(x)

    def forward(self, y):
        return self.Div(y)


class TV3d(TV2d_r):
    """3-D finite differences over the first three axes (pics/tvop_class.py:131-188,
    and TV3d_r :191-253 for trailing dims)."""

    ndim = 3

    def grad(self, x):
        Dx = torch.cat([x[1:], x[-1:]], 0) - x
        Dy = torch.cat([x[:, 1:], x[:, -1:]], 1) - x
        Dz = torch.cat([x[:, :, 1:], x[:, :, -1:]], 2) - x
        return torch.stack([Dx, Dy, Dz], -1)

    def adjDz(self, x):
        res = torch.cat([x[:, :, :1], x[:, :, :-1]], 2) - x
        res[:, :, 0] = -x[:, :, 0]
        res[:, :, -1] = x[:, :, -2]
        return res

    def Div(self, y):
        return self.adjDx(y[..., 0]) + self.adjDy(y[..., 1]) + self.adjDz(y[..., 2])


TV3d_r = TV3d

_TV_AMP = {"reference": _lib.GN_TV_AMP_SQUARE, "modulus": _lib.GN_TV_AMP_MODULUS}


def _tv_prox(y, lambda_tv, step, niter, amp, ndim):
    yd = _dev(y, keep_dtype=True)
    if yd.dtype not in (torch.complex64, torch.complex128):
        yd = yd.to(torch.complex64)
    if yd.dim() < ndim:
        raise ValueError("TV prox over %d axes needs at least %d dimensions" % (ndim, ndim))
    yd = yd.contiguous()
    shape = (ctypes.c_int64 * 3)(*(list(yd.shape[:ndim]) + [1] * (3 - ndim)))
    nt = int(np.prod(yd.shape[ndim:])) if yd.dim() > ndim else 1
    L = _lib.load()
    work = torch.empty(int(L.gn_tv_work_bytes(ndim, shape, nt)), dtype=torch.uint8, device=yd.device)
    f = torch.empty_like(yd)
    _lib.check(L.gn_tv_prox(_code(yd), ndim, shape, nt, yd.data_ptr(), float(lambda_tv), float(step), int(niter),
                            _TV_AMP[amp], work.data_ptr(), f.data_ptr(), _stream(yd.device)))
    return _back(f, y)


def prox_tv2d_r(y, lambda_tv, step=0.1, niter=40, amp="reference"):
    """prox of lambda_tv * TV over the first two axes (pics/proximal_func.py:114-133):
    ``niter`` Chambolle projected dual gradient steps in FP64 on the device,
    f = y - lambda_tv * Div(G).  ``amp="reference"`` normalizes G by
    max(sqrt(Gx**2 + Gy**2), 1) with complex squares and numpy's lexicographic
    complex maximum, exactly as the reference; ``"modulus"`` uses |G|."""
    return _tv_prox(y, lambda_tv, step, niter, amp, 2)


def prox_tv3d(y, lambda_tv, step=0.1, niter=40, amp="reference"):
    """prox of lambda_tv * TV over the first three axes (pics/proximal_func.py:136-155,
    TV3d).  Trailing dims are treated as independent images (the reference's TV3d
    takes exactly three dims)."""
    return _tv_prox(y, lambda_tv, step, niter, amp, 3)


prox_tv3d_r = prox_tv3d


# ------------------------------------------------------------------ ADMM, IST, FISTA

def ADMM_l2Afxnb_tvx(Afunc, invAfunc, b, Nite, step, tv_r, rho, cgd_Nite=3, tvndim=2, verbose=False):
    """ADMM for ||A x - b||^2 + tv_r TV(x) (pics/CS_MRI_solvers_func.py:336-355): x-step =
    cgd_Nite proximal CG iterations, z-step = TV prox at 2 tv_r/rho, u += step (x - z)."""
    if tvndim == 2:
        tvprox = prox_tv2d_r
    elif tvndim == 3:
        tvprox = prox_tv3d
    else:
        raise ValueError("dimension incompatible in ADMM_l2Afxnb_tvx: tvndim must be 2 or 3")
    bd = _dev(b)
    z = invAfunc(bd)
    u = torch.zeros_like(z)
    x = None
    for _ in range(Nite):
        x = _nlcg(Afunc, invAfunc, bd, z - u, z - u, rho, cgd_Nite, 10)
        z = tvprox(x + u, 2.0 * tv_r / rho)
        u = u + step * (x - z)
        if verbose:
            print("gradient in ADMM %g" % norm_f64(x - z))
    return _back(x, b)


def _ist(grad_step, prox, x_acc, Nite, verbose):
    x = None
    for _ in range(Nite):
        x = prox(x_acc)
        r, x_acc = grad_step(x, x_acc)
        if verbose:
            print(norm_f64(r))
    return x


def IST_2(Afunc, invAfunc, b, Nite, step, th, verbose=False):
    """(pics/CS_MRI_solvers_func.py:30-40)"""
    bd = _dev(b)

    def gs(x, x_acc):
        r = Afunc(x) - bd
        return r, x_acc - step * invAfunc(r)
    return _back(_ist(gs, lambda v: prox_l1_soft_thresh2(v, th), step * invAfunc(bd), Nite, verbose), b)


def IST_22(Afunc, invAAfunc, invAfunc, b, Nite, step, th, verbose=False):
    """(pics/CS_MRI_solvers_func.py:42-53)"""
    bd = _dev(b)
    invAb = invAfunc(bd)

    def gs(x, x_acc):
        r = invAAfunc(x) - invAb
        return r, x_acc - step * r
    return _back(_ist(gs, lambda v: prox_l1_soft_thresh2(v, th), step * invAb, Nite, verbose), b)


def IST_3(Afunc, invAfunc, Tfunc, invTfunc, b, Nite, step, th, verbose=False):
    """(pics/CS_MRI_solvers_func.py:63-73)"""
    bd = _dev(b)

    def gs(x, x_acc):
        r = Afunc(x) - bd
        return r, x_acc - step * invAfunc(r)
    return _back(_ist(gs, lambda v: prox_l1_Tf_soft_thresh2(Tfunc, invTfunc, v, th), step * invAfunc(bd), Nite,
                      verbose), b)


def IST_32(Afunc, invAAfunc, invAfunc, Tfunc, invTfunc, b, Nite, step, th, verbose=False):
    """(pics/CS_MRI_solvers_func.py:75-86)"""
    bd = _dev(b)
    invAb = invAfunc(bd)

    def gs(x, x_acc):
        r = invAAfunc(x) - invAb
        return r, x_acc - step * r
    return _back(_ist(gs, lambda v: prox_l1_Tf_soft_thresh2(Tfunc, invTfunc, v, th), step * invAb, Nite, verbose),
                 b)


def _defaults(Nite, step, th):
    return (20 if Nite is None else Nite, 0.1 if step is None else step, 0.04 if th is None else th)


def IST_wrap(Aopt, Topt=None, b=None, Nite=None, step=None, th=None, verbose=False):
    """Dispatch on ``"forward_backward" in dir(Aopt)`` (pics/CS_MRI_solvers_func.py:89-110)."""
    if b is None:
        return None
    Nite, step, th = _defaults(Nite, step, th)
    fb = "forward_backward" in dir(Aopt)
    if Topt is not None:
        if fb:
            return IST_32(Aopt.forward, Aopt.forward_backward, Aopt.backward, Topt.backward, Topt.forward, b, Nite,
                          step, th, verbose)
        return IST_3(Aopt.forward, Aopt.backward, Topt.backward, Topt.forward, b, Nite, step, th, verbose)
    if fb:
        return IST_22(Aopt.forward, Aopt.forward_backward, Aopt.backward, b, Nite, step, th, verbose)
    return IST_2(Aopt.forward, Aopt.backward, b, Nite, step, th, verbose)


def _fista(grad_at, prox, y0, Nite, verbose):
    # y_acc accumulates the gradient steps taken at the extrapolated point y
    y = y0
    y_acc = torch.zeros_like(y0)
    x_pre = y
    t = np.ones(Nite + 1)
    x = None
    for k in range(Nite):
        r, gstep = grad_at(y)
        y_acc = y_acc - gstep
        x = prox(y_acc)
        t[k + 1] = (1 + (1 + 4 * t[k] ** 2) ** 0.5) / 2
        y = x + float((t[k] - 1) / (t[k + 1])) * (x - x_pre)
        x_pre = x
        if verbose:
            print(norm_f64(r))
    return x


def FIST_2(Afunc, invAfunc, b, Nite, step, th, verbose=False):
    """(pics/CS_MRI_solvers_func.py:151-167)"""
    bd = _dev(b)

    def ga(y):
        r = Afunc(y) - bd
        return r, step * invAfunc(r)
    return _back(_fista(ga, lambda v: prox_l1_soft_thresh2(v, th), step * invAfunc(bd), Nite, verbose), b)


def FIST_22(Afunc, invAAfunc, invAfunc, b, Nite, step, th, verbose=False):
    """(pics/CS_MRI_solvers_func.py:169-186)"""
    bd = _dev(b)
    invAb = invAfunc(bd)

    def ga(y):
        r = invAAfunc(y) - invAb
        return r, step * r
    return _back(_fista(ga, lambda v: prox_l1_soft_thresh2(v, th), step * invAb, Nite, verbose), b)


def FIST_3(Afunc, invAfunc, Tfunc, invTfunc, b, Nite, step, th, verbose=False):
    """(pics/CS_MRI_solvers_func.py:189-205)"""
    bd = _dev(b)

    def ga(y):
        r = Afunc(y) - bd
        return r, step * invAfunc(r)
    return _back(_fista(ga, lambda v: prox_l1_Tf_soft_thresh2(Tfunc, invTfunc, v, th), step * invAfunc(bd), Nite,
                        verbose), b)


def FIST_32(Afunc, invAAfunc, invAfunc, Tfunc, invTfunc, b, Nite, step, th, verbose=False):
    """(pics/CS_MRI_solvers_func.py:207-225)"""
    bd = _dev(b)
    invAb = invAfunc(bd)

    def ga(y):
        r = invAAfunc(y) - invAb
        return r, step * r
    return _back(_fista(ga, lambda v: prox_l1_Tf_soft_thresh2(Tfunc, invTfunc, v, th), step * invAb, Nite,
                        verbose), b)


def FIST_wrap(Aopt, Topt=None, b=None, Nite=None, step=None, th=None, verbose=False):
    """(pics/CS_MRI_solvers_func.py:227-247)"""
    if b is None:
        return None
    Nite, step, th = _defaults(Nite, step, th)
    fb = "forward_backward" in dir(Aopt)
    if Topt is not None:
        if fb:
            return FIST_32(Aopt.forward, Aopt.forward_backward, Aopt.backward, Topt.backward, Topt.forward, b,
                           Nite, step, th, verbose)
        return FIST_3(Aopt.forward, Aopt.backward, Topt.backward, Topt.forward, b, Nite, step, th, verbose)
    if fb:
        return FIST_22(Aopt.forward, Aopt.forward_backward, Aopt.backward, b, Nite, step, th, verbose)
    return FIST_2(Aopt.forward, Aopt.backward, b, Nite, step, th, verbose)


def optscaling(FT, b):
    """max |FT.backward(b)| (utilities/utilities_func.py:381-383)."""
    v = FT.backward(_dev(b) if not torch.is_tensor(b) else b)
    return float(v.abs().max().item()) if torch.is_tensor(v) else float(np.max(np.abs(v)))
