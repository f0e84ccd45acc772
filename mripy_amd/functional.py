"""Reference-compatible NUFFT functions running on the HIP path.

Same names, arguments and conventions as mripy's
``nufft{2,3}d{1,2,21}_gaussker`` (fft/nufft_func.py:702,741,780,1167,1209,1261)
and their ``_cuda`` variants (fft/nufft_func_cuda.py:581,609,639,1463,1512,1562):

* coordinates ``x, y[, z]`` in radians (any shape, flattened), scaled by ``df``;
* ``c``/``Fk`` carry trailing coil dimensions, flattened into coils;
* parameters from ``eps`` by the reference rule (3-D clamps nspread to [3, 6]),
  or explicitly with ``rat=``/``nspread=`` keywords;
* type 1 returns ``(1/M) sum_j c_j e^{+-ik.x_j}`` on ``(ms, mt[, mu], ...)``,
  type 2 returns ``sum_k F_k e^{-+ik.x_j}``, type 21 is type 1 of (dcf *) type 2.

``gridfast`` and ``save_ram`` are accepted and ignored: the result does not
depend on them.  numpy inputs give numpy outputs (synchronously, in the input's
complex dtype); torch device tensors give complex64 device tensors.
"""
from __future__ import annotations

import numpy as np
import torch

from .operators import _fingerprint
from .plan import Plan

_plan_cache = {}


def _get_plan(modes, ncoils, eps, iflag, rat, nspread, device):
    key = (tuple(modes), int(ncoils), eps, int(iflag), rat, nspread, str(device))
    P = _plan_cache.get(key)
    if P is None:
        if len(_plan_cache) > 16:
            _plan_cache.clear()
        P = Plan(modes, ncoils=ncoils, eps=eps if (rat is None or nspread is None) else None, rat=rat,
                 nspread=nspread, iflag=iflag, device=device)
        _plan_cache[key] = P
    return P


def _device_of(*arrays):
    for a in arrays:
        if torch.is_tensor(a) and a.is_cuda:
            return a.device
    return torch.device("cuda", torch.cuda.current_device())


def _finish(out, like, shape):
    out = out.reshape(shape)
    if torch.is_tensor(like):
        return out
    o = out.cpu().numpy()
    dt = np.asarray(like).dtype
    if np.iscomplexobj(np.zeros(0, dt)) and dt != o.dtype:
        o = o.astype(dt)
    elif not np.iscomplexobj(np.zeros(0, dt)):
        o = o.astype(np.complex128)
    return o


def _run(kind, coords, data, modes, df, eps, iflag, rat, nspread, dcf=None):
    d = len(modes)
    dev = _device_of(data, *coords)
    M = int(np.prod(np.shape(coords[0])))
    if kind == 1:
        lead = (M,)
        trail = tuple(np.shape(data))[1:] if np.ndim(data) > 1 else ()
    else:
        lead = tuple(modes)
        trail = tuple(np.shape(data))[d:]
    nc = int(np.prod(trail)) if trail else 1
    P = _get_plan(modes, nc, eps, iflag, rat, nspread, dev)
    # re-bin only when the coordinates, df or weights differ from the plan's last binning
    key = tuple(_fingerprint(c) for c in coords) + (float(df), _fingerprint(dcf))
    if getattr(P, "_fn_key", None) != key:
        P.set_points(list(coords), df=df, dcf=dcf)
        P._fn_key = key
    t = torch.as_tensor(data)
    if kind == 1:
        out = P.type1(t.reshape(M, nc))
        return _finish(out, data, tuple(modes) + trail)
    if kind == 2:
        out = P.type2(t.reshape(tuple(modes) + (nc,)))
        return _finish(out, data, (M,) + trail)
    out = P.type21(t.reshape(tuple(modes) + (nc,)))
    return _finish(out, data, tuple(modes) + trail)


def nufft2d1_gaussker(x, y, c, ms, mt, df=1.0, eps=1e-15, iflag=1, gridfast=1, save_ram=1, rat=None, nspread=None):
    """2-D type 1 (adjoint), fft/nufft_func.py:702-738."""
    return _run(1, [x, y], c, (ms, mt), df, eps, iflag, rat, nspread)


def nufft2d2_gaussker(x, y, Fk, ms, mt, df=1.0, eps=1e-15, iflag=1, gridfast=1, save_ram=1, rat=None, nspread=None):
    """2-D type 2 (forward), fft/nufft_func.py:741-777."""
    return _run(2, [x, y], Fk, (ms, mt), df, eps, iflag, rat, nspread)


def nufft2d21_gaussker(x, y, Fk, ms, mt, df=1.0, eps=1e-15, iflag=1, gridfast=1, save_ram=1, rat=None,
                       nspread=None):
    """2-D type 1 of type 2 (no density weighting), fft/nufft_func.py:780-830."""
    return _run(21, [x, y], Fk, (ms, mt), df, eps, iflag, rat, nspread)


def _three(eps, rat, nspread):
    # every 3-D reference driver clamps nspread with max_nspread=6 (fft/nufft_func.py:1169)
    if rat is None or nspread is None:
        from .plan import grid_params
        r, n, _, _ = grid_params((1, 1, 1), eps, 6)
        return (r if rat is None else rat), (n if nspread is None else nspread)
    return rat, nspread


def nufft3d1_gaussker(x, y, z, c, ms, mt, mu, df=1.0, eps=1e-15, iflag=1, gridfast=1, save_ram=1, rat=None,
                      nspread=None):
    """3-D type 1 (adjoint), fft/nufft_func.py:1167-1206."""
    rat, nspread = _three(eps, rat, nspread)
    return _run(1, [x, y, z], c, (ms, mt, mu), df, eps, iflag, rat, nspread)


def nufft3d2_gaussker(x, y, z, Fk, ms, mt, mu, df=1.0, eps=1e-15, iflag=1, gridfast=0, save_ram=0, rat=None,
                      nspread=None):
    """3-D type 2 (forward), fft/nufft_func.py:1209-1258."""
    rat, nspread = _three(eps, rat, nspread)
    return _run(2, [x, y, z], Fk, (ms, mt, mu), df, eps, iflag, rat, nspread)


def nufft3d21_gaussker(x, y, z, Fk, ms, mt, mu, dcf=None, df=1.0, eps=1e-15, iflag=1, gridfast=0, save_ram=1,
                       rat=None, nspread=None):
    """3-D type 1 of dcf * type 2, fft/nufft_func.py:1261-1330."""
    rat, nspread = _three(eps, rat, nspread)
    return _run(21, [x, y, z], Fk, (ms, mt, mu), df, eps, iflag, rat, nspread, dcf=dcf)


# the GPU driver names of fft/nufft_func_cuda.py
nufft2d1_gaussker_cuda = nufft2d1_gaussker
nufft2d2_gaussker_cuda = nufft2d2_gaussker
nufft2d21_gaussker_cuda = nufft2d21_gaussker
nufft3d1_gaussker_cuda = nufft3d1_gaussker
nufft3d2_gaussker_cuda = nufft3d2_gaussker
nufft3d21_gaussker_cuda = nufft3d21_gaussker
