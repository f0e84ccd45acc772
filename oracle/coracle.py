"""ctypes front end of the FP64 C oracle (oracle/gridding.c) -- TEST INFRASTRUCTURE ONLY.

Provides the same ``spread``/``gather`` contract as ``oracle.nufft_oracle`` but
runs the reference's fast loop nests in C with OpenMP, so that full-size
configurations (e.g. 3-D 128^3 x 8 coils x 1M samples) finish in seconds.  The
drivers ``type1``/``type2`` use ``scipy.fft`` with ``workers = cores - 1``, the
analogue of the reference's ``fftw_func.*(threads=config.cpu_count-1)``
(fft/nufft_func.py:1192,1240).  Imported only by tests, smoke() and the
cpu_baseline leg of bench.py.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from . import nufft_oracle as O

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.c_int64
_int = ctypes.c_int


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE, "-B" if force else "all"])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_spread2d.argtypes = [_i64, _dp, _dp, _dp, _int, _int, _int, _int, ctypes.c_double, _dp]
        L.orc_gather2d.argtypes = [_i64, _dp, _dp, _dp, _int, _int, _int, _int, ctypes.c_double, _dp]
        L.orc_spread3d.argtypes = [_i64, _dp, _dp, _dp, _dp, _int, _int, _int, _int, _int,
                                   ctypes.c_double, _dp, _int]
        L.orc_gather3d.argtypes = [_i64, _dp, _dp, _dp, _dp, _int, _int, _int, _int, _int,
                                   ctypes.c_double, _dp, _int]
        L.orc_interp_spread3d.argtypes = [_i64, _dp, _dp, _dp, _dp, _dp, _int, _int, _int, _int, _int,
                                          ctypes.c_double, _dp, _int]
        L.orc_cell_index.argtypes = [_i64, _dp, _int, ctypes.POINTER(ctypes.c_int64), _dp]
        L.orc_max_threads.restype = _int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def cores():
    return len(os.sched_getaffinity(0))


def cell_index(x, nf):
    x = _f64(x)
    m = np.zeros(x.shape, np.int64)
    d = np.zeros(x.shape, np.float64)
    lib().orc_cell_index(x.size, _p(x), nf, m.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _p(d))
    return m, d


def spread(coords, c, nf, ns, tau, threads=None):
    """Type-1 gridding, coil-last (*nf, nc) complex128."""
    coords = [_f64(x) for x in coords]
    M = coords[0].shape[0]
    c2 = _c128(np.asarray(c).reshape(M, -1))
    nc = c2.shape[1]
    grid = np.zeros(tuple(nf) + (nc,), np.complex128)
    threads = threads or cores()
    if len(nf) == 2:
        lib().orc_spread2d(M, _p(coords[0]), _p(coords[1]), _p(c2.view(np.float64)), nc,
                           nf[0], nf[1], ns, tau, _p(grid.view(np.float64)))
    else:
        lib().orc_spread3d(M, _p(coords[0]), _p(coords[1]), _p(coords[2]), _p(c2.view(np.float64)), nc,
                           nf[0], nf[1], nf[2], ns, tau, _p(grid.view(np.float64)), threads)
    return grid


def gather(coords, grid, nf, ns, tau, threads=None):
    """Type-2 gridding (includes the 1/G factor), (M, nc) complex128."""
    coords = [_f64(x) for x in coords]
    M = coords[0].shape[0]
    g = _c128(np.asarray(grid).reshape(tuple(nf) + (-1,)))
    nc = g.shape[-1]
    out = np.zeros((M, nc), np.complex128)
    threads = threads or cores()
    if len(nf) == 2:
        lib().orc_gather2d(M, _p(coords[0]), _p(coords[1]), _p(g.view(np.float64)), nc,
                           nf[0], nf[1], ns, tau, _p(out.view(np.float64)))
    else:
        lib().orc_gather3d(M, _p(coords[0]), _p(coords[1]), _p(coords[2]), _p(g.view(np.float64)), nc,
                           nf[0], nf[1], nf[2], ns, tau, _p(out.view(np.float64)), threads)
    return out


def interp_spread3d(coords, grid, dcf, nf, ns, tau, threads=None):
    coords = [_f64(x) for x in coords]
    M = coords[0].shape[0]
    g = _c128(np.asarray(grid).reshape(tuple(nf) + (-1,)))
    nc = g.shape[-1]
    out = np.zeros_like(g)
    d = _f64(dcf) if dcf is not None else None
    threads = threads or cores()
    lib().orc_interp_spread3d(M, _p(coords[0]), _p(coords[1]), _p(coords[2]),
                              _p(d) if d is not None else None, _p(g.view(np.float64)), nc,
                              nf[0], nf[1], nf[2], ns, tau, _p(out.view(np.float64)), threads)
    return out


def _fftn(a, axes, inverse, workers):
    import scipy.fft as sf
    return (sf.ifftn if inverse else sf.fftn)(a, axes=axes, workers=workers)


def type1(coords, c, p: O.Params, threads=None, fft_workers=None):
    """Adjoint NUFFT through the C spread (same contract as nufft_oracle.type1)."""
    coords = [_f64(x) * p.df for x in coords]
    M = coords[0].shape[0]
    c2 = np.asarray(c).reshape(M, -1)
    has_coil = np.asarray(c).ndim > 1
    grid = spread(coords, c2, p.nf, p.ns, p.tau, threads)
    axes = tuple(range(p.d))
    w = fft_workers or max(1, cores() - 1)
    if p.iflag < 0:
        F = _fftn(grid, axes, False, w) / p.G
    else:
        F = _fftn(grid, axes, True, w)
    F = O.crop(F, p.modes)
    out = F * O.deapod(p.modes, p.tau)[..., None] / M
    return out if has_coil else out[..., 0]


def forward_grid(Fk, p: O.Params, fft_workers=None):
    """``nufft_oracle.forward_grid`` with a multi-threaded FFT: the oversampled grid
    of the type-2 map before the gather, coil-last (*nf, nc) complex128."""
    F2 = np.asarray(Fk).reshape(p.modes + (-1,)) * O.deapod(p.modes, p.tau)[..., None]
    grid = O.place(F2, p.nf)
    axes = tuple(range(p.d))
    w = fft_workers or max(1, cores() - 1)
    return _fftn(grid, axes, True, w) * p.G if p.iflag < 0 else _fftn(grid, axes, False, w)


def type2(coords, Fk, p: O.Params, threads=None, fft_workers=None):
    """Forward NUFFT through the C gather (same contract as nufft_oracle.type2)."""
    coords = [_f64(x) * p.df for x in coords]
    Fk = np.asarray(Fk)
    has_coil = Fk.ndim > p.d
    F2 = Fk.reshape(p.modes + (-1,)) * O.deapod(p.modes, p.tau)[..., None]
    grid = O.place(F2, p.nf)
    axes = tuple(range(p.d))
    w = fft_workers or max(1, cores() - 1)
    if p.iflag < 0:
        grid = _fftn(grid, axes, True, w) * p.G
    else:
        grid = _fftn(grid, axes, False, w)
    out = gather(coords, grid, p.nf, p.ns, p.tau, threads)
    return out if has_coil else out[:, 0]
