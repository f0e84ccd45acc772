"""Trajectory and k-space data I/O for the NUFFT path.

* ``loadmat`` reads MATLAB v5/v7 files with scipy.io (a data parser: nothing in the
  file is executed).
* ``load_ute_mat`` restates the data preparation of the reference's 3-D UTE cones
  reconstruction (test/MRI_recon/test_uterecon.py:85-99): ``ktraj`` (3, ...) in
  samples scaled by the image size, ``dcf``, ``kdata`` (..., nc); the image shape is
  inferred as 2 * int(max(ktraj[a])) per axis, ktraj is divided by it, and kdata is
  flattened to (M, nc).  The result feeds ``NUFFT3d_cuda(im_shape, dcf)`` followed by
  ``normalize_set_ktraj(ktraj)`` exactly as the reference script does.
* ``load_spiral_mat`` reads the 2-D spiral dataset of the reference's data folder
  (data/spiral.mat: ``k`` = kx + i ky in cycles per sample, |k| <= 0.5, ``w``
  density weights, ``data`` with coils last) and returns radian coordinates
  (2 pi k), the weights and (M, nc) samples.  The file carries no matrix size;
  ``im_shape`` defaults to 256 x 256.
"""
from __future__ import annotations

import numpy as np


def loadmat(path, keys=None):
    """Variables of a .mat file as numpy arrays (scipy.io.loadmat; MATLAB's own
    header entries dropped).  ``keys`` restricts the result to those names."""
    import scipy.io as sio
    m = sio.loadmat(path)
    out = {k: v for k, v in m.items() if not k.startswith("__")}
    if keys is not None:
        missing = [k for k in keys if k not in out]
        if missing:
            raise KeyError("%s: no variable(s) %s" % (path, ", ".join(missing)))
        out = {k: out[k] for k in keys}
    return out


def prepare_ute(ktraj, dcf, kdata):
    """The reference's UTE preparation (test/MRI_recon/test_uterecon.py:87-99) on arrays:
    returns (im_shape, ktraj scaled by 1/im_shape, dcf, kdata as (M, nc) complex64)."""
    ktraj = np.array(ktraj, dtype=np.float64)
    kdata = np.asarray(kdata).astype(np.complex64)
    ncoils = kdata.shape[-1]
    im_shape = [2 * int(np.max(ktraj[a])) for a in range(3)]
    for a in range(3):
        ktraj[a] = ktraj[a] * (1.0 / im_shape[a])
    kdata = kdata.reshape((int(np.prod(kdata.shape[:-1])), ncoils)).squeeze()
    return im_shape, ktraj, np.asarray(dcf), kdata


def load_ute_mat(path):
    """``prepare_ute`` of the ``ktraj``, ``dcf`` and ``kdata`` variables of a .mat file."""
    m = loadmat(path, ["ktraj", "dcf", "kdata"])
    im_shape, ktraj, dcf, kdata = prepare_ute(m["ktraj"], m["dcf"], m["kdata"])
    return {"im_shape": im_shape, "ktraj": ktraj, "dcf": dcf, "kdata": kdata}


def load_spiral_mat(path, im_shape=(256, 256)):
    """The 2-D spiral dataset: coordinates in radians (x = 2 pi Re k, y = 2 pi Im k, flattened
    column-major like the file's interleaves), weights (M,), samples (M, nc)."""
    m = loadmat(path, ["k", "w", "data"])
    k = np.asarray(m["k"])
    x = (2 * np.pi * k.real).reshape(-1, order="F")
    y = (2 * np.pi * k.imag).reshape(-1, order="F")
    w = np.asarray(m["w"], dtype=np.float64).reshape(-1, order="F")
    d = np.asarray(m["data"])
    if d.ndim == k.ndim:
        d = d[..., None]
    # the coil axis is last; the sample axes flatten in the same (column-major) order as k
    kdata = np.stack([d[..., c].reshape(-1, order="F") for c in range(d.shape[-1])], 1)
    return {"coords": [x, y], "dcf": w, "kdata": kdata.astype(np.complex64), "im_shape": tuple(im_shape)}
