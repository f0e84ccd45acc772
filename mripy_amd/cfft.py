"""Cartesian FFTs and masked Cartesian operators on the HIP path.

* ``fftnc2c_hip`` / ``ifftnc2c_hip`` / ``fft2c2c_hip`` / ``ifft2c2c_hip``: the cuFFT
  wrappers of fft/cufft.py:27-97 -- complex64, a transform over the first
  ``len(axes)`` axes batched over the remaining (coil) axes, the inverse divided by
  the product of the transformed lengths.  ``*_cuda`` names are aliases.
* ``FFTnd_kmask_hip`` / ``FFT2d_kmask_hip``: ``FFTnd_cuda_kmask`` / ``FFT2d_cuda_kmask``
  (pics/operators_cuda_class.py:7-58): forward = ifftshift(FFT(fftshift(im))) * mask,
  backward = ifftshift(IFFT(fftshift(ksp))), shifts over ``axes``.  When the shift
  axes are the transformed axes (the default) the shifts and the mask run inside
  ``gn_fft_c2c``'s passes before and after rocFFT; otherwise the shifts are separate
  device rolls.

numpy inputs give numpy outputs (complex64, host-synchronous like the reference);
torch device tensors stay on the device.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import GN_FFT_FORWARD, GN_FFT_INVERSE, GN_FFT_SHIFT_IN, GN_FFT_SHIFT_OUT, check


def _dev(x):
    if torch.is_tensor(x) and x.is_cuda:
        return x.device
    return torch.device("cuda", torch.cuda.current_device())


def _c2c(x, rank, direction, flags=0, mask=None):
    is_np = not torch.is_tensor(x)
    dev = _dev(x)
    t = torch.as_tensor(x).to(device=dev, dtype=torch.complex64).contiguous()
    if t.dim() < rank:
        raise ValueError("need at least %d dimensions" % rank)
    lead = tuple(t.shape[:rank])
    batch = int(np.prod(t.shape[rank:])) if t.dim() > rank else 1
    out = torch.empty_like(t)
    m, mb = None, 0
    if mask is not None:
        m = torch.as_tensor(mask).to(device=dev, dtype=torch.float32).contiguous()
        mb = 1 if m.numel() == int(np.prod(lead)) else batch
    shp = (ctypes.c_int64 * 3)(*(list(lead) + [1] * (3 - rank)))
    if t.numel():
        with torch.cuda.device(dev):
            stream = torch.cuda.current_stream(dev).cuda_stream
            check(_lib.load().gn_fft_c2c(rank, shp, batch, direction, flags, m.data_ptr() if m is not None else None,
                                         mb, t.data_ptr(), out.data_ptr(), ctypes.c_void_p(stream)))
    return out.cpu().numpy() if is_np else out


def fftnc2c_hip(x, axes=(0, 1, 2)):
    """Unnormalized forward FFT over the first len(axes) axes (fft/cufft.py:27-43)."""
    return _c2c(x, len(axes), GN_FFT_FORWARD)


def ifftnc2c_hip(x, axes=(0, 1, 2)):
    """Inverse FFT over the first len(axes) axes, divided by their size (fft/cufft.py:46-61)."""
    return _c2c(x, len(axes), GN_FFT_INVERSE)


def fft2c2c_hip(x, axes=(0, 1)):
    """fft/cufft.py:64-80."""
    return _c2c(x, len(axes), GN_FFT_FORWARD)


def ifft2c2c_hip(x, axes=(0, 1)):
    """fft/cufft.py:83-97."""
    return _c2c(x, len(axes), GN_FFT_INVERSE)


fftnc2c_cuda = fftnc2c_hip
ifftnc2c_cuda = ifftnc2c_hip
fft2c2c_cuda = fft2c2c_hip
ifft2c2c_cuda = ifft2c2c_hip


def _roll(x, axes, inverse):
    """fftshift (inverse=False) or ifftshift over axes, on either array type."""
    if torch.is_tensor(x):
        shifts = [(x.shape[a] // 2) * (-1 if inverse else 1) for a in axes]
        return torch.roll(x, shifts, dims=list(axes))
    return np.fft.ifftshift(x, axes) if inverse else np.fft.fftshift(x, axes)


class _KMask:
    rank = 3
    broadcast = "dim_match"

    def __init__(self, mask, axes):
        self.mask = mask
        self.axes = tuple(axes)

    def _fused(self, x):
        return tuple(sorted(self.axes)) == tuple(range(self.rank)) and len(x.shape) >= self.rank

    def _mask_fusable(self, shape):
        """A real mask of the full shape, or of the transformed axes where it broadcasts
        over the batch (dim_match; numpy broadcasting only without batch axes)."""
        m = self.mask
        if (m.is_complex() if torch.is_tensor(m) else np.iscomplexobj(m)):
            return False
        ms = tuple(m.shape)
        if ms == tuple(shape):
            return True
        return ms == tuple(shape[:self.rank]) and (self.broadcast == "dim_match" or len(shape) == self.rank)

    def _apply_mask(self, ksp):
        mask = self.mask
        if torch.is_tensor(ksp):
            mask = torch.as_tensor(mask, device=ksp.device)
        if self.broadcast == "dim_match" and len(ksp.shape) != len(mask.shape):
            from .operators import dim_match
            ks, ms = dim_match(tuple(ksp.shape), tuple(mask.shape))
            return ksp.reshape(ks) * mask.reshape(ms)
        return ksp * mask

    def forward(self, im):
        """k-space <- image: ifftshift(FFT(fftshift(im))) * mask."""
        if self._fused(im) and self._mask_fusable(tuple(im.shape)):
            return _c2c(im, self.rank, GN_FFT_FORWARD, GN_FFT_SHIFT_IN | GN_FFT_SHIFT_OUT, self.mask)
        if self._fused(im):
            ksp = _c2c(im, self.rank, GN_FFT_FORWARD, GN_FFT_SHIFT_IN | GN_FFT_SHIFT_OUT)
        else:
            ksp = _roll(_c2c(_roll(im, self.axes, False), self.rank, GN_FFT_FORWARD), self.axes, True)
        return self._apply_mask(ksp)

    def backward(self, ksp):
        """image <- k-space: ifftshift(IFFT(fftshift(ksp)))."""
        if self._fused(ksp):
            return _c2c(ksp, self.rank, GN_FFT_INVERSE, GN_FFT_SHIFT_IN | GN_FFT_SHIFT_OUT)
        return _roll(_c2c(_roll(ksp, self.axes, False), self.rank, GN_FFT_INVERSE), self.axes, True)


class FFTnd_kmask_hip(_KMask):
    """3-D FFT with a k-space mask (pics/operators_cuda_class.py:7-35): the transform
    covers the first three axes (fftnc2c_cuda's default); the mask broadcasts with
    dim_match (trailing 1s)."""
    rank = 3
    broadcast = "dim_match"

    def __init__(self, mask, axes=(0, 1, 2)):
        super().__init__(mask, axes)


class FFT2d_kmask_hip(_KMask):
    """2-D FFT with a k-space mask (pics/operators_cuda_class.py:37-58): transform over
    the first len(axes) axes; the mask multiplies with numpy broadcasting."""
    rank = 2
    broadcast = "numpy"

    def __init__(self, mask, axes=(0, 1)):
        super().__init__(mask, axes)
        self.rank = len(self.axes)


FFTnd_cuda_kmask = FFTnd_kmask_hip
FFT2d_cuda_kmask = FFT2d_kmask_hip
