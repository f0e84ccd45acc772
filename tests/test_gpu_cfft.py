"""Cartesian FFTs (gn_fft_c2c) and the masked Cartesian operators.

Against numpy.fft in FP64 on the same complex64 inputs (rel-L2 <= 2e-6): the cuFFT
wrapper semantics of fft/cufft.py:27-97 (transform over the first len(axes) axes,
batched over the trailing coil axes, inverse divided by the transformed size) and
FFTnd_cuda_kmask / FFT2d_cuda_kmask (pics/operators_cuda_class.py:7-58:
forward = ifftshift(fftn(fftshift(im))) * mask with dim_match broadcasting,
backward = ifftshift(ifftn(fftshift(ksp)))), for even and odd lengths, with and
without a coil axis, numpy and device-tensor inputs, shifts over the transformed
axes (fused) and over other axes (separate rolls).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    if hasattr(a, "cpu"):
        a = a.cpu().numpy()
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _x(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("shape,rank", [((32, 24, 20), 3), ((16, 12, 10, 4), 3), ((15, 17, 9, 3), 3),
                                        ((64, 48), 2), ((33, 20, 5), 2), ((128,), 1), ((40, 7), 1)])
def test_c2c_matches_numpy(shape, rank):
    import torch
    from mripy_amd import cfft
    print("case: c2c", shape, rank, flush=True)
    x = _x(shape, sum(shape))
    axes = tuple(range(rank))
    xf = x.astype(np.complex128)
    f = cfft.fftnc2c_hip(x, axes)
    assert f.dtype == np.complex64 and rel(f, np.fft.fftn(xf, axes=axes)) <= 2e-6
    b = cfft.ifftnc2c_hip(x, axes)
    assert rel(b, np.fft.ifftn(xf, axes=axes)) <= 2e-6
    xt = torch.as_tensor(x).cuda()
    ft = cfft.fft2c2c_hip(xt, axes) if rank == 2 else cfft.fftnc2c_hip(xt, axes)
    assert ft.is_cuda and rel(ft, np.fft.fftn(xf, axes=axes)) <= 2e-6


def _kmask_ref(x, mask, axes, rank, dim_match):
    k = np.fft.fftshift(x.astype(np.complex128), axes)
    k = np.fft.fftn(k, axes=tuple(range(rank)))
    k = np.fft.ifftshift(k, axes)
    if dim_match and mask.ndim < k.ndim:
        mask = mask.reshape(mask.shape + (1,) * (k.ndim - mask.ndim))
    return k * mask


def _kmask_back_ref(k, axes, rank):
    x = np.fft.fftshift(k.astype(np.complex128), axes)
    x = np.fft.ifftn(x, axes=tuple(range(rank)))
    return np.fft.ifftshift(x, axes)


@pytest.mark.parametrize("shape,maskshape,axes", [((32, 24, 20, 4), (32, 24, 20), (0, 1, 2)),
                                                  ((15, 16, 9, 2), (15, 16, 9, 2), (0, 1, 2)),
                                                  ((16, 16, 8), (16, 16, 8), (0, 1, 2)),
                                                  ((16, 16, 8, 3), (16, 16, 8), (0, 1))])
def test_fftnd_kmask(shape, maskshape, axes):
    import torch
    from mripy_amd.cfft import FFTnd_cuda_kmask
    print("case: FFTnd kmask", shape, maskshape, axes, flush=True)
    rng = np.random.default_rng(len(shape) + shape[0])
    x = _x(shape, 7)
    mask = (rng.random(maskshape) > 0.6).astype(np.float32)
    op = FFTnd_cuda_kmask(mask, axes)
    k = op.forward(x)
    assert rel(k, _kmask_ref(x, mask, axes, 3, True)) <= 2e-6
    assert rel(op.backward(x), _kmask_back_ref(x, axes, 3)) <= 2e-6
    kt = op.forward(torch.as_tensor(x).cuda())
    assert kt.is_cuda and rel(kt, _kmask_ref(x, mask, axes, 3, True)) <= 2e-6


@pytest.mark.parametrize("shape", [(64, 48), (31, 20), (32, 32, 3)])
def test_fft2d_kmask(shape):
    from mripy_amd.cfft import FFT2d_cuda_kmask
    print("case: FFT2d kmask", shape, flush=True)
    rng = np.random.default_rng(shape[0])
    x = _x(shape, 9)
    mask = (rng.random(shape) > 0.5).astype(np.float32)
    op = FFT2d_cuda_kmask(mask)
    assert rel(op.forward(x), _kmask_ref(x, mask, (0, 1), 2, False)) <= 2e-6
    assert rel(op.backward(x), _kmask_back_ref(x, (0, 1), 2)) <= 2e-6


def test_c2c_rejects_bad_arguments():
    import ctypes
    from mripy_amd import _lib
    L = _lib.load()
    shp = (ctypes.c_int64 * 3)(8, 8, 1)
    assert L.gn_fft_c2c(4, shp, 1, 1, 0, None, 0, None, None, None) == _lib.GN_EINVAL
    assert L.gn_fft_c2c(2, shp, 1, 7, 0, None, 0, 1, 1, None) == _lib.GN_EINVAL
