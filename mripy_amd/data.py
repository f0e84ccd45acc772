"""Seeded synthetic MRI data: phantoms, non-Cartesian trajectories, coil maps.

Trajectories are returned in radians (the reference's convention, periodic in
2 pi; pics/operators_cuda_class.py:85-90 uses radians), computed from k-space
positions in cycles/FOV in [-0.5, 0.5) multiplied by 2 pi, and rounded to
float32 so that the device path and the FP64 oracle see identical inputs.
"""
from __future__ import annotations

import numpy as np

GOLDEN = (1 + 5 ** 0.5) / 2

# modified Shepp-Logan (Toft): intensity, a, b, x0, y0, phi(deg)
_SL2D = np.array([
    [1.0, .69, .92, 0, 0, 0],
    [-.8, .6624, .8740, 0, -.0184, 0],
    [-.2, .1100, .3100, .22, 0, -18],
    [-.2, .1600, .4100, -.22, 0, 18],
    [.1, .2100, .2500, 0, .35, 0],
    [.1, .0460, .0460, 0, .1, 0],
    [.1, .0460, .0460, 0, -.1, 0],
    [.1, .0460, .0230, -.08, -.605, 0],
    [.1, .0230, .0230, 0, -.606, 0],
    [.1, .0230, .0460, .06, -.605, 0],
])

# 3-D Shepp-Logan ellipsoids (Kak-Slaney geometry, modified intensities):
# intensity, a, b, c, x0, y0, z0, phi(deg, about z)
_SL3D = np.array([
    [1.0, .6900, .920, .810, 0, 0, 0, 0],
    [-.8, .6624, .874, .780, 0, -.0184, 0, 0],
    [-.2, .1100, .310, .220, .22, 0, 0, -18],
    [-.2, .1600, .410, .280, -.22, 0, 0, 18],
    [.1, .2100, .250, .410, 0, .35, -.15, 0],
    [.1, .0460, .046, .050, 0, .1, .25, 0],
    [.1, .0460, .046, .050, 0, -.1, .25, 0],
    [.1, .0460, .023, .050, -.08, -.605, 0, 0],
    [.1, .0230, .023, .020, 0, -.606, 0, 0],
    [.1, .0230, .046, .020, .06, -.605, 0, 0],
])


def _centres(n):
    """Pixel-centre coordinates in [-1, 1)."""
    return (np.arange(n) + 0.5) / n * 2 - 1


def shepp_logan_2d(n, dtype=np.complex64):
    """Analytic modified Shepp-Logan sampled at pixel centres, (n, n), zero imaginary part."""
    c = _centres(n)
    X, Y = np.meshgrid(c, c, indexing="ij")
    img = np.zeros((n, n))
    for I, a, b, x0, y0, phi in _SL2D:
        p = np.deg2rad(phi)
        xr = (X - x0) * np.cos(p) + (Y - y0) * np.sin(p)
        yr = -(X - x0) * np.sin(p) + (Y - y0) * np.cos(p)
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1] += I
    return img.astype(dtype)


def shepp_logan_3d(n, dtype=np.complex64):
    """3-D ellipsoid Shepp-Logan phantom, (n, n, n)."""
    c = _centres(n)
    X, Y, Z = np.meshgrid(c, c, c, indexing="ij")
    img = np.zeros((n, n, n), dtype=np.float32)
    for I, a, b, cc, x0, y0, z0, phi in _SL3D:
        p = np.deg2rad(phi)
        xr = (X - x0) * np.cos(p) + (Y - y0) * np.sin(p)
        yr = -(X - x0) * np.sin(p) + (Y - y0) * np.cos(p)
        img[(xr / a) ** 2 + (yr / b) ** 2 + ((Z - z0) / cc) ** 2 <= 1] += I
    return img.astype(dtype)


def golden_angle_radial_2d(nspokes, nread):
    """Golden-angle radial: spoke s at angle s*pi/phi, readout r_k = (k - nread/2)/nread.

    Returns float32 (x, y) in radians, shape (nspokes*nread,).
    """
    th = np.arange(nspokes) * np.pi / GOLDEN
    r = (np.arange(nread) - nread // 2) / nread
    x = 2 * np.pi * r[None, :] * np.cos(th)[:, None]
    y = 2 * np.pi * r[None, :] * np.sin(th)[:, None]
    return x.reshape(-1).astype(np.float32), y.reshape(-1).astype(np.float32)


def fibonacci_ute_3d(nspokes, nread, rmax_den=None):
    """Centre-out 3-D radial (UTE) on Fibonacci-sphere directions.

    z_s = 1 - (2s+1)/nspokes, phi_s = s*pi*(3 - sqrt 5); radii r_k = k/(2 nread)
    cycles/FOV (k = 0..nread-1).  Returns float32 (x, y, z) in radians.
    """
    s = np.arange(nspokes)
    zs = 1 - (2 * s + 1) / nspokes
    ph = s * np.pi * (3 - np.sqrt(5))
    rho = np.sqrt(np.maximum(0.0, 1 - zs ** 2))
    d = np.stack([rho * np.cos(ph), rho * np.sin(ph), zs], 0)   # (3, nspokes)
    r = np.arange(nread) / (rmax_den or 2 * nread)
    k = 2 * np.pi * d[:, :, None] * r[None, None, :]            # (3, nspokes, nread)
    return tuple(k[a].reshape(-1).astype(np.float32) for a in range(3))


def archimedean_spiral_2d(ninter, npts, n):
    """Variable-density Archimedean spiral: r(t) = 0.5 t^1.5 cycles/FOV,
    theta_i(t) = 2 pi T t + 2 pi i/ninter with T = n/(2 ninter) turns."""
    t = np.arange(npts) / npts
    T = n / (2.0 * ninter)
    r = 0.5 * t ** 1.5
    th = 2 * np.pi * T * t[None, :] + 2 * np.pi * np.arange(ninter)[:, None] / ninter
    x = 2 * np.pi * r[None, :] * np.cos(th)
    y = 2 * np.pi * r[None, :] * np.sin(th)
    return x.reshape(-1).astype(np.float32), y.reshape(-1).astype(np.float32)


def coil_maps(shape, nc, seed=0, dtype=np.complex64):
    """Smooth complex Gaussian-blob sensitivities, coil-last (*shape, nc), sum_c |S_c|^2 = 1.

    Blob centres lie on a circle (2-D) or Fibonacci sphere (3-D) of radius 0.6 FOV;
    each map carries a smooth linear phase drawn from ``seed``.
    """
    d = len(shape)
    rng = np.random.default_rng(seed)
    axes = [_centres(n) for n in shape]
    grids = np.meshgrid(*axes, indexing="ij")
    if d == 2:
        ang = 2 * np.pi * np.arange(nc) / nc
        cen = 0.6 * np.stack([np.cos(ang), np.sin(ang)], 1)
    else:
        s = np.arange(nc)
        zs = 1 - (2 * s + 1) / nc
        ph = s * np.pi * (3 - np.sqrt(5))
        rho = np.sqrt(np.maximum(0.0, 1 - zs ** 2))
        cen = 0.6 * np.stack([rho * np.cos(ph), rho * np.sin(ph), zs], 1)
    S = np.zeros(tuple(shape) + (nc,), dtype=np.complex128)
    for c in range(nc):
        r2 = sum((g - cen[c, a]) ** 2 for a, g in enumerate(grids))
        phase = sum(rng.uniform(-np.pi, np.pi) * g for g in grids) + rng.uniform(-np.pi, np.pi)
        S[..., c] = np.exp(-r2 / (2 * 0.5 ** 2)) * np.exp(1j * phase)
    S /= np.sqrt(np.sum(np.abs(S) ** 2, axis=-1, keepdims=True))
    return S.astype(dtype)


def complex_normal(shape, seed=0, sigma=1.0, dtype=np.complex64):
    """(N(0,1) + i N(0,1)) * sigma from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    return (sigma * (rng.standard_normal(shape) + 1j * rng.standard_normal(shape))).astype(dtype)


# ---------------------------------------------------------------------------
# benchmark / test configurations
# ---------------------------------------------------------------------------

def config_2d_radial(n=128, nspokes=128, nread=256, seed=0):
    """Plain 2-D: n^2 Shepp-Logan, golden-angle radial nspokes x nread, one coil."""
    x, y = golden_angle_radial_2d(nspokes, nread)
    img = shepp_logan_2d(n)
    c = complex_normal(x.shape[0], seed)
    return dict(coords=[x, y], image=img, samples=c, modes=(n, n))


def config_3d_ute(n=128, nc=8, nspokes=4096, nread=256, seed=0):
    """Plain 3-D multi-coil: n^3 phantom x nc Gaussian-blob maps, Fibonacci UTE radial."""
    x, y, z = fibonacci_ute_3d(nspokes, nread)
    ph = shepp_logan_3d(n)
    S = coil_maps((n, n, n), nc, seed=seed)
    img = (ph[..., None] * S).astype(np.complex64)
    c = complex_normal((x.shape[0], nc), seed)
    return dict(coords=[x, y, z], image=img, samples=c, modes=(n, n, n), maps=S, phantom=ph)


def config_2d_spiral_cs(n=256, nc=32, ninter=48, npts=2048, sigma=1e-3, seed=5):
    """Multi-coil 2-D CS reconstruction workload: n^2 Shepp-Logan phantom, nc
    Gaussian-blob maps, variable-density Archimedean spiral ninter x npts, complex
    noise of std ``sigma`` from ``seed`` (added to the simulated k-space by the
    caller, which owns the forward model)."""
    x, y = archimedean_spiral_2d(ninter, npts, n)
    S = coil_maps((n, n), nc, seed=seed)
    ph = shepp_logan_2d(n)
    noise = complex_normal((x.shape[0], nc), seed=seed, sigma=sigma)
    return dict(coords=[x, y], maps=S, phantom=ph, noise=noise, modes=(n, n))


def config_batched_uneven(ninst=64, seed=4, sizes=(128, 192, 256, 320), spokes=(32, 400), coils=(1, 32),
                          with_images=True):
    """Uneven 2-D multi-instance workload: per instance N drawn from ``sizes``, a
    golden-angle radial trajectory of U[spokes] spokes x 2N readout points, U[coils]
    coils with smooth Gaussian-blob maps, image = Shepp-Logan x maps (default_rng(seed))."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(ninst):
        N = int(rng.choice(sizes))
        nsp = int(rng.integers(spokes[0], spokes[1] + 1))
        nc = int(rng.integers(coils[0], coils[1] + 1))
        x, y = golden_angle_radial_2d(nsp, 2 * N)
        inst = dict(modes=(N, N), coords=[x, y], ncoils=nc)
        if with_images:
            S = coil_maps((N, N), nc, seed=seed * 1000 + i)
            inst["image"] = (shepp_logan_2d(N)[..., None] * S).astype(np.complex64)
        out.append(inst)
    return out
