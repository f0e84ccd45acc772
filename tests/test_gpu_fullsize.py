"""GPU parity at the full workload sizes and at the reference's own defaults.

* 3-D UTE multi-coil: 128^3 image x 8 coils, 4096 x 256 Fibonacci spokes
  (M = 1,048,576), rat 2, nspread 3 -- type 2 and type 1 through the C ABI
  against the FP64 C oracle on the same float32 inputs, rel-L2 <= 1e-5.
* Differentiable NUFFT at full geometry: 128^3 x 4 coils, 1,048,576 UTE samples
  perturbed by U(-0.1/N, 0.1/N), loss ||A(x) f - y||^2 with y = A f0 + 0.01 noise.
  Image gradient against 2 A^H r in FP64 (rel-L2 <= 1e-4); trajectory gradient
  against the FP64 analytic gradient on 4096 random samples evaluated on the full
  FP64 grid (exact: d L / d x_j depends only on sample j; rel-L2 <= 1e-3); central
  finite differences of the FP64 loss at 12 samples (<= 1e-3 relative), where a
  perturbation of sample j changes only y_j, so one-sample FP64 gathers suffice.
* The reference's known-answer inputs (fft/nufft_test_func.py:15-66:
  RandomState(0), x = 100 rand, sizes fft/nufft_func.py:1350-1352) at its default
  eps = 1e-15 (rat 3; nspread 14 in 2-D, 6 in 3-D) through the reference-named
  functions on the HIP path, against the brute-force NUDFT and the FP64 oracle.
"""
import numpy as np
import pytest

from oracle import coracle as C
from oracle import nufft_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def test_3d_ute_8coil_full_size_against_fp64_oracle():
    import torch
    from mripy_amd import data
    from mripy_amd.plan import Plan
    print("case: 3-D UTE 128^3 x 8 coils, M=1048576, nspread 3", flush=True)
    cfg = data.config_3d_ute(n=128, nc=8, nspokes=4096, nread=256)
    coords = cfg["coords"]
    p = O.Params((128, 128, 128), rat=2, nspread=3)
    P = Plan((128, 128, 128), ncoils=8, rat=2, nspread=3)
    P.set_points([torch.as_tensor(c, device="cuda") for c in coords])
    y = P.type2(torch.as_tensor(cfg["image"], device="cuda")).cpu().numpy()
    im = P.type1(torch.as_tensor(cfg["samples"], device="cuda")).cpu().numpy()
    torch.cuda.synchronize()
    yo = C.type2(coords, cfg["image"], p)
    e2 = rel(y, yo)
    del yo
    io = C.type1(coords, cfg["samples"], p)
    e1 = rel(im, io)
    print("full-size rel-L2: type 2 %.3e, type 1 %.3e" % (e2, e1), flush=True)
    assert e2 <= 1e-5 and e1 <= 1e-5, (e2, e1)


def test_differentiable_full_geometry_4coil():
    import torch
    from mripy_amd import data
    from mripy_amd.autograd import NUFFT
    n, nc = 128, 4
    print("case: autograd 128^3 x 4 coils, M=1048576 perturbed UTE", flush=True)
    x, y, z = data.fibonacci_ute_3d(4096, 256)
    rng = np.random.default_rng(2)
    coords = [(c + 2 * np.pi * rng.uniform(-0.1 / n, 0.1 / n, c.shape[0])).astype(np.float32) for c in (x, y, z)]
    S = data.coil_maps((n, n, n), nc, seed=0)
    f0 = (data.shepp_logan_3d(n)[..., None] * S).astype(np.complex64)
    p = O.Params((n, n, n), rat=2, nspread=3)
    y0 = C.type2(coords, f0, p)
    ymeas = (y0 + data.complex_normal(y0.shape, seed=1, sigma=0.01)).astype(np.complex64)
    del y0
    f = (0.9 * f0 + 0.05 * data.complex_normal(f0.shape, seed=3)).astype(np.complex64)

    op = NUFFT((n, n, n), ncoils=nc, rat=2, nspread=3)
    ct = torch.tensor(np.stack(coords), device="cuda", requires_grad=True)
    ft = torch.tensor(f, device="cuda", requires_grad=True)
    yt = torch.tensor(ymeas, device="cuda")
    loss = (op(ft, ct) - yt).abs().pow(2).sum()
    loss.backward()
    g_img = ft.grad.cpu().numpy()
    g_crd = ct.grad.cpu().numpy()
    print("gpu loss %.6e" % float(loss), flush=True)

    M = coords[0].shape[0]
    grid = C.forward_grid(f, p)                      # FP64 oversampled grid of A f
    r = C.gather(coords, grid, p.nf, p.ns, p.tau) - ymeas
    ref_img = 2 * M * C.type1(coords, r, p)          # 2 A^H r, A^H = M type 1
    e_img = rel(g_img, ref_img)
    del ref_img
    sub = np.sort(rng.choice(M, 4096, replace=False))
    ref_crd = O.traj_grad_on_grid([c[sub] for c in coords], grid, 2 * r[sub], p)
    e_crd = float(np.linalg.norm(g_crd[:, sub] - ref_crd) / np.linalg.norm(ref_crd))
    print("rel-L2: image gradient %.3e, trajectory gradient (4096 samples) %.3e" % (e_img, e_crd), flush=True)
    assert e_img <= 1e-4
    assert e_crd <= 1e-3

    # central differences of the FP64 loss; only sample j's term depends on x_j
    step = 2 * np.pi * 1e-3 / n
    h = 2 * np.pi / p.nf[0]
    c64 = [c.astype(np.float64) for c in coords]

    def loss_j(j, a, dx):
        cs = [np.array([c[j]]) for c in c64]
        cs[a] = cs[a] + dx
        yj = C.gather(cs, grid, p.nf, p.ns, p.tau)[0]
        return float(np.sum(np.abs(yj - ymeas[j]) ** 2))

    checked = 0
    gmax = np.abs(ref_crd).max()
    for j in rng.permutation(M):
        a = int(rng.integers(3))
        u = np.mod(c64[a][j], 2 * np.pi) / h
        frac = u - np.floor(u)
        if not (2 * step / h < frac < 1 - 2 * step / h):
            continue
        fd = (-loss_j(j, a, 2 * step) + 8 * loss_j(j, a, step) - 8 * loss_j(j, a, -step)
              + loss_j(j, a, -2 * step)) / (12 * step)
        if abs(fd) < 1e-2 * gmax:
            continue
        assert abs(g_crd[a, j] - fd) <= 1e-3 * abs(fd), (j, a, g_crd[a, j], fd)
        checked += 1
        if checked == 12:
            break
    assert checked == 12


@pytest.mark.parametrize("iflag", [1, -1])
@pytest.mark.parametrize("df", [1.0, 2.0])
def test_reference_kat_2d_default_eps(df, iflag):
    from mripy_amd import functional as F
    print("case: KAT 2-D (16,16,25000) eps=1e-15 df=%g iflag=%d" % (df, iflag), flush=True)
    rng = np.random.RandomState(0)
    mc = 25000
    x = 100 * rng.rand(mc)
    y = 100 * rng.rand(mc)
    c = np.sin(2 * x) + 1j * np.cos(2 * y)
    g = F.nufft2d1_gaussker(x, y, c, 16, 16, df=df, iflag=iflag)
    ref = O.nudft_type1([x, y], c, (16, 16), df=df, iflag=iflag)
    orc = O.nufft2d1_gaussker(x, y, c, 16, 16, df=df, iflag=iflag)
    assert np.allclose(g, ref, rtol=1e-2, atol=1e-2)      # the reference's own assertion
    assert rel(g, ref) <= 1e-5 and rel(g, orc) <= 1e-5, (rel(g, ref), rel(g, orc))
    # the forward of the same geometry at eps=1e-15 (rat 3, nspread 14)
    Fk = np.sin(np.arange(256)).reshape(16, 16) + 1j * np.cos(np.arange(256) * 0.3).reshape(16, 16)
    gy = F.nufft2d2_gaussker(x, y, Fk, 16, 16, df=df, iflag=iflag)
    ry = O.nudft_type2([x, y], Fk, (16, 16), df=df, iflag=iflag)
    assert rel(gy, ry) <= 1e-5, rel(gy, ry)


def _nudft_type1_chunked(coords, c, modes, iflag, chunk=2048):
    """Brute-force type 1 summed over sample chunks (bounded memory)."""
    M = coords[0].shape[0]
    out = 0
    for s in range(0, M, chunk):
        e = min(M, s + chunk)
        out = out + O.nudft_type1([x[s:e] for x in coords], c[s:e], modes, iflag=iflag) * (e - s)
    return out / M


@pytest.mark.parametrize("iflag", [1, -1])
def test_reference_kat_3d_default_eps(iflag):
    from mripy_amd import functional as F
    print("case: KAT 3-D (32,32,16,20480) eps=1e-15 iflag=%d" % iflag, flush=True)
    rng = np.random.RandomState(0)
    mc = 20480
    x = 100 * rng.rand(mc)
    y = 100 * rng.rand(mc)
    z = 100 * rng.rand(mc)
    c = np.sin(2 * x) + 1j * np.cos(2 * y)
    g = F.nufft3d1_gaussker(x, y, z, c, 32, 32, 16, iflag=iflag)
    orc = C.type1([x, y, z], c, O.Params((32, 32, 16), eps=1e-15, iflag=iflag, max_nspread=6))
    ref = _nudft_type1_chunked([x, y, z], c, (32, 32, 16), iflag)
    e_ref, e_orc = rel(g, ref), rel(g, orc)
    print("rel-L2: vs NUDFT %.3e, vs oracle %.3e" % (e_ref, e_orc), flush=True)
    assert np.allclose(g, ref, rtol=1e-2, atol=1e-2)      # the reference's own assertion
    assert e_orc <= 1e-5, e_orc
    # tau uses ms = 32 for every axis (fft/nufft_func.py:248), so with nspread clamped
    # to 6 the short third axis limits the reference kernel itself to ~8e-3 here
    assert e_ref <= 1e-2, e_ref
