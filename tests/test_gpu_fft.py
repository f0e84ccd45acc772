"""GPU tests of the pruned in-place FFT passes fused with pad/crop.

The same plan is run with the pruned engine and with rocFFT (separate pad and
crop): results agree to FP32 roundoff (rel-L2 <= 2e-6), for 2-D and 3-D, radix
2/3/4/5 lengths, coil-last and coil-major layouts, both signs, 1 to 20 coils (the
row passes take coils in chunks of up to 16), and every transform kind; the
trajectory gradient uses the same passes.  Lengths with other factors fall back
to rocFFT.
"""
import numpy as np
import pytest

from oracle import nufft_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.detach().cpu().numpy().astype(np.complex128).ravel()
    b = b.detach().cpu().numpy().astype(np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


CASES = [
    ((20, 24), 3, 1, "coil_last"),          # nf 40, 48: radix 2*4*5, 4*4*3
    ((64, 96), 1, -1, "coil_major"),        # 128, 192
    ((48, 40), 20, 1, "coil_last"),         # 20 coils: chunked row passes
    ((16, 12, 10), 2, 1, "coil_last"),      # 3-D nf 32, 24, 20
    ((32, 32, 32), 5, -1, "coil_major"),
    ((24, 20, 30), 8, 1, "coil_last"),      # 48, 40, 60
]


@pytest.mark.parametrize("modes,nc,iflag,layout", CASES)
def test_pruned_engine_matches_rocfft(modes, nc, iflag, layout):
    import torch
    from mripy_amd.plan import Plan
    print("case: pruned FFT vs rocFFT", modes, nc, iflag, layout, flush=True)
    d = len(modes)
    rng = np.random.default_rng(sum(modes) + nc)
    M = 3000 if d == 2 else 6000
    cs = [(np.pi * (2 * rng.random(M) - 1)).astype(np.float32) for _ in range(d)]
    dcf = rng.random(M).astype(np.float32)
    P = Plan(modes, ncoils=nc, rat=2, nspread=3, iflag=iflag, layout=layout).set_points(cs, dcf=dcf)
    img = torch.as_tensor((rng.standard_normal(P.image_shape()) + 1j * rng.standard_normal(P.image_shape()))
                          .astype(np.complex64)).cuda()
    y = torch.as_tensor((rng.standard_normal(P.sample_shape()) + 1j * rng.standard_normal(P.sample_shape()))
                        .astype(np.complex64)).cuda()
    outs = {}
    for eng in ("auto", "rocfft"):
        P.set_fft_engine(eng)
        outs[eng] = (P.type2(img), P.type1(y), P.adjoint(y), P.type21(img))
    for a, b in zip(outs["auto"], outs["rocfft"]):
        assert rel(a, b) <= 2e-6
    # and against the FP64 oracle
    p = O.Params(modes, rat=2, nspread=3, iflag=iflag)
    im_np = img.cpu().numpy()
    if layout == "coil_major":
        im_np = np.moveaxis(im_np, 0, -1)
    got = outs["auto"][0].cpu().numpy()
    if layout == "coil_major":
        got = got.T
    want = O.type2([c.astype(np.float64) for c in cs], im_np, p)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 1e-5


def test_traj_grad_engines_agree():
    import torch
    from mripy_amd.plan import Plan
    rng = np.random.default_rng(3)
    M, modes = 4000, (16, 20, 18)
    cs = [(np.pi * (2 * rng.random(M) - 1)).astype(np.float32) for _ in range(3)]
    P = Plan(modes, ncoils=3, rat=2, nspread=3).set_points(cs)
    F = torch.as_tensor((rng.standard_normal(modes + (3,)) + 1j * rng.standard_normal(modes + (3,)))
                        .astype(np.complex64)).cuda()
    w = torch.as_tensor((rng.standard_normal((M, 3)) + 1j * rng.standard_normal((M, 3))).astype(np.complex64)).cuda()
    g1 = P.traj_grad(F, w)
    P.set_fft_engine("rocfft")
    g2 = P.traj_grad(F, w)
    assert float(torch.linalg.vector_norm(g1 - g2) / torch.linalg.vector_norm(g2)) <= 2e-6


def test_unsupported_length_falls_back_to_rocfft():
    from mripy_amd.plan import Plan
    rng = np.random.default_rng(4)
    modes = (7, 9)          # nf 14 = 2 * 7: not a product of 2, 3, 5
    cs = [(np.pi * (2 * rng.random(500) - 1)).astype(np.float32) for _ in range(2)]
    P = Plan(modes, ncoils=1, rat=2, nspread=3).set_points(cs)
    F = (rng.standard_normal(modes + (1,)) + 1j * rng.standard_normal(modes + (1,))).astype(np.complex64)
    p = O.Params(modes, rat=2, nspread=3)
    got = P.type2(F)
    want = O.type2([c.astype(np.float64) for c in cs], F, p)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 1e-5
