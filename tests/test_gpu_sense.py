"""SENSE transforms fused into the NUFFT's first and last FFT passes.

joint2operators(espirit(S), NUFFT) (pics/operators_class.py:437-472, 496-514) runs
the plan's GN_SENSE_* kinds: forward = NUFFT.forward(x * S), backward =
sum_c conj(S_c) NUFFT.backward(y)_c (keepdims), forward_backward =
espirit.backward(NUFFT.forward_backward(espirit.forward(x))).  Checked against the
same composition with separate broadcasts (rel-L2 <= 1e-6: only the FP32 rounding
order differs) and against the FP64 oracle (rel-L2 <= 1e-5), for 2-D and 3-D, both
FFT engines, both layouts, and coil counts whose combine does or does not fit one
row-pass block.
"""
import numpy as np
import pytest

from oracle import nufft_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    if hasattr(a, "cpu"):
        a = a.cpu().numpy()
    if hasattr(b, "cpu"):
        b = b.cpu().numpy()
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _case(modes, nc, M, seed):
    from mripy_amd import data
    rng = np.random.default_rng(seed)
    d = len(modes)
    coords = [(np.pi * (2 * rng.random(M) - 1)).astype(np.float32) for _ in range(d)]
    S = data.coil_maps(modes, nc, seed=seed)
    x = (rng.standard_normal(modes + (1,)) + 1j * rng.standard_normal(modes + (1,))).astype(np.complex64)
    y = (rng.standard_normal((M, nc)) + 1j * rng.standard_normal((M, nc))).astype(np.complex64)
    dcf = rng.random(M).astype(np.float32)
    return coords, S, x, y, dcf


CASES = [((48, 40), 3, "auto"), ((64, 256), 40, "auto"), ((40, 36), 5, "rocfft"), ((24, 20, 16), 4, "auto"),
         ((16, 12, 10), 3, "rocfft")]


@pytest.mark.parametrize("modes,nc,engine", CASES)
def test_fused_sense_matches_composition_and_oracle(modes, nc, engine):
    import torch
    from mripy_amd.operators import NUFFT2d_hip, NUFFT3d_hip, espirit, joint2operators
    print("case: SENSE", modes, nc, engine, flush=True)
    d = len(modes)
    M = 3000 if d == 2 else 5000
    coords, S, x, y, dcf = _case(modes, nc, M, seed=sum(modes) + nc)
    cls = NUFFT2d_hip if d == 2 else NUFFT3d_hip
    nuf = cls(modes, dcf, *coords, rat=2, nspread=3)
    St = torch.as_tensor(S).cuda()
    fused = joint2operators(espirit(St), nuf)
    plain = joint2operators(espirit(St), nuf, fused=False)
    xt, yt = torch.as_tensor(x).cuda(), torch.as_tensor(y).cuda()
    nuf._ensure(nc, xt.device).set_fft_engine(engine)
    outs_f = (fused.forward(xt), fused.backward(yt), fused.forward_backward(xt))
    outs_p = (plain.forward(xt), plain.backward(yt), plain.forward_backward(xt))
    assert outs_f[0].shape == (M, nc) and outs_f[1].shape == modes + (1,) and outs_f[2].shape == modes + (1,)
    for a, b in zip(outs_f, outs_p):
        assert rel(a, b) <= 1e-6, rel(a, b)
    # FP64 oracle of the composition
    p = O.Params(modes, rat=2, nspread=3)
    c64 = [c.astype(np.float64) for c in coords]
    S64 = S.astype(np.complex128)
    fwd = O.type2(c64, x.astype(np.complex128) * S64, p)
    bwd = np.sum(np.conj(S64) * O.type1(c64, y * dcf[:, None], p), axis=-1, keepdims=True)
    fb = np.sum(np.conj(S64) * O.type21(c64, x.astype(np.complex128) * S64, p, dcf), axis=-1, keepdims=True)
    for a, b in zip(outs_f, (fwd, bwd, fb)):
        assert rel(a, b) <= 1e-5, rel(a, b)


@pytest.mark.parametrize("layout", ["coil_last", "coil_major"])
def test_plan_sense_kinds_and_exact_adjoint(layout):
    import torch
    from mripy_amd.plan import Plan
    print("case: Plan SENSE kinds", layout, flush=True)
    modes, nc, M = (32, 24, 20), 5, 6000
    coords, S, x, y, dcf = _case(modes, nc, M, seed=7)
    P = Plan(modes, ncoils=nc, rat=2, nspread=3, layout=layout)
    P.set_points(coords)
    maps = S if layout == "coil_last" else np.moveaxis(S, -1, 0)
    P.set_sense(torch.as_tensor(np.ascontiguousarray(maps)).cuda())
    xt = torch.as_tensor(x[..., 0]).cuda()
    yt = torch.as_tensor(y if layout == "coil_last" else np.ascontiguousarray(y.T)).cuda()
    Ax = P.sense_type2(xt)
    AHy = P.sense_adjoint(yt)
    # <A x, y> = <x, A^H y>
    lhs = torch.vdot(Ax.reshape(-1).to(torch.complex128), yt.reshape(-1).to(torch.complex128))
    rhs = torch.vdot(xt.reshape(-1).to(torch.complex128), AHy.reshape(-1).to(torch.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5
    # against the plain kinds with the expansion / combine done by hand
    mt = torch.as_tensor(np.ascontiguousarray(maps)).cuda()
    coil_img = (xt[..., None] * mt) if layout == "coil_last" else (xt[None] * mt)
    assert rel(Ax, P.type2(coil_img)) <= 1e-6
    cimg = P.type1(yt)
    comb = (torch.conj(mt) * cimg).sum(-1 if layout == "coil_last" else 0)
    assert rel(P.sense_type1(yt), comb) <= 1e-6
    assert rel(P.sense_type21(xt), P.sense_type1(P.sense_type2(xt))) <= 1e-6
    with pytest.raises(Exception):
        Plan(modes, ncoils=nc, rat=2, nspread=3).set_points(coords).sense_type2(xt)
