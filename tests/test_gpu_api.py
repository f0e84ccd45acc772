"""GPU tests of the drop-in API: counting sort, functional API, operator classes,
layouts, edge cases.  Tolerance rel-L2 <= 1e-5 against the FP64 oracle."""
import numpy as np
import pytest

from oracle import coracle as C
from oracle import nufft_oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def rnd_coords(rng, M, d, span=1.0):
    return [(span * np.pi * (2 * rng.random(M) - 1)).astype(np.float32) for _ in range(d)]


def rnd_c(rng, shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


# ------------------------------------------------------------ counting sort (N1)

@pytest.mark.parametrize("dim", [2, 3])
def test_counting_sort_is_stable_argsort_of_tile_ids(dim):
    from mripy_amd.plan import Plan
    print("case: counting sort dim", dim, flush=True)
    rng = np.random.default_rng(20 + dim)
    M = 50000
    modes = (40, 36) if dim == 2 else (20, 18, 22)
    coords = rnd_coords(rng, M, dim, span=0.4)     # dense centre: many equal keys
    P = Plan(modes, ncoils=1, rat=2, nspread=3)
    P.set_points(coords)
    ns = 3
    gt = P.gather_tiles
    key = np.zeros(M, np.int64)
    for a in range(dim):
        nf = P.nf[a]
        _, m = O.cell_index(coords[a].astype(np.float64), nf)
        s = np.mod(m - ns, nf)
        nt = -(-nf // gt[a])
        key = key * nt + s // gt[a]
    want = np.argsort(key, kind="stable")
    assert (P.sample_order() == want).all()


# ------------------------------------------------------------ functional API

def test_functional_2d_matches_reference_semantics():
    from mripy_amd import functional as F
    print("case: functional 2-D", flush=True)
    rng = np.random.default_rng(1)
    M = 4000
    x, y = rnd_coords(rng, M, 2, span=5.0)
    c = rnd_c(rng, (M, 2)).astype(np.complex128)
    Fk = rnd_c(rng, (20, 24, 2)).astype(np.complex128)
    for eps in (1e-5, 1e-12):
        p = O.Params((20, 24), eps=eps, df=2.0, iflag=-1)
        a = F.nufft2d1_gaussker(x, y, c, 20, 24, df=2.0, eps=eps, iflag=-1)
        assert a.dtype == np.complex128 and a.shape == (20, 24, 2)
        assert rel(a, O.type1([x, y], c, p)) <= TOL
        b = F.nufft2d2_gaussker(x, y, Fk, 20, 24, df=2.0, eps=eps, iflag=-1)
        assert b.shape == (M, 2)
        assert rel(b, O.type2([x, y], Fk, p)) <= TOL
        t = F.nufft2d21_gaussker_cuda(x, y, Fk, 20, 24, df=2.0, eps=eps, iflag=-1)
        assert rel(t, O.type21([x, y], Fk, p)) <= TOL


def test_functional_3d_single_and_multicoil():
    from mripy_amd import functional as F
    print("case: functional 3-D", flush=True)
    rng = np.random.default_rng(2)
    M = 6000
    x, y, z = rnd_coords(rng, M, 3)
    c1 = rnd_c(rng, (M,))
    Fk = rnd_c(rng, (12, 10, 14, 3))
    dcf = rng.random(M).astype(np.float32)
    p = O.Params((12, 10, 14), eps=1e-5, max_nspread=6)
    assert rel(F.nufft3d1_gaussker_cuda(x, y, z, c1, 12, 10, 14, eps=1e-5), O.type1([x, y, z], c1, p)) <= TOL
    assert rel(F.nufft3d2_gaussker(x, y, z, Fk, 12, 10, 14, eps=1e-5), C.type2([x, y, z], Fk, p)) <= TOL
    assert rel(F.nufft3d21_gaussker(x, y, z, Fk, 12, 10, 14, dcf=dcf, eps=1e-5),
               O.type21([x, y, z], Fk, p, dcf)) <= TOL


# ------------------------------------------------------------ operator classes

def test_nufft3d_operator_methods():
    from mripy_amd.operators import NUFFT3d_cuda
    print("case: NUFFT3d_cuda operator", flush=True)
    rng = np.random.default_rng(3)
    M, nc = 8000, 2
    x, y, z = rnd_coords(rng, M, 3)
    dcf = rng.random(M).astype(np.float32)
    op = NUFFT3d_cuda((16, 12, 14), dcf, x, y, z)
    p = O.Params((16, 12, 14), eps=1e-5, max_nspread=6)
    im = rnd_c(rng, (16, 12, 14, nc))
    kd = rnd_c(rng, (M, nc))
    assert rel(op.forward(im), C.type2([x, y, z], im, p)) <= TOL
    assert rel(op.backward(kd), C.type1([x, y, z], kd * dcf[:, None], p)) <= TOL
    assert rel(op.forward_backward(im), O.type21([x, y, z], im, p, dcf)) <= TOL
    assert rel(op.adjoint(kd), C.type1([x, y, z], kd, p) * M) <= TOL
    # density_weighting squeezes like the reference
    assert op.density_weighting(kd, dcf).shape == (M, nc)
    # set_params: the next call runs with the new window (and a rebuilt plan)
    assert (op.rat, op.nspread) == (2, 5)
    op.set_params(nspread=3)
    p3 = O.Params((16, 12, 14), rat=2, nspread=3)
    assert rel(op.forward(im), C.type2([x, y, z], im, p3)) <= TOL
    with pytest.raises(ValueError):
        op.set_params(nspread=9)
    kn = op.normalize_set_ktraj(np.stack([x, y, z]))
    assert np.allclose(max(np.max(v) for v in kn), np.pi)


def test_sense_joint_operator_torch_device():
    import torch
    from mripy_amd import data
    from mripy_amd.operators import NUFFT2d_hip, espirit, joint2operators
    print("case: espirit + joint2operators on torch", flush=True)
    rng = np.random.default_rng(4)
    x, y = data.golden_angle_radial_2d(40, 64)
    S = data.coil_maps((32, 32), 4, seed=3)
    nft = NUFFT2d_hip((32, 32), np.ones(x.shape[0], np.float32), x, y, rat=2, nspread=3)
    A = joint2operators(espirit(torch.as_tensor(S).cuda()), nft)
    img = torch.as_tensor(rnd_c(rng, (32, 32))).cuda()[..., None]
    k = A.forward(img)
    assert k.is_cuda and k.shape == (x.shape[0], 4)
    p = O.Params((32, 32), rat=2, nspread=3)
    ref = O.type2([x, y], img.cpu().numpy() * S, p)
    assert rel(k.cpu().numpy(), ref) <= TOL
    back = A.backward(k)
    ref_b = np.sum(np.conj(S) * O.type1([x, y], ref, p), axis=-1, keepdims=True)
    assert rel(back.cpu().numpy(), ref_b) <= TOL
    fb = A.forward_backward(img)
    assert rel(fb.cpu().numpy(), ref_b) <= TOL


# ------------------------------------------------------------ layouts and edges

def test_coil_major_layout_equals_coil_last():
    import torch
    from mripy_amd.plan import Plan
    print("case: coil-major layout", flush=True)
    rng = np.random.default_rng(5)
    M, nc, modes = 5000, 3, (14, 16, 12)
    cs = rnd_coords(rng, M, 3)
    a = Plan(modes, ncoils=nc, rat=2, nspread=3).set_points(cs)
    b = Plan(modes, ncoils=nc, rat=2, nspread=3, layout="coil_major").set_points(cs)
    im = torch.as_tensor(rnd_c(rng, modes + (nc,))).cuda()
    y = torch.as_tensor(rnd_c(rng, (M, nc))).cuda()
    assert torch.equal(a.type2(im), b.type2(im.permute(3, 0, 1, 2).contiguous()).t())
    assert torch.equal(a.type1(y), b.type1(y.t().contiguous()).permute(1, 2, 3, 0))


def test_edge_cases():
    import torch
    from mripy_amd.plan import Plan
    print("case: edge cases", flush=True)
    P = Plan((8, 8), ncoils=1, rat=2, nspread=3)
    with pytest.raises(RuntimeError):
        P.type2(np.zeros((8, 8, 1), np.complex64))       # before set_points
    P.set_points([np.zeros(0, np.float32), np.zeros(0, np.float32)])
    assert P.type2(np.ones((8, 8, 1), np.complex64)).shape == (0, 1)
    with pytest.raises(ValueError):
        P.type1(np.zeros((0, 1), np.complex64))            # 1/M undefined, like the reference
    assert np.abs(P.adjoint(np.zeros((0, 1), np.complex64))).max() == 0
    with pytest.raises(ValueError):
        P.set_points([np.array([0.1, np.nan], np.float32), np.zeros(2, np.float32)])
    # one sample at the origin: type 1 is the constant image
    P.set_points([np.zeros(1, np.float32), np.zeros(1, np.float32)])
    img = P.type1(np.ones((1, 1), np.complex64))
    p = O.Params((8, 8), rat=2, nspread=3)
    ref = O.type1([np.zeros(1), np.zeros(1)], np.ones(1), p)
    assert np.abs(ref - 1).max() < 1e-2          # nspread 3 kernel error of the constant image
    assert rel(img[..., 0], ref) <= TOL
    # cell boundaries, large and negative coordinates (periodic)
    h = 2 * np.pi / 16
    xs = np.array([0, h, -h, 3 * h, 2 * np.pi - h, 1000.0, -1000.0, np.pi], np.float64)
    P = Plan((8, 8), ncoils=1, rat=2, nspread=3).set_points([xs, xs[::-1].copy()])
    f = (np.arange(64).reshape(8, 8) * (1 + 1j)).astype(np.complex64)
    assert rel(P.type2(f[..., None])[:, 0], O.type2([xs, xs[::-1]], f, p)) <= TOL
    del torch


def test_bitwise_reproducible_across_plans():
    import torch
    from mripy_amd.plan import Plan
    print("case: reproducibility", flush=True)
    rng = np.random.default_rng(6)
    cs = rnd_coords(rng, 40000, 3, span=0.2)
    y = torch.as_tensor(rnd_c(rng, (40000, 2))).cuda()
    outs = []
    for _ in range(2):
        P = Plan((24, 24, 24), ncoils=2, rat=2, nspread=3).set_points(cs)
        outs.append(P.adjoint(y))
    assert torch.equal(outs[0], outs[1])


def test_operator_sees_in_place_trajectory_and_dcf_updates():
    """The operator re-bins when its trajectory or density weights change in place
    (numpy arrays through a content hash, torch tensors through their version
    counter), as the reference does by passing them to the driver on every call."""
    import torch
    from mripy_amd.operators import NUFFT2d_hip
    print("case: operator in-place updates", flush=True)
    rng = np.random.default_rng(9)
    modes, M = (32, 28), 2000
    x, y = [(np.pi * (2 * rng.random(M) - 1)).astype(np.float32) for _ in range(2)]
    dcf = rng.random(M).astype(np.float32)
    img = (rng.standard_normal(modes + (2,)) + 1j * rng.standard_normal(modes + (2,))).astype(np.complex64)
    kd = (rng.standard_normal((M, 2)) + 1j * rng.standard_normal((M, 2))).astype(np.complex64)
    op = NUFFT2d_hip(modes, dcf, x, y, rat=2, nspread=3)
    op.forward(img)
    op.ktrajx *= 0.5                     # numpy, in place
    fresh = NUFFT2d_hip(modes, dcf, op.ktrajx.copy(), y, rat=2, nspread=3)
    assert np.array_equal(op.forward(img), fresh.forward(img))
    op.dcf[:] = 1.0                      # numpy weights, in place
    fresh = NUFFT2d_hip(modes, np.ones(M, np.float32), op.ktrajx.copy(), y, rat=2, nspread=3)
    assert np.array_equal(op.backward(kd), fresh.backward(kd))
    # torch trajectory on the device, updated in place through a view
    xt = torch.as_tensor(x).cuda()
    yt = torch.as_tensor(y).cuda()
    opt = NUFFT2d_hip(modes, None, xt, yt, rat=2, nspread=3)
    it = torch.as_tensor(img).cuda()
    a = opt.forward(it).clone()
    xt.mul_(0.25)
    b = opt.forward(it)
    fresh = NUFFT2d_hip(modes, None, xt.clone(), yt, rat=2, nspread=3)
    assert torch.equal(b, fresh.forward(it)) and not torch.equal(a, b)
