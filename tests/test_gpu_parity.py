"""GPU parity: the HIP path through the C ABI against the FP64 oracle.

Tolerance: rel-L2 <= 1e-5 for complex64/FP32 results against the FP64 oracle on
the same (float32-representable) inputs.  Sizes are chosen so the NumPy/C
oracle finishes in seconds.
"""
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


def _plan(modes, nc, **kw):
    from mripy_amd.plan import Plan
    return Plan(modes, ncoils=nc, **kw)


def _rand_coords(rng, M, d, spread=1.0):
    return [(spread * np.pi * (2 * rng.random(M) - 1)).astype(np.float32) for _ in range(d)]


def test_cfg1_radial_2d_forward_adjoint():
    from mripy_amd import data
    print("case: cfg1 2-D radial 128^2, 128x256", flush=True)
    cfg = data.config_2d_radial()
    x, y = cfg["coords"]
    p = O.Params((128, 128), rat=2, nspread=3)
    P = _plan((128, 128), 1, rat=2, nspread=3)
    P.set_points([x, y])
    img = cfg["image"][..., None]
    yg = P.type2(img)[:, 0]
    yo = O.type2([x, y], cfg["image"], p)
    assert rel(yg, yo) <= TOL, rel(yg, yo)
    c = cfg["samples"][:, None]
    ig = P.type1(c)[..., 0]
    io = O.type1([x, y], cfg["samples"], p)
    assert rel(ig, io) <= TOL, rel(ig, io)


@pytest.mark.parametrize("ns,rat", [(3, 2), (5, 2), (6, 3)])
def test_random_2d_multicoil(ns, rat):
    print("case: 2-D random 40x36 ns=%d rat=%d" % (ns, rat), flush=True)
    rng = np.random.default_rng(1)
    M, nc = 5000, 3
    x, y = _rand_coords(rng, M, 2, spread=3.0)  # several periods: exercises the wrap
    p = O.Params((40, 36), rat=rat, nspread=ns)
    P = _plan((40, 36), nc, rat=rat, nspread=ns)
    P.set_points([x, y])
    img = (rng.standard_normal((40, 36, nc)) + 1j * rng.standard_normal((40, 36, nc))).astype(np.complex64)
    c = (rng.standard_normal((M, nc)) + 1j * rng.standard_normal((M, nc))).astype(np.complex64)
    assert rel(P.type2(img), O.type2([x, y], img, p)) <= TOL
    assert rel(P.type1(c), O.type1([x, y], c, p)) <= TOL


@pytest.mark.parametrize("modes,ns", [((32, 32, 32), 3), ((24, 20, 18), 5), ((17, 19, 21), 3)])
def test_random_3d_multicoil(modes, ns):
    print("case: 3-D random", modes, "ns", ns, flush=True)
    rng = np.random.default_rng(2)
    M, nc = 20000, 3
    coords = _rand_coords(rng, M, 3)
    p = O.Params(modes, rat=2, nspread=ns)
    P = _plan(modes, nc, rat=2, nspread=ns)
    P.set_points(coords)
    img = (rng.standard_normal(modes + (nc,)) + 1j * rng.standard_normal(modes + (nc,))).astype(np.complex64)
    c = (rng.standard_normal((M, nc)) + 1j * rng.standard_normal((M, nc))).astype(np.complex64)
    assert rel(P.type2(img), C.type2(coords, img, p)) <= TOL
    assert rel(P.type1(c), C.type1(coords, c, p)) <= TOL


def test_iflag_df_dcf_3d():
    print("case: 3-D iflag=-1 df=2 dcf", flush=True)
    rng = np.random.default_rng(3)
    modes, M, nc = (16, 16, 16), 8000, 2
    coords = _rand_coords(rng, M, 3)
    dcf = rng.random(M).astype(np.float32)
    p = O.Params(modes, rat=2, nspread=3, df=2.0, iflag=-1)
    P = _plan(modes, nc, rat=2, nspread=3, iflag=-1)
    P.set_points(coords, df=2.0, dcf=dcf)
    img = (rng.standard_normal(modes + (nc,)) + 1j * rng.standard_normal(modes + (nc,))).astype(np.complex64)
    c = (rng.standard_normal((M, nc)) + 1j * rng.standard_normal((M, nc))).astype(np.complex64)
    assert rel(P.type2(img), C.type2(coords, img, p)) <= TOL
    assert rel(P.type1(c), C.type1(coords, c * dcf[:, None], p)) <= TOL
    assert rel(P.type21(img), O.type21(coords, img, p, dcf)) <= TOL


def test_adjointness_and_determinism_3d():
    print("case: 3-D adjointness + bitwise determinism", flush=True)
    import torch
    rng = np.random.default_rng(4)
    modes, M, nc = (32, 32, 32), 30000, 2
    coords = _rand_coords(rng, M, 3, spread=0.3)  # dense centre: many samples per tile
    P = _plan(modes, nc, rat=2, nspread=3)
    P.set_points(coords)
    f = torch.randn(modes + (nc,), dtype=torch.complex64, device="cuda")
    y = torch.randn((M, nc), dtype=torch.complex64, device="cuda")
    Af = P.type2(f)
    AHy = P.adjoint(y)
    lhs = torch.vdot(Af.reshape(-1).to(torch.complex128), y.reshape(-1).to(torch.complex128))
    rhs = torch.vdot(f.reshape(-1).to(torch.complex128), AHy.reshape(-1).to(torch.complex128))
    assert abs(lhs - rhs) / abs(lhs) < 1e-5
    AHy2 = P.adjoint(y)
    assert torch.equal(AHy, AHy2)
    assert torch.equal(P.type2(f), Af)


@pytest.mark.parametrize("dim,nc", [(3, 1), (3, 4), (3, 8), (3, 12), (2, 2), (2, 8), (2, 19)])
def test_matrix_core_and_vector_spreads_agree(dim, nc):
    """Block-aligned spread tiles (multiples of 4 per axis in 3-D, of 8 in 2-D) run the
    matrix-core spread; other tiles the vector spread.  Both match the oracle, and each
    other to FP32 roundoff, for every coil-group width (8, 4, 2, 1 and a partial group)."""
    print("case: spread kernels", dim, "D, nc", nc, flush=True)
    rng = np.random.default_rng(10 + nc)
    modes = (24, 20, 16) if dim == 3 else (40, 36)
    M = 12000 if dim == 3 else 6000
    coords = _rand_coords(rng, M, dim)
    p = O.Params(modes, rat=2, nspread=3)
    c = (rng.standard_normal((M, nc)) + 1j * rng.standard_normal((M, nc))).astype(np.complex64)
    out = []
    for tiles in (((4, 4, 8), (4, 4, 6)) if dim == 3 else ((8, 16), (8, 12))):
        P = _plan(modes, nc, rat=2, nspread=3)
        P.set_tiles(tiles)
        P.set_points(coords)
        out.append(P.type1(c))
    ref = C.type1(coords, c, p)
    assert rel(out[0], ref) <= TOL
    assert rel(out[1], ref) <= TOL
    assert rel(out[0], out[1]) <= 1e-6


def test_dense_center_split_tiles_3d():
    """Samples packed into a few cells: tiles hold thousands of pairs, so the spread
    splits them over several work items (partial tiles + the Kahan reduce)."""
    print("case: dense centre, split spread tiles", flush=True)
    rng = np.random.default_rng(12)
    modes, nc, M = (32, 32, 32), 2, 40000
    coords = [(0.05 * rng.standard_normal(M)).astype(np.float32) for _ in range(3)]
    p = O.Params(modes, rat=2, nspread=3)
    P = _plan(modes, nc, rat=2, nspread=3)
    P.set_points(coords)
    img = (rng.standard_normal(modes + (nc,)) + 1j * rng.standard_normal(modes + (nc,))).astype(np.complex64)
    c = (rng.standard_normal((M, nc)) + 1j * rng.standard_normal((M, nc))).astype(np.complex64)
    assert rel(P.type1(c), C.type1(coords, c, p)) <= TOL
    assert rel(P.type2(img), C.type2(coords, img, p)) <= TOL


@pytest.mark.parametrize("dim", [2, 3])
def test_tiles_larger_than_one_workgroup(dim):
    """Spread tiles of more than 8 blocks (32x32 in 2-D, 8x8x16 in 3-D) span several
    workgroups and run the vector spread; they match the oracle like the defaults."""
    print("case: large spread tiles", dim, flush=True)
    rng = np.random.default_rng(30 + dim)
    modes = (48, 40) if dim == 2 else (24, 20, 32)
    M, nc = (8000, 3) if dim == 2 else (12000, 2)
    coords = _rand_coords(rng, M, dim)
    p = O.Params(modes, rat=2, nspread=3)
    P = _plan(modes, nc, rat=2, nspread=3)
    P.set_tiles((32, 32) if dim == 2 else (8, 8, 16))
    P.set_points(coords)
    c = (rng.standard_normal((M, nc)) + 1j * rng.standard_normal((M, nc))).astype(np.complex64)
    assert rel(P.type1(c), C.type1(coords, c, p)) <= TOL


@pytest.mark.parametrize("dim,modes,nc", [(3, (32, 32, 32), 3), (2, (64, 48), 2), (2, (600, 120), 2)])
def test_unwritten_grid_cells_stay_out_of_results(dim, modes, nc):
    """Samples inside a small ball leave most spread tiles empty.  With the pruned FFT
    passes the spread leaves empty tiles unwritten and the first adjoint pass reads
    zeros for them.  Alternating type 2, type 1 and the rocFFT engine on one plan
    leaves stale values of other transforms in those cells, and none may reach a
    result.  (600, 120)
    has 150 spread tiles along the leading axis, so one mask bit covers three tiles
    and the spread writes zeros for the empty tiles of occupied groups."""
    print("case: unwritten grid cells", dim, modes, nc, flush=True)
    rng = np.random.default_rng(40 + len(modes) + modes[0])
    M = 6000
    u = rng.standard_normal((M, dim))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    r = 0.35 * np.pi * rng.random(M) ** (1.0 / dim)
    coords = [(u[:, a] * r).astype(np.float32) for a in range(dim)]
    p = O.Params(modes, rat=2, nspread=3)
    P = _plan(modes, nc, rat=2, nspread=3)
    P.set_points(coords)
    img = (rng.standard_normal(modes + (nc,)) + 1j * rng.standard_normal(modes + (nc,))).astype(np.complex64)
    c = (rng.standard_normal((M, nc)) + 1j * rng.standard_normal((M, nc))).astype(np.complex64)
    ref2 = C.type2(coords, img, p)
    ref1 = C.type1(coords, c, p)
    for engine in ("auto", "auto", "rocfft", "auto"):
        P.set_fft_engine(engine)
        e2 = rel(P.type2(img), ref2)
        e1 = rel(P.type1(c), ref1)
        assert e2 <= TOL and e1 <= TOL, (engine, e2, e1)
