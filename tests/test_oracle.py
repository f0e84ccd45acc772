"""CPU tests that pin the FP64 oracle (no GPU needed).

Pins: the reference's own known-answer test (fft/nufft_test_func.py:15-66:
RandomState(0) inputs, type 1 vs the brute-force NUDFT at rtol=atol=1e-2) and
far tighter closed forms -- brute-force NUDFT at eps=1e-15, exact adjointness of
type 2 and type 1, brute-force type-2 DFT, fast (E1 E2^k E3) vs direct weights,
the Cartesian-lattice special case against numpy.fft, 2 pi periodicity, a
single-sample closed form, linearity, the C restatement against the NumPy one,
and a planted mistake that the NUDFT check must catch.
"""
import numpy as np
import pytest

from oracle import coracle as C
from oracle import nufft_oracle as O


def rel(a, b):
    return float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


# --------------------------------------------------------------------- params

@pytest.mark.parametrize("modes,eps,mx,expect", [
    ((128, 128, 128), 1e-3, 6, (3, 2, (256, 256, 256))),
    ((128, 128, 128), 1e-5, 6, (5, 2, (256, 256, 256))),
    ((128, 128, 128), 5e-6, 6, (6, 2, (256, 256, 256))),
    ((16, 16), 1e-15, None, (14, 3, (48, 48))),
    ((32, 32, 16), 1e-15, 6, (6, 3, (96, 96, 48))),
    ((4, 4), 1e-5, None, (5, 2, (10, 10))),           # nf = max(rat*m, 2*ns)
])
def test_grid_params_table(modes, eps, mx, expect):
    ns, nf, tau, rat = O.grid_params(modes, eps=eps, max_nspread=mx)
    assert (ns, rat, nf) == expect
    assert tau == pytest.approx(np.pi * ns / (rat * (rat - 0.5)) / modes[0] ** 2, rel=1e-15)


def test_eps_range():
    for bad in (0.2, 1e-40):
        with pytest.raises(ValueError):
            O.grid_params((8, 8), eps=bad)


def test_freqs():
    k1, k2 = O.freqs((5, 4))
    assert k1[:, 0].tolist() == [-2, -1, 0, 1, 2]
    assert k2[0].tolist() == [-2, -1, 0, 1]


# --------------------------------------------------------------- type 1 KATs

@pytest.mark.parametrize("df", [1.0, 2.0])
@pytest.mark.parametrize("iflag", [1, -1])
def test_kat_2d1_reference(df, iflag):
    # compare_nufft2d1 inputs and tolerance (fft/nufft_test_func.py:31-47), (16,16,25000)
    rng = np.random.RandomState(0)
    mc = 25000
    x = 100 * rng.rand(mc)
    y = 100 * rng.rand(mc)
    c = np.sin(2 * x) + 1j * np.cos(2 * y)
    F1 = O.nudft_type1([x, y], c, (16, 16), df=df, iflag=iflag)
    F2 = O.nufft2d1_gaussker(x, y, c, 16, 16, df=df, iflag=iflag)
    assert np.allclose(F1, F2, rtol=1e-2, atol=1e-2)
    assert rel(F2, F1) < 1e-11          # eps=1e-15 default: far tighter than the reference test


@pytest.mark.parametrize("iflag", [1, -1])
def test_kat_3d1_reference(iflag):
    # compare_nufft3d1 (fft/nufft_test_func.py:49-66) at a CPU-friendly size
    rng = np.random.RandomState(0)
    mc = 4096
    x = 100 * rng.rand(mc)
    y = 100 * rng.rand(mc)
    z = 100 * rng.rand(mc)
    c = np.sin(2 * x) + 1j * np.cos(2 * y)
    F1 = O.nudft_type1([x, y, z], c, (16, 16, 8), iflag=iflag)
    F2 = O.nufft3d1_gaussker(x, y, z, c, 16, 16, 8, iflag=iflag)
    assert np.allclose(F1, F2, rtol=1e-2, atol=1e-2)
    # tau uses ms for every axis (fft/nufft_func.py:248), so a short third axis is less
    # accurate; on a cubic grid the clamped nspread 6 at rat 3 gives ~1e-7
    F1 = O.nudft_type1([x, y, z], c, (12, 12, 12), iflag=iflag)
    F2 = O.nufft3d1_gaussker(x, y, z, c, 12, 12, 12, iflag=iflag)
    assert rel(F2, F1) < 1e-6


@pytest.mark.parametrize("eps,tol", [(1e-3, 1e-2), (1e-5, 5e-4), (1e-8, 2e-6)])
def test_accuracy_matches_eps(eps, tol):
    rng = np.random.RandomState(1)
    x, y = 2 * np.pi * rng.rand(2, 3000)
    c = rng.randn(3000) + 1j * rng.randn(3000)
    F1 = O.nudft_type1([x, y], c, (24, 20))
    F2 = O.nufft2d1_gaussker(x, y, c, 24, 20, eps=eps)
    assert rel(F2, F1) < tol


def test_planted_sign_mistake_is_caught():
    # a wrong sign convention in the oracle must fail the brute-force check
    rng = np.random.RandomState(0)
    x, y = 100 * rng.rand(2, 2000)
    c = np.sin(2 * x) + 1j * np.cos(2 * y)
    F1 = O.nudft_type1([x, y], c, (16, 16), iflag=1)
    wrong = O.nufft2d1_gaussker(x, y, c, 16, 16, iflag=-1)
    assert not np.allclose(F1, wrong, rtol=1e-2, atol=1e-2)


# --------------------------------------------------------------- type 2 / 21

def test_type2_bruteforce_and_adjointness_2d():
    rng = np.random.RandomState(2)
    M = 3000
    x, y = 20 * rng.rand(2, M) - 10
    F = rng.randn(12, 10) + 1j * rng.randn(12, 10)
    c = rng.randn(M) + 1j * rng.randn(M)
    p = O.Params((12, 10), eps=1e-15)
    A_F = O.type2([x, y], F, p)
    assert rel(A_F, O.nudft_type2([x, y], F, (12, 10))) < 1e-11
    AH_c = O.type1([x, y], c, p) * M
    lhs = np.vdot(c, A_F)
    rhs = np.vdot(AH_c, F)
    assert abs(lhs - rhs) / abs(lhs) < 1e-13


@pytest.mark.parametrize("iflag,df", [(-1, 1.0), (-1, 1.5), (1, 1.5)])
def test_type2_bruteforce_sign_and_df(iflag, df):
    # the iflag < 0 branch (inverse FFT times G, fft/nufft_func.py:768-771) and the df
    # scaling of the coordinates against the brute-force type-2 sum at eps=1e-15
    rng = np.random.RandomState(11)
    M = 2500
    x, y = 20 * rng.rand(2, M) - 10
    F = rng.randn(12, 10) + 1j * rng.randn(12, 10)
    p = O.Params((12, 10), eps=1e-15, iflag=iflag, df=df)
    got = O.type2([x, y], F, p)
    assert rel(got, O.nudft_type2([x, y], F, (12, 10), df=df, iflag=iflag)) < 1e-11
    # 3-D as well, through the NumPy and the C restatements
    xyz = list(12 * rng.rand(3, 1500) - 6)
    p3 = O.Params((8, 8, 8), eps=1e-15, iflag=iflag, df=df, max_nspread=6)
    F3 = rng.randn(8, 8, 8) + 1j * rng.randn(8, 8, 8)
    want = O.nudft_type2(xyz, F3, (8, 8, 8), df=df, iflag=iflag)
    assert rel(O.type2(xyz, F3, p3), want) < 1e-6
    assert rel(C.type2(xyz, F3, p3), want) < 1e-6


def test_adjointness_3d_multicoil_rat2():
    rng = np.random.RandomState(3)
    M, nc = 2000, 2
    xyz = 2 * np.pi * rng.rand(3, M)
    F = rng.randn(8, 6, 10, nc) + 1j * rng.randn(8, 6, 10, nc)
    c = rng.randn(M, nc) + 1j * rng.randn(M, nc)
    p = O.Params((8, 6, 10), rat=2, nspread=3)
    A_F = O.type2(list(xyz), F, p)
    AH_c = O.adjoint_exact(list(xyz), c, p)
    assert abs(np.vdot(c, A_F) - np.vdot(AH_c, F)) / abs(np.vdot(c, A_F)) < 1e-13


def test_type21_is_composition():
    rng = np.random.RandomState(4)
    M = 1500
    xyz = 2 * np.pi * rng.rand(3, M)
    F = rng.randn(6, 6, 6) + 1j * rng.randn(6, 6, 6)
    dcf = rng.rand(M)
    p = O.Params((6, 6, 6), rat=2, nspread=3)
    got = O.type21(list(xyz), F, p, dcf)
    want = O.type1(list(xyz), O.type2(list(xyz), F, p) * dcf, p)
    assert rel(got, want) < 1e-14
    # the C restatement of build_grid_3d21 agrees
    coords = [v * p.df for v in xyz]
    F2 = F * O.deapod(p.modes, p.tau)
    grid = np.fft.fftn(O.place(F2[..., None], p.nf), axes=(0, 1, 2))
    g21 = C.interp_spread3d(coords, grid, dcf, p.nf, p.ns, p.tau)
    ref = O.spread(coords, O.gather(coords, grid, p.nf, p.ns, p.tau) * dcf[:, None], p.nf, p.ns, p.tau)
    assert rel(g21, ref) < 1e-13


# ------------------------------------------------------------- special cases

def test_fast_and_direct_weights_agree():
    rng = np.random.RandomState(5)
    x = 100 * rng.rand(5000) - 50
    for nf, ns in [(32, 3), (64, 5), (48, 14)]:
        tau = np.pi * ns / (2 * 1.5) / (nf // 2) ** 2
        i1, w1 = O.window_direct(x, nf, ns, tau)
        i2, w2 = O.window_fast(x, nf, ns, tau)
        assert (i1 == i2).all()
        assert np.abs(w1 - w2).max() < 1e-12 * max(1.0, np.abs(w1).max())


def test_cartesian_lattice_reduces_to_fft():
    # samples on x = 2 pi n/m: type 1 is (1/M) times an inverse DFT with the k shift
    m = 16
    n = np.arange(m)
    X, Y = np.meshgrid(2 * np.pi * n / m, 2 * np.pi * n / m, indexing="ij")
    rng = np.random.RandomState(6)
    c = rng.randn(m, m) + 1j * rng.randn(m, m)
    F = O.nufft2d1_gaussker(X.ravel(), Y.ravel(), c.ravel(), m, m, eps=1e-12)
    # sum_j c_j e^{i k x_j} with k = -m/2..m/2-1 == m^2 ifft2(c) rolled by m/2
    ref = np.fft.fftshift(np.fft.ifft2(c)) * m * m / (m * m)
    assert rel(F, ref) < 1e-9


def test_periodicity():
    rng = np.random.RandomState(7)
    x, y = 2 * np.pi * rng.rand(2, 800)
    c = rng.randn(800) + 1j * rng.randn(800)
    a = O.nufft2d1_gaussker(x, y, c, 10, 12, eps=1e-10)
    b = O.nufft2d1_gaussker(x + 4 * np.pi, y - 2 * np.pi, c, 10, 12, eps=1e-10)
    assert rel(b, a) < 1e-12


def test_single_sample_at_origin_gives_constant_image():
    F = O.nufft3d1_gaussker(np.zeros(1), np.zeros(1), np.zeros(1), np.ones(1), 6, 6, 6, eps=1e-12)
    assert np.abs(F - 1.0).max() < 1e-6  # 3-D clamps nspread to 6


def test_linearity():
    rng = np.random.RandomState(8)
    x, y = 2 * np.pi * rng.rand(2, 500)
    c1 = rng.randn(500) + 1j * rng.randn(500)
    c2 = rng.randn(500) + 1j * rng.randn(500)
    p = O.Params((9, 7), rat=2, nspread=3)
    a = O.type1([x, y], 2 * c1 - 3j * c2, p)
    b = 2 * O.type1([x, y], c1, p) - 3j * O.type1([x, y], c2, p)
    assert rel(a, b) < 1e-13


# ------------------------------------------------------ C restatement parity

@pytest.mark.parametrize("dim", [2, 3])
def test_c_oracle_matches_numpy_oracle(dim):
    rng = np.random.RandomState(9)
    M, nc = 2500, 3
    coords = list(100 * rng.rand(dim, M) - 50)
    modes = (12, 10) if dim == 2 else (10, 8, 12)
    p = O.Params(modes, eps=1e-5, max_nspread=6 if dim == 3 else None)
    c = rng.randn(M, nc) + 1j * rng.randn(M, nc)
    F = rng.randn(*modes, nc) + 1j * rng.randn(*modes, nc)
    assert rel(C.type1(coords, c, p), O.type1(coords, c, p)) < 1e-13
    assert rel(C.type2(coords, F, p), O.type2(coords, F, p)) < 1e-13


def test_cell_index_python_semantics():
    # x % 2pi and xi // h exactly as Python floats do, including exact multiples
    nf = 256
    h = 2 * np.pi / nf
    xs = np.array([0.0, -0.0, h, 2 * h, -h, 2 * np.pi, -2 * np.pi, 3 * np.pi, np.nextafter(2 * np.pi, 0),
                   1e-300, -1e-300, 100.0, -100.0])
    xi, m = O.cell_index(xs, nf)
    for v, a, b in zip(xs, xi, m):
        r = v % (2 * np.pi)
        assert a == r
        assert b == 1 + int(r // h)
    mc, _ = C.cell_index(xs, nf)
    assert (mc == m).all()


def test_pipe_menon_uniform_on_a_lattice_and_radial_ramp():
    # samples on a Cartesian lattice (a quarter period off the grid cells, so that every
    # window covers the same taps) have equal density: every weight is 1
    m = 16
    n = np.arange(m) - m // 2 + 0.25
    X, Y = np.meshgrid(2 * np.pi * n / m, 2 * np.pi * n / m, indexing="ij")
    w = O.pipe_menon([X.ravel(), Y.ravel()], (m, m), niter=5)
    assert np.abs(w - 1).max() < 1e-12
    # a radial trajectory is dense at the centre: weights grow with |k| away from it
    th = np.arange(64) * np.pi / 64
    r = (np.arange(64) - 32) / 64
    x = (2 * np.pi * r[None, :] * np.cos(th)[:, None]).ravel()
    y = (2 * np.pi * r[None, :] * np.sin(th)[:, None]).ravel()
    w = O.pipe_menon([x, y], (32, 32), niter=10)
    rad = np.hypot(x, y)
    inner = w[(rad > 0.3) & (rad < 0.6)].mean()
    outer = w[(rad > 1.8) & (rad < 2.4)].mean()
    assert outer > 2.5 * inner
