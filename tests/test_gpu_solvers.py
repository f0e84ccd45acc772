"""GPU tests of the reconstruction callers (mripy_amd.solvers) and their device
primitives, against the FP64 solver oracle.

* FP64 reductions (gn_dot_f64, gn_sqdist_f64) vs NumPy FP64 on the same vectors:
  relative error <= 1e-12 (different summation order only).
* TV prox kernel (gn_tv_prox) vs the oracle's prox_tv2d_r and prox_tv3d: <= 1e-12 for complex128
  data, <= 1e-6 for complex64 data (FP64 inside, rounded on output); the two-level
  step image against the closed-form TV-denoising solution.
* CG, ADMM-TV, IST/FISTA on a 2-D spiral SENSE problem vs the same algorithms in
  FP64: rel-L2 <= 1e-4.  The solvers amplify the FP32 roundoff of every operator
  application (~3e-7 relative): an FP64 run with that much noise injected per
  operator call ends 1.5e-5 (CG) and 2.4e-6 (ADMM) from the clean run.
* The full 256^2 x 32-coil spiral workload against committed FP64 oracle runs
  (tests/golden/make_solvers.py), same bound.
"""
import os

import numpy as np
import pytest

from oracle import nufft_oracle as O
from oracle import solvers_oracle as SO

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-4


def rel(a, b):
    if hasattr(a, "cpu"):
        a = a.cpu().numpy()
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ------------------------------------------------------------------ reductions (N8)

@pytest.mark.parametrize("dtype", ["float32", "complex64", "complex128"])
def test_fp64_reductions_match_numpy(dtype):
    import torch
    from mripy_amd import solvers as S
    print("case: reductions", dtype, flush=True)
    rng = np.random.default_rng(0)
    for n in (0, 1, 7, 1000, 262145, 10_000_019):
        a = rng.standard_normal(n) + (1j * rng.standard_normal(n) if dtype != "float32" else 0)
        b = rng.standard_normal(n) + (1j * rng.standard_normal(n) if dtype != "float32" else 0)
        a, b = a.astype(dtype), b.astype(dtype)
        ta, tb = torch.as_tensor(a).cuda(), torch.as_tensor(b).cuda()
        a64, b64 = a.astype(np.complex128), b.astype(np.complex128)
        want = np.vdot(a64, b64)
        got = S.dot_f64(ta, tb)
        assert abs(got - want) <= 1e-12 * max(1.0, np.linalg.norm(a64) * np.linalg.norm(b64))
        want2 = float(np.sum(np.abs(a64 - b64) ** 2))
        assert abs(S.sqdist_f64(ta, tb) - want2) <= 1e-12 * max(1.0, want2)
        assert abs(S.norm_f64(ta) - np.linalg.norm(a64)) <= 1e-12 * max(1.0, np.linalg.norm(a64))
    # deterministic: same bits run to run
    x = torch.as_tensor(rng.standard_normal(3_000_001).astype(np.float32)).cuda()
    assert S.sqdist_f64(x) == S.sqdist_f64(x)


def test_reduction_errors():
    import torch
    from mripy_amd import solvers as S
    a = torch.zeros(4, dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError):
        S.dot_f64(a, torch.zeros(5, dtype=torch.complex64, device="cuda"))
    with pytest.raises(TypeError):
        S.sqdist_f64(torch.zeros(4, dtype=torch.float16, device="cuda"))


# ------------------------------------------------------------------ TV prox

def test_tv_prox_kernel_matches_oracle():
    from mripy_amd import solvers as S
    print("case: TV prox vs oracle", flush=True)
    rng = np.random.default_rng(1)
    for shape, lam in [((37, 29, 3), 0.3), ((64, 64), 2.0), ((2, 5, 1), 0.05), ((48, 40, 1), 40.0)]:
        y = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 3
        want = SO.prox_tv2d_r(y, lam)
        got = S.prox_tv2d_r(y.astype(np.complex128), lam)
        assert got.dtype == np.complex128 and got.shape == y.shape
        assert rel(got, want) <= 1e-12
        y32 = y.astype(np.complex64)
        want32 = SO.prox_tv2d_r(y32.astype(np.complex128), lam)
        assert rel(S.prox_tv2d_r(y32, lam), want32) <= 1e-6
    with pytest.raises(ValueError):
        S.prox_tv2d_r(np.zeros((1, 5), np.complex64), 0.1)
    with pytest.raises(ValueError):
        S.prox_tv2d_r(np.ones((4, 5), np.complex64), -1.0)


def test_tv3d_prox_kernel_matches_oracle():
    from mripy_amd import solvers as S
    print("case: 3-D TV prox vs oracle", flush=True)
    rng = np.random.default_rng(11)
    for shape, lam in [((13, 9, 7), 0.4), ((2, 3, 2), 0.1), ((20, 16, 24), 3.0)]:
        y = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 2
        want = SO.prox_tv3d(y, lam)
        assert rel(S.prox_tv3d(y, lam), want) <= 1e-12
        assert rel(S.prox_tv3d(y.astype(np.complex64), lam), SO.prox_tv3d(y.astype(np.complex64).astype(complex), lam)) <= 1e-6
    # trailing dims are independent images
    y4 = (rng.standard_normal((6, 5, 4, 2)) + 1j * rng.standard_normal((6, 5, 4, 2)))
    got = S.prox_tv3d(y4, 0.3)
    for t in range(2):
        assert rel(got[..., t], SO.prox_tv3d(y4[..., t], 0.3)) <= 1e-12


def test_tv_prox_step_closed_form_both_amplitude_rules():
    from mripy_amd import solvers as S
    print("case: TV prox closed form", flush=True)
    sx, sy, k, lam = 16, 5, 6, 0.8
    y = np.full((sx, sy), 3.0)
    y[:k] = 1.0
    want = np.full((sx, sy), 3.0 - lam / (sx - k))
    want[:k] = 1.0 + lam / k
    for amp in ("reference", "modulus"):
        f = S.prox_tv2d_r(y.astype(np.complex128), lam, niter=4000, amp=amp)
        assert np.abs(f - want).max() < 1e-9


def test_tv_operator_class_matches_oracle():
    import torch
    from mripy_amd import solvers as S
    rng = np.random.default_rng(2)
    x = rng.standard_normal((9, 7, 2)) + 1j * rng.standard_normal((9, 7, 2))
    p = rng.standard_normal((9, 7, 2, 2)) + 1j * rng.standard_normal((9, 7, 2, 2))
    tg, to = S.TV2d_r(), SO.TV2d_r()
    xt, pt = torch.as_tensor(x).cuda(), torch.as_tensor(p).cuda()
    assert rel(tg.grad(xt).cpu(), to.grad(x)) < 1e-15
    assert rel(tg.Div(pt).cpu(), to.Div(p)) < 1e-15
    assert rel(tg.amp(pt).cpu(), to.amp(p)) < 1e-15
    x3 = rng.standard_normal((5, 6, 4)) + 1j * rng.standard_normal((5, 6, 4))
    p3 = rng.standard_normal((5, 6, 4, 3)) + 1j * rng.standard_normal((5, 6, 4, 3))
    t3g, t3o = S.TV3d(), SO.TV3d()
    assert rel(t3g.grad(torch.as_tensor(x3).cuda()).cpu(), t3o.grad(x3)) < 1e-15
    assert rel(t3g.Div(torch.as_tensor(p3).cuda()).cpu(), t3o.Div(p3)) < 1e-15


# ------------------------------------------------------------------ solvers on a small SENSE problem

def small_problem(n=64, nc=8, ninter=12, npts=512, seed=5):
    import torch
    from mripy_amd import data
    from mripy_amd.operators import NUFFT2d_hip, espirit, joint2operators
    x, y = data.archimedean_spiral_2d(ninter, npts, n)
    Smap = data.coil_maps((n, n), nc, seed=seed)
    ph = data.shepp_logan_2d(n)
    p = O.Params((n, n), rat=2, nspread=3)
    nuf_o = SO.NUFFT2d((n, n), None, x, y, p)
    A_o = SO.joint2operators(SO.espirit(Smap.astype(np.complex128)), nuf_o)
    b = A_o.forward(ph[..., None].astype(np.complex128))
    b = b + data.complex_normal(b.shape, seed=seed, sigma=1e-3)
    b = b / SO.optscaling(nuf_o, b)
    nuf = NUFFT2d_hip((n, n), None, x, y, rat=2, nspread=3)
    A = joint2operators(espirit(torch.as_tensor(Smap).cuda()), nuf)
    return A, A_o, nuf, nuf_o, b, torch.as_tensor(b.astype(np.complex64)).cuda()


def test_line_search_reference_signature():
    from mripy_amd import solvers as S
    print("case: line search", flush=True)
    A, A_o, _, _, b, bt = small_problem()
    f_o = lambda v: np.linalg.norm(A_o.forward(v) - b) ** 2
    df_o = lambda v: 2 * A_o.backward(A_o.forward(v) - b)
    f = lambda v: S.sqdist_f64(A.forward(v), bt)
    df = lambda v: 2 * A.backward(A.forward(v) - bt)
    x_o = A_o.backward(b)
    x = A.backward(bt)
    for scale in (1.0, 1e3, 1e5):      # long steps force backtracking
        a_o, i_o = SO.BacktrackingLineSearch(f_o, df_o, x_o, -scale * df_o(x_o))
        a, i = S.BacktrackingLineSearch(f, df, x, -scale * df(x))
        assert (a, i) == (a_o, i_o)


def test_cgd2_matches_fp64_oracle():
    from mripy_amd import solvers as S
    print("case: CGD2", flush=True)
    A, A_o, _, _, b, bt = small_problem()
    want = SO.prox_l2_Afxnb_CGD2(A_o.forward, A_o.backward, b, 30)
    got = S.prox_l2_Afxnb_CGD2(A.forward, A.backward, bt, 30)
    assert got.is_cuda and got.shape == want.shape
    assert rel(got.cpu().numpy(), want) <= TOL
    # numpy in, numpy out
    got_np = S.prox_l2_Afxnb_CGD2(A.forward, A.backward, b.astype(np.complex64), 30)
    assert isinstance(got_np, np.ndarray) and rel(got_np, want) <= TOL


def test_cgd2_fletcher_reeves_beta_hand_trace():
    """The GPU CG on the hand-derived trace of tests/test_solvers_oracle.py:
    A = diag(1/2, 1/4), b = (1, 1), one iteration after the first step; the
    un-squared beta = ||g1||/||g0|| puts x at (1.625 + 0.75 beta, 1.12890625 + 0.46875 beta)."""
    import torch
    from mripy_amd import solvers as S
    print("case: CGD2 hand trace", flush=True)
    a = torch.tensor([0.5, 0.25], dtype=torch.complex64, device="cuda")
    b = torch.ones(2, dtype=torch.complex64, device="cuda")
    x = S.prox_l2_Afxnb_CGD2(lambda v: a * v, lambda r: a.conj() * r, b, 1).cpu().numpy()
    beta = np.sqrt(0.3088531494140625 / 0.7822265625)
    want = np.array([1.25 + 0.375 + 0.75 * beta, 0.71875 + 0.41015625 + 0.46875 * beta])
    assert np.abs(x - want).max() < 1e-6, (x, want)


def test_cgd_proximal_matches_fp64_oracle():
    from mripy_amd import solvers as S
    print("case: CGD", flush=True)
    A, A_o, _, _, b, bt = small_problem(seed=6)
    rng = np.random.default_rng(3)
    x0 = (0.1 * (rng.standard_normal((64, 64, 1)) + 1j * rng.standard_normal((64, 64, 1))))
    want = SO.prox_l2_Afxnb_CGD(A_o.forward, A_o.backward, b, x0, 1e-2, 5)
    got = S.prox_l2_Afxnb_CGD(A.forward, A.backward, bt, x0.astype(np.complex64), 1e-2, 5)
    assert rel(got, want) <= TOL


def test_admm_tv_matches_fp64_oracle():
    from mripy_amd import solvers as S
    print("case: ADMM-TV", flush=True)
    A, A_o, _, _, b, bt = small_problem()
    want = SO.ADMM_l2Afxnb_tvx(A_o.forward, A_o.backward, b, 20, 0.5, 1e-3, 1e-2)
    got = S.ADMM_l2Afxnb_tvx(A.forward, A.backward, bt, 20, 0.5, 1e-3, 1e-2)
    assert rel(got.cpu().numpy(), want) <= TOL


class _NoFB:
    """The same operator without forward_backward (takes the other dispatch branch)."""

    def __init__(self, op):
        self.forward, self.backward = op.forward, op.backward


@pytest.mark.parametrize("wrap", ["IST_wrap", "FIST_wrap"])
def test_ist_fista_wraps_match_fp64_oracle(wrap):
    from mripy_amd import solvers as S
    print("case:", wrap, flush=True)
    A, A_o, nuf, nuf_o, b, bt = small_problem(n=48, nc=4, ninter=8, npts=512)
    T, T_o = S.TV2d_r(), SO.TV2d_r()
    th = 0.02
    for topt in (False, True):
        for fb in (True, False):
            op, op_o = (A, A_o) if fb else (_NoFB(A), _NoFB(A_o))
            want = getattr(SO, wrap)(op_o, T_o if topt else None, b=b, Nite=15, step=0.5, th=th)
            got = getattr(S, wrap)(op, T if topt else None, b=bt, Nite=15, step=0.5, th=th)
            assert rel(got.cpu().numpy(), want) <= TOL, (topt, fb)
    assert getattr(S, wrap)(A, b=None) is None


def test_optscaling():
    from mripy_amd import solvers as S
    _, _, nuf, nuf_o, b, bt = small_problem()
    assert abs(S.optscaling(nuf, bt) - SO.optscaling(nuf_o, b)) <= 1e-5 * SO.optscaling(nuf_o, b)


# ------------------------------------------------------------------ full spiral CS workload

def test_spiral_cs_cg_and_admm_vs_committed_fp64_runs():
    import torch
    from mripy_amd import data
    from mripy_amd import solvers as S
    from mripy_amd.operators import NUFFT2d_hip, espirit, joint2operators
    print("case: 256^2 x 32-coil spiral CG + ADMM-TV vs committed FP64 runs", flush=True)
    ref = np.load(os.path.join(HERE, "golden", "solvers_spiral_cs.npz"))
    cfg = data.config_2d_spiral_cs()
    x, y = cfg["coords"]
    p = O.Params(cfg["modes"], rat=2, nspread=3)
    nuf_o = SO.NUFFT2d(cfg["modes"], None, x, y, p, threads=8)
    A_o = SO.joint2operators(SO.espirit(cfg["maps"].astype(np.complex128)), nuf_o)
    b = A_o.forward(cfg["phantom"][..., None].astype(np.complex128)) + cfg["noise"].astype(np.complex128)
    nuf = NUFFT2d_hip(cfg["modes"], None, x, y, rat=2, nspread=3)
    A = joint2operators(espirit(torch.as_tensor(cfg["maps"]).cuda()), nuf)
    bt = torch.as_tensor(b.astype(np.complex64)).cuda()
    scale = S.optscaling(nuf, bt)
    assert abs(scale - float(ref["scale"])) <= 1e-5 * float(ref["scale"])
    bt = bt / float(ref["scale"])
    cg = S.prox_l2_Afxnb_CGD2(A.forward, A.backward, bt, 30)
    assert rel(cg.cpu().numpy(), ref["cg"]) <= TOL
    admm = S.ADMM_l2Afxnb_tvx(A.forward, A.backward, bt, 20, 0.5, 1e-3, 1e-2)
    assert rel(admm.cpu().numpy(), ref["admm"]) <= TOL
