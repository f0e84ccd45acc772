"""CPU tests that pin the FP64 solver oracle (oracle/solvers_oracle.py).

Pins: nonlinear CG on dense operators whose Hessian is the identity (one step of
length 1 must land on numpy.linalg.lstsq / the proximal normal equations, and the
gradient test must stop the loop); the exhausted-line-search stop; Armijo
line-search acceptance; Div as
the exact adjoint of grad (a dropped boundary term fails it); the TV prox of a
constant (fixed point) and of a two-level step image (closed-form 1-D TV
denoising: each level moves by lambda / its length); IST/FISTA on the identity and
on a dense full-rank operator against their fixed points; IST_wrap/FIST_wrap
dispatch equivalence; ADMM-TV with a vanishing TV weight against least squares;
and a planted sign mistake in grad that the closed form must catch.
"""
import cmath

import numpy as np
import pytest

from oracle import solvers_oracle as S


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def dense_problem(m=40, n=12, seed=0, noise=0.05):
    rng = np.random.default_rng(seed)
    A = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))) / np.sqrt(m)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    b = A @ x + noise * (rng.standard_normal(m) + 1j * rng.standard_normal(m))
    return A, b


class Dense:
    """forward = A x, backward = A^H y, forward_backward = A^H A x."""

    def __init__(self, A):
        self.A = A

    def forward(self, x):
        return self.A @ x

    def backward(self, y):
        return self.A.conj().T @ y

    def forward_backward(self, x):
        return self.backward(self.forward(x))


class DenseNoFB:
    def __init__(self, A):
        self.A = A

    def forward(self, x):
        return self.A @ x

    def backward(self, y):
        return self.A.conj().T @ y


# ---------------------------------------------------------------- CG, line search

def orthonormal_columns(m, n, seed):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n)))
    return Q


def test_cgd2_one_newton_step_is_exact():
    # A^H A = I/2 makes the Hessian 2 A^H A the identity: the first trial step alpha = 1
    # along -df lands on the least-squares solution and the gradient test stops the loop
    A = orthonormal_columns(40, 12, 0) / np.sqrt(2)
    rng = np.random.default_rng(1)
    b = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    x_ls = np.linalg.lstsq(A, b, rcond=None)[0]
    x = S.prox_l2_Afxnb_CGD2(lambda v: A @ v, lambda r: A.conj().T @ r, b, Nite=30)
    assert rel(x, x_ls) < 1e-13
    # a scaled operator needs backtracking and further iterations, and still decreases f
    A2 = 3 * A
    f = lambda v: np.linalg.norm(A2 @ v - b) ** 2
    x2 = S.prox_l2_Afxnb_CGD2(lambda v: A2 @ v, lambda r: A2.conj().T @ r, b, Nite=30)
    assert f(x2) < f(A2.conj().T @ b)


def test_cgd_proximal_one_newton_step_is_exact():
    # 2 A^H A + rho I = I: the minimizer of ||Ax-b||^2 + rho/2||x-x0||^2 in one step
    rho = 0.5
    A = orthonormal_columns(30, 10, 2) * np.sqrt((1 - rho) / 2)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(30) + 1j * rng.standard_normal(30)
    x0 = rng.standard_normal(10) + 1j * rng.standard_normal(10)
    want = np.linalg.solve(2 * A.conj().T @ A + rho * np.eye(10), 2 * A.conj().T @ b + rho * x0)
    x = S.prox_l2_Afxnb_CGD(lambda v: A @ v, lambda r: A.conj().T @ r, b, x0, rho, Nite=3)
    assert rel(x, want) < 1e-13


# A worked two-step trace of prox_l2_Afxnb_CGD2 (pics/proximal_func.py:314-343) on
# A = diag(1/2, 1/4), b = (1, 1), Nite = 1, derived by hand:
#   x0 = A^H b = (0.5, 0.25);  g0 = 2 A^H (A x0 - b) = (-0.75, -0.46875)
#   trial step alpha = 1 along -g0: x0 - g0 = (1.25, 0.71875) with f = 0.81353759765625
#   < f(x0) = 1.44140625, so the line search accepts it at once:  x1 = (1.25, 0.71875)
#   g1 = 2 A^H (A x1 - b) = (-0.375, -0.41015625)
#   beta = ||g1|| / ||g0|| = sqrt(0.3088531494140625 / 0.7822265625)   (norms, not squares)
#   s1 = -g1 + beta (-g0);  x1 + s1 has f = 0.41723 < f(x1): accepted, x2 = x1 + s1.
CG_TRACE_A = np.diag([0.5, 0.25]).astype(np.complex128)
CG_TRACE_BETA = np.sqrt(0.3088531494140625 / 0.7822265625)


def cg_trace_expected(beta):
    return np.array([1.25 + 0.375 + 0.75 * beta, 0.71875 + 0.41015625 + 0.46875 * beta])


def test_cgd2_fletcher_reeves_beta_hand_trace():
    A, b = CG_TRACE_A, np.ones(2, np.complex128)
    x = S.prox_l2_Afxnb_CGD2(lambda v: A @ v, lambda r: A.conj().T @ r, b, Nite=1)
    assert np.abs(x - cg_trace_expected(CG_TRACE_BETA)).max() < 1e-14
    # the textbook squared ratio ||g1||^2/||g0||^2 lands elsewhere: the trace tells them apart
    assert np.abs(cg_trace_expected(CG_TRACE_BETA ** 2) - cg_trace_expected(CG_TRACE_BETA)).max() > 0.1


def test_cgd2_stops_when_the_line_search_is_exhausted():
    # the un-squared Fletcher-Reeves beta lets the direction turn uphill; the line search
    # then backtracks ls_Nite times and the loop ends (pics/proximal_func.py:331)
    A, b = dense_problem()
    calls = []
    orig = S.BacktrackingLineSearch

    def spy(*a, **k):
        r = orig(*a, **k)
        calls.append(r[1])
        return r
    S.BacktrackingLineSearch = spy
    try:
        S.prox_l2_Afxnb_CGD2(lambda v: A @ v, lambda r: A.conj().T @ r, b, Nite=200)
    finally:
        S.BacktrackingLineSearch = orig
    assert calls[-1] == 10 and all(c < 10 for c in calls[:-1]) and len(calls) < 200


def test_line_search_acceptance():
    A, b = dense_problem(20, 8, seed=3)
    f = lambda v: np.linalg.norm(A @ v - b) ** 2
    df = lambda v: 2 * A.conj().T @ (A @ v - b)
    rng = np.random.default_rng(4)
    for k in range(20):
        x = rng.standard_normal(8) + 1j * rng.standard_normal(8)
        p = -df(x) * 10.0 ** (k % 5)      # over-long steps force backtracking
        alpha, i = S.BacktrackingLineSearch(f, df, x, p)
        assert alpha == pytest.approx(0.2 ** i, rel=1e-15)
        derphi = np.real(np.vdot(df(x), p))
        f0, ft = f(x), f(x + alpha * p)
        assert i == 10 or ft <= f0 or ft - f0 <= 1e-4 * alpha * derphi
        if i > 0:   # the previous step was rejected
            fp = f(x + alpha / 0.2 * p)
            assert fp > f0 and fp - f0 > 1e-4 * alpha / 0.2 * derphi


# ---------------------------------------------------------------- TV

def test_div_is_adjoint_of_grad():
    rng = np.random.default_rng(5)
    tv = S.TV2d_r()
    for shape in [(7, 5), (6, 9, 3), (2, 2, 1)]:
        x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        p = rng.standard_normal(shape + (2,)) + 1j * rng.standard_normal(shape + (2,))
        lhs = np.vdot(tv.grad(x), p)
        rhs = np.vdot(x, tv.Div(p))
        assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_div3d_is_adjoint_of_grad3d():
    rng = np.random.default_rng(11)
    tv = S.TV3d()
    for shape in [(5, 4, 6), (2, 3, 2)]:
        x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        p = rng.standard_normal(shape + (3,)) + 1j * rng.standard_normal(shape + (3,))
        assert abs(np.vdot(tv.grad(x), p) - np.vdot(x, tv.Div(p))) < 1e-12 * abs(np.vdot(tv.grad(x), p))


def test_tv3d_prox_step_closed_form():
    # a jump along axis 2, constant along axes 0 and 1: every line is the 1-D problem
    sz, k, lam = 12, 5, 0.6
    y = np.full((4, 3, sz), 3.0)
    y[:, :, :k] = 1.0
    f = S.prox_tv3d(y.astype(np.complex128), lam, niter=4000)
    want = np.full(y.shape, 3.0 - lam / (sz - k))
    want[:, :, :k] = 1.0 + lam / k
    assert np.abs(f - want).max() < 1e-9


def test_tv_prox_of_constant_is_constant():
    y = np.full((8, 6, 2), 0.7 - 0.3j)
    assert np.abs(S.prox_tv2d_r(y, 0.5) - y).max() < 1e-15


def step_image(sx=16, sy=5, k=6, a=1.0, b=3.0):
    y = np.full((sx, sy), b, dtype=np.float64)
    y[:k] = a
    return y


def step_closed_form(sx, sy, k, a, b, lam):
    # min 1/2||f-y||^2 + lam * sum|grad f| for a jump b > a along axis 0: each column is
    # the 1-D problem with solution f = a + lam/k (first k rows), b - lam/(sx-k) (rest)
    f = np.full((sx, sy), b - lam / (sx - k))
    f[:k] = a + lam / k
    return f


@pytest.mark.parametrize("k,lam", [(6, 0.8), (9, 1.5), (3, 0.4)])
def test_tv_prox_step_closed_form(k, lam):
    y = step_image(k=k)
    f = S.prox_tv2d_r(y.astype(np.complex128), lam, niter=4000)
    want = step_closed_form(16, 5, k, 1.0, 3.0, lam)
    assert np.abs(f - want).max() < 1e-6
    # the reference's 40 iterations are a partial solve
    f40 = S.prox_tv2d_r(y.astype(np.complex128), lam)
    assert np.abs(f40 - y).max() > 1e-2 and np.abs(f40 - want).max() > 1e-3


def test_planted_grad_sign_mistake_is_caught():
    y = step_image().astype(np.complex128)
    orig = S.TV2d_r.grad
    try:
        S.TV2d_r.grad = lambda self, x: -orig(self, x)
        f = S.prox_tv2d_r(y, 0.8, niter=4000)
    finally:
        S.TV2d_r.grad = orig
    assert np.abs(f - step_closed_form(16, 5, 6, 1.0, 3.0, 0.8)).max() > 1e-2


def test_tv_complex_amplitude_rule():
    # sqrt(Gx^2 + Gy^2) with complex squares and numpy's lexicographic complex maximum
    rng = np.random.default_rng(6)
    G = (rng.standard_normal((5, 4, 2)) + 1j * rng.standard_normal((5, 4, 2))) * 1.5
    d = S.TV2d_r().amp(G)
    out = G / np.maximum(d, np.ones(G.shape))
    for i in range(5):
        for j in range(4):
            s = cmath.sqrt(G[i, j, 0] ** 2 + G[i, j, 1] ** 2)
            keep = s.real > 1 or (s.real == 1 and s.imag >= 0)
            want = G[i, j] / s if keep else G[i, j]
            assert np.allclose(out[i, j], want, rtol=1e-14, atol=0)


# ---------------------------------------------------------------- IST / FISTA / ADMM

@pytest.mark.parametrize("fn", ["IST_wrap", "FIST_wrap"])
def test_ist_fista_identity_fixed_point(fn):
    rng = np.random.default_rng(7)
    b = rng.standard_normal(30) + 1j * rng.standard_normal(30)
    op = Dense(np.eye(30))
    x = getattr(S, fn)(op, b=b, Nite=400, step=0.5, th=0.05)
    assert rel(x, b) < 1e-9


@pytest.mark.parametrize("fn", ["IST_wrap", "FIST_wrap"])
def test_ist_fista_dense_fixed_point_is_least_squares(fn):
    A, b = dense_problem(40, 8, seed=8)
    x_ls = np.linalg.lstsq(A, b, rcond=None)[0]
    # the FISTA variant accumulates gradients taken at the extrapolated point and needs
    # a smaller step than IST to stay stable on this operator (|A|^2 ~ 2)
    x = getattr(S, fn)(Dense(A), b=b, Nite=3000, step=0.5 if fn == "IST_wrap" else 0.1, th=0.02)
    assert rel(x, x_ls) < 1e-8


@pytest.mark.parametrize("fn", ["IST_wrap", "FIST_wrap"])
def test_wrap_dispatch_variants_agree(fn):
    A, b = dense_problem(25, 6, seed=9)
    x1 = getattr(S, fn)(Dense(A), b=b, Nite=30)
    x2 = getattr(S, fn)(DenseNoFB(A), b=b, Nite=30)
    assert rel(x1, x2) < 1e-12
    T = Dense(np.diag(np.linspace(0.5, 2.0, 6)).astype(np.complex128))
    y1 = getattr(S, fn)(Dense(A), T, b=b, Nite=30)
    y2 = getattr(S, fn)(DenseNoFB(A), T, b=b, Nite=30)
    assert rel(y1, y2) < 1e-12
    assert getattr(S, fn)(Dense(A), b=None) is None


def test_admm_tv_with_vanishing_weight_is_least_squares():
    rng = np.random.default_rng(10)
    n = 6 * 5
    A = (rng.standard_normal((60, n)) + 1j * rng.standard_normal((60, n))) / np.sqrt(60)
    xt = rng.standard_normal((6, 5, 1)) + 1j * rng.standard_normal((6, 5, 1))
    b = A @ xt.reshape(-1)
    Af = lambda v: A @ v.reshape(-1)
    AHf = lambda r: (A.conj().T @ r).reshape(6, 5, 1)
    x = S.ADMM_l2Afxnb_tvx(Af, AHf, b, Nite=60, step=0.5, tv_r=1e-12, rho=1.0, cgd_Nite=20)
    assert rel(x, xt) < 1e-3
