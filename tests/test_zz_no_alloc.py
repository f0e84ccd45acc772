"""gn_execute_type21_noalloc: the normal operator with its intermediate samples in the
library's static device scratch instead of plan memory.  Through the C ABI, its
output must equal gn_execute(GN_TYPE21) bit for bit (the same kernels in the same
order; only the intermediate buffer differs), on two calls in a row, the second
with a smaller problem, and it must reject a problem larger than the scratch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(modes, nc, M, seed):
    import torch
    from mripy_amd.plan import Plan
    rng = np.random.default_rng(seed)
    d = len(modes)
    coords = [(np.pi * (2 * rng.random(M) - 1)).astype(np.float32) for _ in range(d)]
    dcf = rng.random(M).astype(np.float32)
    P = Plan(modes, ncoils=nc, rat=2, nspread=3).set_points(coords, dcf=dcf)
    x = torch.as_tensor((rng.standard_normal(modes + (nc,)) + 1j * rng.standard_normal(modes + (nc,)))
                        .astype(np.complex64)).cuda()
    return P, x


def test_noalloc_type21_equals_type21_twice():
    import torch
    from mripy_amd import _lib
    L = _lib.load()
    for modes, nc, M, seed in (((32, 24, 20), 3, 30000, 1), ((48, 40), 2, 5000, 2)):
        print("case: no-alloc type 21", modes, nc, M, flush=True)
        P, x = _case(modes, nc, M, seed)
        want = P.type21(x)
        got = torch.empty_like(want)
        torch.cuda.synchronize()
        rc = L.gn_execute_type21_noalloc(P._h, x.data_ptr(), got.data_ptr())
        assert rc == 0, L.gn_last_error()
        torch.cuda.synchronize()
        assert torch.equal(got, want)
    # larger than the 4,194,304 complex samples of the scratch: a clear error
    P, x = _case((16, 16, 16), 9, 500000, 3)
    out = torch.empty_like(x)
    assert L.gn_execute_type21_noalloc(P._h, x.data_ptr(), out.data_ptr()) == _lib.GN_EINVAL
    assert b"static scratch" in L.gn_last_error()
