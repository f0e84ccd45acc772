"""Host-side behaviour of the operator classes (no GPU calls): set_params validation and
the plan invalidation it implies."""
import numpy as np
import pytest


def test_set_params_validates_and_invalidates_plan():
    from mripy_amd.operators import NUFFT2d_hip, NUFFT3d_cuda
    x = np.zeros(8, np.float32)
    op = NUFFT3d_cuda((16, 16, 16), None, x, x, x)
    assert (op.rat, op.nspread) == (2, 5)          # eps = 1e-5, as the reference operator
    op._plan, op._plan_key, op._points_key = object(), "k", "p"
    assert op.set_params(nspread=5) is op and op._plan is not None   # unchanged: plan kept
    op.set_params(rat=3, nspread=6)
    assert (op.rat, op.nspread) == (3, 6) and op._plan is None and op._points_key is None
    with pytest.raises(ValueError):
        op.set_params(nspread=9)                   # 3-D limit 8
    with pytest.raises(ValueError):
        op.set_params(rat=1)
    op2 = NUFFT2d_hip((32, 32), None, x, x, eps=1e-15)
    assert (op2.rat, op2.nspread) == (3, 14)
    op2.set_params(nspread=16)
    assert op2.nspread == 16
    with pytest.raises(ValueError):
        op2.set_params(nspread=17)
