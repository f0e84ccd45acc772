"""The C-ABI library loads and exports every symbol include/*.h declares (no GPU needed)."""
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = "".join(open(p).read() for p in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))))
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)  # declarations only, not comments
    return sorted(set(re.findall(r"\b(gn_[a-z_0-9]+)\s*\(", src)))


def test_header_symbols_exported():
    from mripy_amd import _lib
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 19
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"


def test_version_and_params_without_gpu():
    from mripy_amd import _lib
    from mripy_amd.plan import grid_params
    from oracle import nufft_oracle as O
    assert _lib.load().gn_version() == 1
    for modes, eps, mx in [((128, 128, 128), 1e-5, 6), ((128, 128), 1e-3, None), ((16, 16), 1e-15, None),
                           ((32, 32, 16), 1e-15, 6), ((64, 48), 5e-6, None)]:
        rat, ns, nf, tau = grid_params(modes, eps, mx)
        o_ns, o_nf, o_tau, o_rat = O.grid_params(modes, eps=eps, max_nspread=mx)
        assert (rat, ns, nf) == (o_rat, o_ns, o_nf)
        assert abs(tau - o_tau) <= 1e-15 * o_tau


def test_eps_out_of_range_raises():
    import pytest
    from mripy_amd.plan import grid_params
    with pytest.raises(ValueError):
        grid_params((16, 16), 0.2)
    with pytest.raises(ValueError):
        grid_params((16, 16), 1e-40)
