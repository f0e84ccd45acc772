"""Gridding alone, density compensation and the data-file paths on the HIP path.

* GN_GRID_SPREAD / GN_GRID_GATHER (the reference's build_grid_{2d,3d}{1,2}_fast)
  against the C oracle's spread and gather (rel-L2 <= 1e-5), 2-D and 3-D, coil-major
  grids.
* Pipe-Menon weights on the device against the FP64 oracle's iteration (rel-L2 <=
  1e-4 after 10 iterations), and uniform on a Cartesian lattice.
* The reference's 2-D spiral dataset (data/spiral.mat, copied to
  tests/golden/spiral.mat; no reference output exists for it): the loader's
  coordinates, weights and 8-coil samples; NUFFT2d.backward with the file's weights
  against the C oracle (rel-L2 <= 1e-5); the operator's normal operator and a
  5-iteration CG on it against the FP64 solver oracle (rel-L2 <= 1e-4).
* The UTE preparation of the reference script (im_shape = 2 max(k), k / im_shape)
  on a synthetic .mat file written here.
"""
import os

import numpy as np
import pytest

from oracle import coracle as C
from oracle import nufft_oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    if hasattr(a, "cpu"):
        a = a.cpu().numpy()
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("modes,nc", [((40, 36), 3), ((20, 16, 12), 2)])
def test_grid_spread_and_gather_match_oracle(modes, nc):
    from mripy_amd.plan import Plan
    print("case: gridding alone", modes, nc, flush=True)
    rng = np.random.default_rng(len(modes) + nc)
    d = len(modes)
    M = 4000
    coords = [(np.pi * (2 * rng.random(M) - 1)).astype(np.float32) for _ in range(d)]
    p = O.Params(modes, rat=2, nspread=3)
    P = Plan(modes, ncoils=nc, rat=2, nspread=3).set_points(coords)
    y = (rng.standard_normal((M, nc)) + 1j * rng.standard_normal((M, nc))).astype(np.complex64)
    g = P.grid_spread(y)
    ref = C.spread(coords, y, p.nf, p.ns, p.tau)           # coil-last (*nf, nc)
    assert g.shape == (nc,) + p.nf
    assert rel(np.moveaxis(g, 0, -1), ref) <= 1e-5
    grid = (rng.standard_normal((nc,) + p.nf) + 1j * rng.standard_normal((nc,) + p.nf)).astype(np.complex64)
    s = P.grid_gather(grid)
    assert rel(s, C.gather(coords, np.moveaxis(grid, 0, -1), p.nf, p.ns, p.tau)) <= 1e-5


def test_pipe_menon_matches_fp64_iteration():
    from mripy_amd import data
    from mripy_amd.dcf import pipe_menon
    print("case: Pipe-Menon weights", flush=True)
    x, y = data.golden_angle_radial_2d(96, 128)
    w = pipe_menon([x, y], (64, 64), niter=10)
    ref = O.pipe_menon([x.astype(np.float64), y.astype(np.float64)], (64, 64), niter=10)
    assert rel(w, ref) <= 1e-4
    n = np.arange(16) - 8 + 0.25
    X, Y = np.meshgrid(2 * np.pi * n / 16, 2 * np.pi * n / 16, indexing="ij")
    wl = pipe_menon([X.ravel().astype(np.float32), Y.ravel().astype(np.float32)], (16, 16), niter=5)
    assert np.abs(wl - 1).max() < 1e-5


def test_spiral_mat_gridding_reconstruction():
    import torch
    from mripy_amd import io
    from mripy_amd import solvers as S
    from mripy_amd.operators import NUFFT2d_hip
    from oracle import solvers_oracle as SO
    print("case: data/spiral.mat", flush=True)
    d = io.load_spiral_mat(os.path.join(HERE, "golden", "spiral.mat"))
    x, y = [c.astype(np.float32) for c in d["coords"]]
    M = x.shape[0]
    assert d["kdata"].shape == (M, 8) and M == 1182 * 60
    assert np.abs(np.hypot(x, y)).max() <= np.pi * 1.0000001
    dcf = d["dcf"].astype(np.float32)
    nuf = NUFFT2d_hip(d["im_shape"], dcf, x, y, rat=2, nspread=3)
    im = nuf.backward(d["kdata"])
    p = O.Params(d["im_shape"], rat=2, nspread=3)
    ref = C.type1([x, y], d["kdata"] * dcf[:, None], p)
    assert im.shape == d["im_shape"] + (8,) and rel(im, ref) <= 1e-5
    # CG on the 8-coil normal equations (no coil maps: each coil independently)
    nuf_o = SO.NUFFT2d(d["im_shape"], dcf, x, y, p, threads=8)
    b = d["kdata"].astype(np.complex128)
    want = SO.prox_l2_Afxnb_CGD2(nuf_o.forward, nuf_o.backward, b, 5)
    bt = torch.as_tensor(d["kdata"]).cuda()
    got = S.prox_l2_Afxnb_CGD2(nuf.forward, nuf.backward, bt, 5)
    assert rel(got, want) <= 1e-4


def test_ute_mat_preparation(tmp_path):
    import scipy.io as sio
    from mripy_amd import io
    print("case: UTE .mat preparation", flush=True)
    rng = np.random.default_rng(0)
    ktraj = np.stack([rng.uniform(-16, 16, (40, 20)), rng.uniform(-24, 24, (40, 20)),
                      rng.uniform(-8, 8, (40, 20))])
    ktraj[0, 0, 0], ktraj[1, 0, 0], ktraj[2, 0, 0] = 16.0, 24.0, 8.0
    kdata = (rng.standard_normal((40, 20, 1, 4)) + 1j * rng.standard_normal((40, 20, 1, 4)))
    dcf = rng.random((40, 20))
    path = str(tmp_path / "ute.mat")
    sio.savemat(path, {"ktraj": ktraj, "dcf": dcf, "kdata": kdata})
    u = io.load_ute_mat(path)
    assert u["im_shape"] == [32, 48, 16]
    assert np.allclose(u["ktraj"][0], ktraj[0] / 32) and np.allclose(u["ktraj"][2], ktraj[2] / 16)
    assert u["kdata"].shape == (800, 4) and u["kdata"].dtype == np.complex64
