"""Golden fixtures: the oracle reproduces them exactly (CPU), the HIP path to 1e-5 rel-L2 (GPU)."""
import os

import numpy as np
import pytest

from oracle import nufft_oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["g2d_radial", "g2d_eps_iflag", "g3d_ute"]


def load(name):
    return dict(np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False))


def rel(a, b):
    return float(np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def coords_of(g):
    return [g[k] for k in ("x", "y", "z") if k in g]


def params_of(g):
    modes = g["image"].shape[:len(coords_of(g))]
    return O.Params(modes, rat=int(g["rat"]), nspread=int(g["nspread"]), df=float(g["df"]), iflag=int(g["iflag"]))


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden(name):
    g = load(name)
    p = params_of(g)
    cs = coords_of(g)
    assert rel(O.type2(cs, g["image"], p), g["fwd"]) < 1e-13
    assert rel(O.type1(cs, g["samples"], p), g["adj"]) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_hip_matches_golden(name):
    from mripy_amd.plan import Plan
    print("case: golden", name, flush=True)
    g = load(name)
    cs = coords_of(g)
    d = len(cs)
    img = g["image"]
    nc = img.shape[d] if img.ndim > d else 1
    P = Plan(img.shape[:d], ncoils=nc, rat=int(g["rat"]), nspread=int(g["nspread"]), iflag=int(g["iflag"]))
    P.set_points(cs, df=float(g["df"]))
    fwd = P.type2(img.reshape(P.image_shape())).reshape(g["fwd"].shape)
    adj = P.type1(g["samples"].reshape(P.sample_shape())).reshape(g["adj"].shape)
    assert rel(fwd, g["fwd"]) <= 1e-5
    assert rel(adj, g["adj"]) <= 1e-5
    if "dcf" in g:
        P.set_points(cs, df=float(g["df"]), dcf=g["dcf"])
        adj_d = P.type1(g["samples"].reshape(P.sample_shape())).reshape(g["adj"].shape)
        assert rel(adj_d, g["adj_dcf"]) <= 1e-5
        t21 = P.type21(img.reshape(P.image_shape())).reshape(g["t21"].shape)
        assert rel(t21, g["t21"]) <= 1e-5
