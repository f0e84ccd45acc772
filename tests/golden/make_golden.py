"""Write the golden fixtures under tests/golden/ with the FP64 oracle.

Inputs are float32-representable (coordinates float32, data complex64) so the
HIP path and the oracle see identical values; outputs are complex128.
Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mripy_amd import data  # noqa: E402
from oracle import nufft_oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def case_2d_radial():
    x, y = data.golden_angle_radial_2d(48, 96)
    img = data.shepp_logan_2d(48)
    c = data.complex_normal(x.shape[0], seed=0)
    p = O.Params((48, 48), rat=2, nspread=3)
    return dict(x=x, y=y, image=img, samples=c, rat=2, nspread=3, iflag=1, df=1.0,
                fwd=O.type2([x, y], img, p), adj=O.type1([x, y], c, p))


def case_2d_eps_iflag():
    rng = np.random.default_rng(11)
    M, nc = 3000, 2
    x, y = (np.pi * (6 * rng.random((2, M)) - 3)).astype(np.float32)
    img = (rng.standard_normal((30, 26, nc)) + 1j * rng.standard_normal((30, 26, nc))).astype(np.complex64)
    c = (rng.standard_normal((M, nc)) + 1j * rng.standard_normal((M, nc))).astype(np.complex64)
    p = O.Params((30, 26), eps=1e-5, iflag=-1, df=1.5)
    return dict(x=x, y=y, image=img, samples=c, rat=p.rat, nspread=p.ns, iflag=-1, df=1.5,
                fwd=O.type2([x, y], img, p), adj=O.type1([x, y], c, p))


def case_3d_ute():
    x, y, z = data.fibonacci_ute_3d(160, 32)
    S = data.coil_maps((16, 16, 16), 2, seed=1)
    img = (data.shepp_logan_3d(16)[..., None] * S).astype(np.complex64)
    c = data.complex_normal((x.shape[0], 2), seed=2)
    dcf = np.linspace(0.2, 1.0, x.shape[0]).astype(np.float32)
    p = O.Params((16, 16, 16), rat=2, nspread=3)
    return dict(x=x, y=y, z=z, image=img, samples=c, dcf=dcf, rat=2, nspread=3, iflag=1, df=1.0,
                fwd=O.type2([x, y, z], img, p), adj=O.type1([x, y, z], c, p),
                adj_dcf=O.type1([x, y, z], c * dcf[:, None], p), t21=O.type21([x, y, z], img, p, dcf))


def main():
    for name, fn in [("g2d_radial", case_2d_radial), ("g2d_eps_iflag", case_2d_eps_iflag), ("g3d_ute", case_3d_ute)]:
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **fn())
        print("wrote", name)


if __name__ == "__main__":
    main()
