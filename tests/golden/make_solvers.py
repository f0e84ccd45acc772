"""Write tests/golden/solvers_spiral_cs.npz: the FP64 oracle's CG and ADMM-TV
reconstructions of the 2-D spiral CS workload (256^2 Shepp-Logan, 32 Gaussian-blob
coil maps, 48 x 2048 variable-density Archimedean spiral, noise 1e-3 from seed 5,
rat 2, nspread 3).

b = NUFFT(S * phantom) + noise, scaled by optscaling(NUFFT, b) as
test/CS_MRI/pics_TV_ADMM_2d_uselib.py:47-48 does; then
  cg   = prox_l2_Afxnb_CGD2(A.forward, A.backward, b, 30)
  admm = ADMM_l2Afxnb_tvx(A.forward, A.backward, b, 20, step=0.5, tv_r=1e-3, rho=1e-2)
with A = joint2operators(espirit(S), NUFFT), all in complex128 by
oracle/solvers_oracle.py.  Run from the repository root:
    python tests/golden/make_solvers.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mripy_amd import data  # noqa: E402
from oracle import nufft_oracle as O  # noqa: E402
from oracle import solvers_oracle as S  # noqa: E402


def problem(threads=1):
    cfg = data.config_2d_spiral_cs()
    x, y = cfg["coords"]
    p = O.Params(cfg["modes"], rat=2, nspread=3)
    nuf = S.NUFFT2d(cfg["modes"], None, x, y, p, threads=threads)
    A = S.joint2operators(S.espirit(cfg["maps"].astype(np.complex128)), nuf)
    b = A.forward(cfg["phantom"][..., None].astype(np.complex128)) + cfg["noise"].astype(np.complex128)
    return cfg, nuf, A, b


def main():
    threads = len(os.sched_getaffinity(0))
    cfg, nuf, A, b = problem(threads)
    scale = S.optscaling(nuf, b)
    b = b / scale
    t = time.time()
    cg = S.prox_l2_Afxnb_CGD2(A.forward, A.backward, b, 30)
    print("cg %.1fs" % (time.time() - t), flush=True)
    t = time.time()
    admm = S.ADMM_l2Afxnb_tvx(A.forward, A.backward, b, 20, 0.5, 1e-3, 1e-2)
    print("admm %.1fs" % (time.time() - t), flush=True)
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "solvers_spiral_cs.npz"), scale=scale,
                        cg=cg.astype(np.complex64), admm=admm.astype(np.complex64))


if __name__ == "__main__":
    main()
