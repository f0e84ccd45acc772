"""Density compensation weights for non-Cartesian trajectories (the reference has no
estimator: its scripts read ``dcf`` from data files, test/MRI_recon/test_uterecon.py:87).

``pipe_menon`` is the iterative estimate of Pipe and Menon (Magn. Reson. Med. 41:179,
1999) on the plan's own gridding kernel:

    w_0 = 1,   w_{i+1} = w_i / | G^T G w_i |,

where G^T is the spread onto the oversampled grid and G the gather back to the
samples (the Gaussian of the NUFFT, ``Plan.grid_spread`` / ``Plan.grid_gather``, so
the effective kernel is its self-convolution).  The weights are normalised to a
maximum of 1.  Both steps run on the device; the iteration runs in FP32 on the
device and returns float32.
"""
from __future__ import annotations

import numpy as np
import torch

from .plan import Plan


def pipe_menon(coords, modes, niter=10, rat=2, nspread=3, device=None):
    """Weights (M,) for the sample coordinates (radians; d arrays of M)."""
    P = Plan(modes, ncoils=1, rat=rat, nspread=nspread, device=device)
    P.set_points(coords)
    M = P.M
    w = torch.ones((M, 1), dtype=torch.complex64, device=P.device)
    for _ in range(int(niter)):
        c = P.grid_gather(P.grid_spread(w))
        w = w / c.abs().clamp_min(1e-30)
    w = w.real.reshape(-1)
    w = w / w.max()
    return w if any(torch.is_tensor(c) for c in coords) else w.cpu().numpy()
