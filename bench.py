#!/usr/bin/env python
"""Throughput of one forward + one adjoint Gaussian-gridding NUFFT on one MI355X.

Workload (the JSON line's ``value``): 3-D multi-coil UTE -- 128^3 complex64
ellipsoid phantom times 8 Gaussian-blob coil maps, 4096 Fibonacci-sphere spokes x
256 centre-out points (1,048,576 samples), oversampling 2, 6 taps per axis
(nspread 3), complex64.  A step is forward (type 2) of the image followed by the
adjoint (type 1) of the resulting samples, on device-resident data.
value = coil-samples per second.

Further legs, each reported as its own object in the same JSON line:
  * ``nspread5``: the same workload at the operator's default eps = 1e-5 (nspread 5);
  * ``differentiable``: 128^3 x 4 coils, 1,048,576 perturbed UTE samples, one
    forward + backward of ||A(x) f - y||^2 with image and trajectory gradients,
    with fixed coordinates and with coordinates updated (re-binned) every step;
  * ``cg_sense``: 256^2 x 32-coil spiral SENSE (48 x 2048 points): the operator pair,
    30 CG iterations and 20 ADMM-TV iterations, against the committed FP64 runs;
  * ``batched_uneven``: 64 uneven 2-D instances, plain plans vs one batched launch
    sequence (autotuned tiles) vs its graph replay;
  * ``copy``: a 1 GiB device copy, the box's sustained HBM bandwidth;
  * ``cpu_baseline``: the FP64 C restatement of the reference loop nests on the
    host cores (rank 0, N=1), and the rel-L2 of the timed GPU results against
    those FP64 results.

Per-stage times come from HIP events recorded on the plan's stream inside the
timed region; ``roofline`` is computed for the dominant stage from its
algorithmic bytes per launch.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
COPY_BYTES = 1 << 30
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_3d_ute_8coil.txt")
METRIC = "NUFFT fwd+adj samples/s & % HBM roofline; batched/diff vs plain at <=1e-5 rel-L2"


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def measured_traffic(stage, tiles, nspread):
    """HBM bytes per launch of ``stage`` from the committed FETCH_SIZE/WRITE_SIZE
    counter summary (lines ``traffic <stage> <bytes> tiles <t0,t1,t2> nspread <n>``),
    or None when no measurement exists for this kernel configuration."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                w = line.split()
                if (len(w) == 7 and w[0] == "traffic" and w[1] == stage and w[4] == ",".join(map(str, tiles))
                        and int(w[6]) == nspread):
                    return float(w[2])
    except OSError:
        pass
    return None


def algorithmic_bytes(d, N, G, M, nc, modes=None, nf=None):
    """Per-launch algorithmic HBM bytes of each stage (complex64 = 8 B, coordinates 4 B/axis).

    gather / spread: the grid, the samples and the coordinates.  With the pruned FFT
    passes (``modes`` and ``nf`` given) pad and crop are the first / last FFT pass
    (image <-> cells with n0 < m0, n1 < m1 on the leading axes) and fft2 / fft1 the
    remaining column passes, each at its minimal read + write; otherwise pad and
    crop are plain passes and fft2 / fft1 one read + one write of the grid."""
    if modes is None:
        return {"pad": 8 * N * nc + 8 * G * nc, "fft2": 16 * G * nc,
                "gather": 8 * G * nc + 4 * d * M + 8 * M * nc, "spread": 8 * M * nc + 4 * d * M + 8 * G * nc,
                "fft1": 16 * G * nc, "crop": 16 * N * nc}
    if d == 3:
        t1 = modes[0] * modes[1] * nf[2]      # after the last-axis pass
        t2 = modes[0] * nf[1] * nf[2]         # after the middle-axis pass
        cols = 8 * nc * ((t1 + t2) + (t2 + G))
    else:
        t1 = modes[0] * nf[1]
        cols = 8 * nc * (t1 + G)
    return {"pad": 8 * N * nc + 8 * t1 * nc, "fft2": cols,
            "gather": 8 * G * nc + 4 * d * M + 8 * M * nc, "spread": 8 * M * nc + 4 * d * M + 8 * G * nc,
            "fft1": cols, "crop": 8 * t1 * nc + 8 * N * nc}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def timed(fn, steps, warmup, stream=None):
    """ms per call of fn over ``steps`` calls after ``warmup`` calls (HIP events)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def copy_bandwidth(dev):
    """Sustained device copy bandwidth (read + write bytes / time) of a 1 GiB copy."""
    import torch
    a = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    ms = timed(lambda: b.copy_(a), 10, 3)
    del a, b
    torch.cuda.empty_cache()
    return {"bytes": 2 * COPY_BYTES, "ms": ms, "GBps": 2 * COPY_BYTES / (ms * 1e-3) / 1e9,
            "kernel": "torch copy_ of 1 GiB float32 (1 GiB read + 1 GiB write)"}


# ----------------------------------------------------------------------------- 3-D legs

def leg_3d(cfg, dev, nspread, steps, warmup, tiles=None, keep=False):
    import numpy as np
    import torch

    from mripy_amd._lib import GN_TYPE1, GN_TYPE2
    from mripy_amd.plan import Plan

    modes = cfg["modes"]
    nc = cfg["image"].shape[-1]
    M = cfg["coords"][0].shape[0]
    d, N = 3, int(np.prod(modes))
    P = Plan(modes, ncoils=nc, rat=2, nspread=nspread, device=dev, tiles=tiles)
    G = int(np.prod(P.nf))
    coords = [torch.as_tensor(c, device=dev) for c in cfg["coords"]]
    sp = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        P.set_points(coords)
        torch.cuda.synchronize()
        sp.append((time.perf_counter() - t0) * 1e3)
    img = torch.as_tensor(cfg["image"], device=dev).contiguous()
    ybuf = torch.empty((M, nc), dtype=torch.complex64, device=dev)
    obuf = torch.empty(modes + (nc,), dtype=torch.complex64, device=dev)

    def step():
        P.execute_into(GN_TYPE2, img, ybuf)
        P.execute_into(GN_TYPE1, ybuf, obuf)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    P.set_timing(True)
    P.stage_times(reset=True)
    stream = torch.cuda.current_stream(dev)
    t0 = time.perf_counter()
    ms_step = timed(step, steps, 0, stream)
    wall = time.perf_counter() - t0
    st = P.stage_times(reset=True)
    P.set_timing(False)
    ab = algorithmic_bytes(d, N, G, M, nc, modes, P.nf)
    stages = {k: {"ms": v[0] / max(1, v[1]), "launches": v[1]} for k, v in st.items()}
    for k in stages:
        stages[k]["GBps"] = ab[k] / (stages[k]["ms"] * 1e-3) / 1e9 if stages[k]["ms"] > 0 else None
    # the normal operator (type 21 with density weights 1: type 2, then type 1, one call)
    from mripy_amd._lib import GN_TYPE21
    obuf2 = torch.empty_like(obuf)
    ms_t21 = timed(lambda: P.execute_into(GN_TYPE21, img, obuf2), max(3, steps // 2), 1, stream)
    del obuf2
    out = {"ms_per_step": ms_step, "value": M * nc / (ms_step * 1e-3), "stages": stages, "ab": ab,
           "type21_ms": ms_t21,
           "set_points_ms": sorted(sp)[1], "tiles": list(P.tiles), "gather_tiles": list(P.gather_tiles),
           "nf": list(P.nf), "wall_s_timed": wall, "path_bytes": sum(algorithmic_bytes(d, N, G, M, nc).values())}
    if keep:
        out["_keep"] = (P, img, ybuf, obuf)
    return out


def leg_differentiable(dev, steps, warmup):
    """Forward + backward of ||A(x) f - y||^2 at 128^3 x 4 coils, M = 1,048,576."""
    import numpy as np
    import torch

    from mripy_amd import data
    from mripy_amd.autograd import NUFFT
    n, nc = 128, 4
    x, y, z = data.fibonacci_ute_3d(4096, 256)
    rng = np.random.default_rng(2)
    coords = [(c + 2 * np.pi * rng.uniform(-0.1 / n, 0.1 / n, c.shape[0])).astype(np.float32) for c in (x, y, z)]
    S = data.coil_maps((n, n, n), nc, seed=0)
    f0 = (data.shepp_logan_3d(n)[..., None] * S).astype(np.complex64)
    op = NUFFT((n, n, n), ncoils=nc, rat=2, nspread=3, device=dev)
    ct = torch.tensor(np.stack(coords), device=dev, requires_grad=True)
    with torch.no_grad():
        y0 = op(torch.as_tensor(f0, device=dev), ct)
        yt = y0 + torch.as_tensor(data.complex_normal(tuple(y0.shape), seed=1, sigma=0.01), device=dev)
    ft = torch.tensor((0.9 * f0 + 0.05 * data.complex_normal(f0.shape, seed=3)).astype(np.complex64), device=dev,
                      requires_grad=True)

    def step():
        ft.grad = None
        ct.grad = None
        loss = (op(ft, ct) - yt).abs().pow(2).sum()
        loss.backward()

    ms_fixed = timed(step, steps, warmup)
    delta = torch.full_like(ct, 1e-7)

    def step_moving():
        with torch.no_grad():
            ct.add_(delta)          # a trajectory update: the plan re-bins in the next forward
        step()

    ms_moving = timed(step_moving, max(3, steps // 2), 1)
    M = coords[0].shape[0]
    return {"workload": "differentiable NUFFT 128^3 x 4 coils, 1,048,576 UTE samples perturbed by U(-0.1/N, 0.1/N); "
                        "loss ||A(x) f - y||^2, backward gives image and trajectory gradients (type 2 + exact "
                        "adjoint + trajectory-gradient kernel)",
            "ms_fwd_bwd_fixed_coords": ms_fixed, "ms_fwd_bwd_with_rebin": ms_moving,
            "coil_samples_per_s_fixed": M * nc / (ms_fixed * 1e-3)}


def leg_cg_sense(dev):
    """CG-SENSE and ADMM-TV on the 256^2 x 32-coil spiral workload."""
    import numpy as np
    import torch

    from mripy_amd import data
    from mripy_amd import solvers as S
    from mripy_amd.operators import NUFFT2d_hip, espirit, joint2operators
    cfg = data.config_2d_spiral_cs()
    x, y = cfg["coords"]
    nuf = NUFFT2d_hip(cfg["modes"], None, x, y, rat=2, nspread=3)
    A = joint2operators(espirit(torch.as_tensor(cfg["maps"], device=dev)), nuf)
    ph = torch.as_tensor(cfg["phantom"][..., None], device=dev)
    b = A.forward(ph) + torch.as_tensor(cfg["noise"], device=dev)
    scale = S.optscaling(nuf, b)
    b = b / scale
    xi = A.backward(b)
    ms_pair = timed(lambda: A.backward(A.forward(xi)), 10, 3)
    ms_normal = timed(lambda: A.forward_backward(xi), 10, 3)
    plain = joint2operators(espirit(torch.as_tensor(cfg["maps"], device=dev)), nuf, fused=False)
    ms_pair_unfused = timed(lambda: plain.backward(plain.forward(xi)), 10, 3)
    # one untimed run of each solver first (first-call allocations), then the timed run
    S.prox_l2_Afxnb_CGD2(A.forward, A.backward, b, 30)
    S.ADMM_l2Afxnb_tvx(A.forward, A.backward, b, 20, 0.5, 1e-3, 1e-2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cg = S.prox_l2_Afxnb_CGD2(A.forward, A.backward, b, 30)
    torch.cuda.synchronize()
    t_cg = time.perf_counter() - t0
    t0 = time.perf_counter()
    admm = S.ADMM_l2Afxnb_tvx(A.forward, A.backward, b, 20, 0.5, 1e-3, 1e-2)
    torch.cuda.synchronize()
    t_admm = time.perf_counter() - t0
    out = {"workload": "256^2 Shepp-Logan x 32 Gaussian-blob coils, 48 x 2048 variable-density spiral, noise 1e-3; "
                       "A = joint2operators(espirit, NUFFT2d), rat 2, nspread 3",
           "ms_operator_pair": ms_pair, "ms_operator_pair_unfused_sense": ms_pair_unfused,
           "ms_forward_backward": ms_normal, "cg30_s": t_cg, "admm20_s": t_admm,
           "note": "cg30_s: prox_l2_Afxnb_CGD2 with Nite=30, which ends early at the reference's stop test "
                   "(delta <= 1e-3 delta0, pics/proximal_func.py:314-343); admm20_s: 20 outer ADMM-TV iterations",
           "coil_samples_per_s_pair": x.shape[0] * 32 / (ms_pair * 1e-3)}
    ref_path = os.path.join(ROOT, "tests", "golden", "solvers_spiral_cs.npz")
    if os.path.exists(ref_path):
        ref = np.load(ref_path)

        def rel(a, r):
            a = a.detach().cpu().numpy().astype(np.complex128).ravel()
            r = np.asarray(r, np.complex128).ravel()
            return float(np.linalg.norm(a - r) / np.linalg.norm(r))
        out["rel_l2_vs_fp64_runs"] = {"cg": rel(cg, ref["cg"]), "admm": rel(admm, ref["admm"]),
                                      "note": "committed FP64 oracle runs (tests/golden/solvers_spiral_cs.npz) "
                                              "whose k-space was simulated in FP64; here it is simulated on the GPU"}
    return out


def batched_uneven(dev, steps, warmup, tune=True):
    """64 uneven 2-D instances (N in {128,192,256,320}, U[32,400] spokes x 2N, U[1,32]
    coils, default_rng(4)): fwd+adj of every instance as 64 plain plans, as one
    batched execution with autotuned tiles, and as a batched graph replay."""
    import numpy as np
    import torch

    from mripy_amd import data
    from mripy_amd._lib import GN_TYPE1, GN_TYPE2, GN_TYPE21
    from mripy_amd.batch import BatchPlan
    from mripy_amd.plan import Plan

    inst = data.config_batched_uneven()
    n = len(inst)
    plans = [Plan(i["modes"], ncoils=i["ncoils"], rat=2, nspread=3, device=dev).set_points(i["coords"]) for i in inst]
    imgs = [torch.as_tensor(i["image"], device=dev) for i in inst]

    def bufs():
        return ([torch.empty(p.sample_shape(), dtype=torch.complex64, device=dev) for p in plans],
                [torch.empty(p.image_shape(), dtype=torch.complex64, device=dev) for p in plans])

    ys_p, ims_p = bufs()

    def plain():
        for i in range(n):
            plans[i].execute_into(GN_TYPE2, imgs[i], ys_p[i])
            plans[i].execute_into(GN_TYPE1, ys_p[i], ims_p[i])

    res, out = {}, {}
    tuned = None
    for name, graph in (("batched", False), ("graph", True)):
        B = BatchPlan(inst, rat=2, nspread=3, device=dev, graph=False)
        ys, ims = bufs()
        if tune:
            if tuned is None:
                tuned = B.autotune(GN_TYPE21, imgs, ims)
            else:
                B.set_tiles(*tuned[0])
        B.set_graph(graph)

        def run(B=B, ys=ys, ims=ims):
            B.execute_into(GN_TYPE2, imgs, ys)
            B.execute_into(GN_TYPE1, ys, ims)
        out[name] = (B, ys, ims, run)
    for name, fn in (("plain", plain), ("batched", out["batched"][3]), ("graph", out["graph"][3])):
        res[name] = timed(fn, steps, warmup)

    def rel(a, b):
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    worst = 0.0
    for name in ("batched", "graph"):
        _, ys, ims, _ = out[name]
        for i in range(n):
            worst = max(worst, rel(ys[i], ys_p[i]), rel(ims[i], ims_p[i]))
    info = out["batched"][0].info()
    coil_samples = sum(p.M * p.ncoils for p in plans)
    path_bytes = 0
    for p in plans:
        N, G = int(np.prod(p.modes)), int(np.prod(p.nf))
        path_bytes += sum(algorithmic_bytes(2, N, G, p.M, p.ncoils).values())
    best = min(res["batched"], res["graph"])
    return {
        "workload": "64 uneven 2-D instances: N in {128,192,256,320}, U[32,400] golden-angle spokes x 2N, "
                    "U[1,32] coils (default_rng(4)), rat 2, nspread 3, fwd+adj",
        "coil_samples": coil_samples,
        "plain_ms": res["plain"], "batched_ms": res["batched"], "graph_ms": res["graph"],
        "speedup_batched": res["plain"] / res["batched"], "speedup_graph": res["plain"] / res["graph"],
        "max_rel_vs_plain": worst,
        "launches_per_fwd_adj": {"plain": n * 2 * 3, "batched": info["launches_type2"] + info["launches_type1"]},
        "tiles": [list(info["tiles"]), list(info["gather_tiles"])],
        "autotuned": tuned is not None, "autotune_ms": tuned[1] if tuned else None,
        "path_roofline": {"algorithmic_bytes": path_bytes,
                          "frac_batched": path_bytes / (best * 1e-3) / HBM_PEAK,
                          "frac_plain": path_bytes / (res["plain"] * 1e-3) / HBM_PEAK},
    }


# ----------------------------------------------------------------------------- CPU baseline

def cpu_baseline(cfg, gpu_out, reps=5):
    """The FP64 C oracle (the reference loop nests restated, oracle/gridding.c) timed on
    the host cores: one warm-up, then the median of ``reps`` runs.

    * 3-D (the ``value`` workload): one fwd+adj of all 1,048,576 samples, coils over
      threads like the reference's prange, scipy FFT with cores-1 workers; its FP64
      results also give the rel-L2 of the GPU's timed results;
    * 3-D at nspread 5: every 8th spoke, gridding time scaled back to all samples;
    * 2-D 128^2 radial (128 x 256), single coil: single-threaded gridding like
      build_grid_2d*_fast, numpy FFT like fft/nufft_func.py:723,770."""
    import numpy as np

    from mripy_amd import data
    from oracle import coracle as C
    from oracle import nufft_oracle as O

    cores = C.cores()
    C.lib()
    modes = cfg["modes"]
    nc = cfg["image"].shape[-1]
    threads = min(nc, cores)   # numba prange over coils (fft/nufft_func.py:935,1036)

    def fwd_adj(coords, p, img, threads, fft_workers):
        t0 = time.perf_counter()
        grid = C.forward_grid(img, p, fft_workers)
        t1 = time.perf_counter()
        y = C.gather(coords, grid, p.nf, p.ns, p.tau, threads)
        t2 = time.perf_counter()
        g1 = C.spread(coords, y, p.nf, p.ns, p.tau, threads)
        t3 = time.perf_counter()
        F = C._fftn(g1, tuple(range(p.d)), True, fft_workers)
        out = O.crop(F, p.modes) * O.deapod(p.modes, p.tau)[..., None] / coords[0].shape[0]
        t4 = time.perf_counter()
        return (t1 - t0) + (t4 - t3), (t3 - t1), y, out

    def median(fn):
        fn()
        runs = sorted((fn() for _ in range(reps)), key=lambda r: r[0] + r[1])
        return runs[len(runs) // 2]

    res = {}
    # 3-D, nspread 3, full workload
    coords = [c.astype(np.float64) for c in cfg["coords"]]
    M = coords[0].shape[0]
    p = O.Params(modes, rat=2, nspread=3)
    img = cfg["image"].astype(np.complex128)
    t_fixed, t_grid, y64, out64 = median(lambda: fwd_adj(coords, p, img, threads, max(1, cores - 1)))
    t3d = t_fixed + t_grid
    if gpu_out is not None:
        _, _, ybuf, obuf = gpu_out
        yg = ybuf.cpu().numpy()
        og = obuf.cpu().numpy()
        res["rel_l2_vs_fp64"] = {
            "type2": float(np.linalg.norm(yg - y64) / np.linalg.norm(y64)),
            "fwd_adj": float(np.linalg.norm(og - out64) / np.linalg.norm(out64)),
            "note": "the timed GPU outputs (forward samples, and the adjoint of those samples) against the FP64 "
                    "C oracle's forward and adjoint of the same float32 inputs"}
    del y64, out64
    # 3-D, nspread 5, every 8th spoke
    frac = 8
    sub = [c.reshape(-1, 256)[::frac].reshape(-1) for c in coords]
    p5 = O.Params(modes, rat=2, nspread=5)
    f5, g5, _, _ = median(lambda: fwd_adj(sub, p5, img, threads, max(1, cores - 1)))
    t5 = f5 + g5 * frac
    # 2-D cfg1
    c1 = data.config_2d_radial()
    co2 = [c.astype(np.float64) for c in c1["coords"]]
    p2 = O.Params((128, 128), rat=2, nspread=3)
    im2 = c1["image"][..., None].astype(np.complex128)

    def fwd_adj_2d():
        t0 = time.perf_counter()
        g = O.place(im2 * O.deapod(p2.modes, p2.tau)[..., None], p2.nf)
        g = np.fft.fftn(g, axes=(0, 1))
        t1 = time.perf_counter()
        yy = C.gather(co2, g, p2.nf, p2.ns, p2.tau)
        g1 = C.spread(co2, yy, p2.nf, p2.ns, p2.tau)
        t2 = time.perf_counter()
        F = np.fft.ifftn(g1, axes=(0, 1))
        o = O.crop(F, p2.modes) * O.deapod(p2.modes, p2.tau)[..., None] / co2[0].shape[0]
        t3 = time.perf_counter()
        del o
        return (t1 - t0) + (t3 - t2), (t2 - t1)
    f2, g2 = median(fwd_adj_2d)
    t2d = f2 + g2
    res.update({
        "value": M * nc / t3d,
        "unit": "coil-samples/s",
        "cores": cores,
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": ("one fwd+adj of the full 3-D workload (all %d samples, 8 coils, nspread 3) by the FP64 C "
                   "restatement of the reference loop nests (oracle/gridding.c): gridding with %d threads over "
                   "coils, 256^3x8 FFTs (scipy, %d workers), pad/crop; median of %d after 1 warm-up: %.2fs "
                   "gridding + %.2fs fft/pre/post" % (M, threads, max(1, cores - 1), reps, t_grid, t_fixed)),
        "legs": {
            "3d_nspread3": {"s_per_fwd_adj": t3d, "coil_samples_per_s": M * nc / t3d},
            "3d_nspread5": {"s_per_fwd_adj": t5, "coil_samples_per_s": M * nc / t5,
                            "sample": "every %dth spoke (%d samples), gridding time x%d" % (frac, sub[0].shape[0],
                                                                                           frac)},
            "2d_radial_128": {"s_per_fwd_adj": t2d, "coil_samples_per_s": co2[0].shape[0] / t2d,
                              "threads": 1, "sample": "128^2, 128 x 256 golden-angle radial, 1 coil, nspread 3; "
                                                      "single-threaded gridding, numpy FFT"},
        },
    })
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--nspread", type=int, default=3)
    ap.add_argument("--tiles", type=str, default=None, help="spread tile shape, e.g. 4,4,8")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched uneven-instance leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the main workload (skip nspread 5, differentiable, CG-SENSE, copy legs)")
    ap.add_argument("--copy", action="store_true",
                    help="with --no-extra: still run the 1 GiB copy leg (the counter calibration kernel)")
    ap.add_argument("--only-batched", action="store_true",
                    help="run only the batched uneven-instance leg (prints its JSON object)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the NUFFT path shards by coil with no exchange, and this "
              "build runs on a single MI355X (requested --gpus %d, WORLD_SIZE %d)." % (args.gpus, world),
              file=sys.stderr)
        sys.exit(2)

    import torch

    from mripy_amd import data

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.only_batched:
        log("batched uneven instances only")
        print(json.dumps(batched_uneven(dev, steps=args.steps, warmup=args.warmup)))
        return

    log("generating the 3-D UTE workload")
    cfg = data.config_3d_ute(n=128, nc=8, nspokes=4096, nread=256)
    M, nc = cfg["coords"][0].shape[0], 8
    main_leg = leg_3d(cfg, dev, args.nspread, args.steps, args.warmup,
                      tiles=[int(t) for t in args.tiles.split(",")] if args.tiles else None, keep=True)
    ms_step = main_leg["ms_per_step"]
    log("timed %d steps: %.3f ms/step (set_points %.2f ms)" % (args.steps, ms_step, main_leg["set_points_ms"]))
    stages, ab = main_leg["stages"], main_leg["ab"]
    dom = max(stages, key=lambda k: stages[k]["ms"])
    dur = stages[dom]["ms"] * 1e-3
    achieved = ab[dom] / dur
    path_bytes = main_leg["path_bytes"]
    kernels = {"pad": "k_fft_rows_pad (pass A)", "fft2": "k_fft_cols passes B, C", "gather": "k_gather",
               "spread": "k_spread_mfma + k_spread_reduce", "fft1": "k_fft_cols passes C', B'",
               "crop": "k_fft_rows_crop (pass A')"}
    out = {
        "metric": METRIC,
        "value": main_leg["value"],
        "unit": "coil-samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": "3-D UTE radial NUFFT fwd+adj: 128^3 image x 8 coils, 4096x256 Fibonacci "
                               "spokes (M=1048576), rat 2, nspread %d (%d taps/axis), complex64"
                               % (args.nspread, 2 * args.nspread),
                   "modes": list(cfg["modes"]), "ncoils": nc, "samples": M, "nspread": args.nspread,
                   "tiles": main_leg["tiles"], "gather_tiles": main_leg["gather_tiles"],
                   "parallelism": "single GPU"},
        "path_roofline": {"algorithmic_bytes": path_bytes, "achieved_GBps": path_bytes / (ms_step * 1e-3) / 1e9,
                          "frac": path_bytes / (ms_step * 1e-3) / HBM_PEAK},
        "roofline": {"kernel": dom, "kernels": kernels[dom], "bound": "hbm", "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic(dom, main_leg["tiles"], args.nspread),
                     "bytes_per_launch": ab[dom], "avg_launch_ms": stages[dom]["ms"]},
        "stages": stages,
        "set_points_ms": main_leg["set_points_ms"],
        "type21_ms": main_leg["type21_ms"],
        "wall_s_timed": main_leg["wall_s_timed"],
    }
    keep = main_leg.pop("_keep")
    if not args.no_extra or args.copy:
        log("copy bandwidth")
        cp = copy_bandwidth(dev)
        out["copy"] = cp
        out["roofline"]["frac_vs_copy"] = achieved / (cp["GBps"] * 1e9)
        out["path_roofline"]["frac_vs_copy"] = path_bytes / (ms_step * 1e-3) / (cp["GBps"] * 1e9)
    if not args.no_extra:
        log("nspread 5")
        l5 = leg_3d(cfg, dev, 5, max(5, args.steps // 2), 2)
        out["nspread5"] = {"workload": "as the main workload at nspread 5 (10 taps/axis; the operator default "
                                       "eps = 1e-5)", "ms_per_step": l5["ms_per_step"], "value": l5["value"],
                           "stages": l5["stages"], "set_points_ms": l5["set_points_ms"],
                           "tiles": l5["tiles"], "gather_tiles": l5["gather_tiles"]}
        log("differentiable")
        out["differentiable"] = leg_differentiable(dev, max(5, args.steps // 2), 2)
        log("CG-SENSE / ADMM-TV")
        out["cg_sense"] = leg_cg_sense(dev)
    if not args.no_batched:
        log("batched uneven instances")
        out["batched_uneven"] = batched_uneven(dev, steps=max(3, args.steps // 4), warmup=2)
        log("batched: plain %.2f ms, batched %.2f ms, graph %.2f ms"
            % (out["batched_uneven"]["plain_ms"], out["batched_uneven"]["batched_ms"],
               out["batched_uneven"]["graph_ms"]))
    if not args.no_cpu:
        log("CPU baseline")
        try:
            out["cpu_baseline"] = cpu_baseline(cfg, keep if args.nspread == 3 else None)
            rl = out["cpu_baseline"].pop("rel_l2_vs_fp64", None)
            if rl is not None:
                out["rel_l2_vs_fp64"] = rl
        except Exception as e:  # the baseline is reported, never gating
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
