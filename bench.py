#!/usr/bin/env python
"""Throughput of one forward + one adjoint Gaussian-gridding NUFFT on one MI355X.

Workload (default): 3-D multi-coil UTE -- 128^3 complex64 ellipsoid phantom times
8 Gaussian-blob coil maps, 4096 Fibonacci-sphere spokes x 256 centre-out points
(1,048,576 samples), oversampling 2, 6 taps per axis (nspread 3), complex64.
A step is forward (type 2) of the image followed by the adjoint (type 1) of the
resulting samples, on device-resident data.  value = coil-samples per second.

Prints ONE JSON line.  Per-stage times come from HIP events recorded on the
plan's stream inside the timed region; ``roofline`` is computed for the
dominant hand-written kernel from its algorithmic bytes per launch.
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


def cpu_baseline(cfg, modes, nc, fraction, reps=3):
    """Time the FP64 C oracle (the reference loop nests restated) on the same workload:
    every ``fraction``-th spoke (1 = all), full grids and FFTs; one warm-up pass, then
    the median of ``reps`` passes.  The gridding time is scaled back to the full
    sample count when a fraction is used."""
    import numpy as np

    from oracle import coracle as C
    from oracle import nufft_oracle as O

    cores = C.cores()
    nread = 256
    coords = [c.reshape(-1, nread)[::fraction].reshape(-1).astype(np.float64) for c in cfg["coords"]]
    Ms = coords[0].shape[0]
    p = O.Params(modes, rat=2, nspread=3)
    img = cfg["image"].astype(np.complex128)
    threads = min(nc, cores)   # numba prange over coils (fft/nufft_func.py:935,1036)
    C.lib()

    def one():
        t0 = time.perf_counter()
        F2 = img * O.deapod(p.modes, p.tau)[..., None]
        grid = O.place(F2, p.nf)
        grid = C._fftn(grid, (0, 1, 2), False, max(1, cores - 1))
        t1 = time.perf_counter()
        y = C.gather(coords, grid, p.nf, p.ns, p.tau, threads)
        t2 = time.perf_counter()
        g1 = C.spread(coords, y, p.nf, p.ns, p.tau, threads)
        t3 = time.perf_counter()
        F = C._fftn(g1, (0, 1, 2), True, max(1, cores - 1))
        out = O.crop(F, p.modes) * O.deapod(p.modes, p.tau)[..., None] / Ms
        t4 = time.perf_counter()
        del out
        return (t1 - t0) + (t4 - t3), (t3 - t1)

    one()
    runs = sorted((one() for _ in range(reps)), key=lambda r: r[0] + r[1])
    t_fixed, t_grid = runs[len(runs) // 2]
    M_full = cfg["coords"][0].shape[0]
    t_full = t_fixed + t_grid * (M_full / Ms)
    what = "all %d samples" % M_full if fraction == 1 else "%d of %d samples (every %dth spoke) scaled x%d" % (
        Ms, M_full, fraction, fraction)
    return {
        "value": M_full * nc / t_full,
        "unit": "coil-samples/s",
        "cores": cores,
        "kind": "port",
        "sample": ("one fwd+adj of the full workload by the FP64 C restatement of the reference loop nests "
                   "(oracle/gridding.c): gridding on %s with %d threads over coils, 256^3x%d FFTs (scipy, "
                   "%d workers), pad/crop; median of %d after 1 warm-up: %.2fs gridding + %.2fs fft/pre/post"
                   % (what, threads, nc, max(1, cores - 1), reps, t_grid, t_fixed)),
    }


def batched_uneven(dev, steps, warmup):
    """64 uneven 2-D instances (N in {128,192,256,320}, U[32,400] spokes x 2N, U[1,32]
    coils, default_rng(4)): fwd+adj of every instance as 64 plain plans, as one
    batched execution, and as a batched graph replay; per-instance agreement."""
    import numpy as np
    import torch

    from mripy_amd import data
    from mripy_amd._lib import GN_TYPE1, GN_TYPE2
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

    res = {}
    out = {}
    for name, graph in (("batched", False), ("graph", True)):
        B = BatchPlan(inst, rat=2, nspread=3, device=dev, graph=graph)
        ys, ims = bufs()

        def run(B=B, ys=ys, ims=ims):
            B.execute_into(GN_TYPE2, imgs, ys)
            B.execute_into(GN_TYPE1, ys, ims)
        out[name] = (B, ys, ims, run)
    for name, fn in (("plain", plain), ("batched", out["batched"][3]), ("graph", out["graph"][3])):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        res[name] = e0.elapsed_time(e1) / steps

    def rel(a, b):
        return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    worst = 0.0
    for name in ("batched", "graph"):
        _, ys, ims, _ = out[name]
        for i in range(n):
            worst = max(worst, rel(ys[i], ys_p[i]), rel(ims[i], ims_p[i]))
    info = out["batched"][0].info()
    coil_samples = sum(p.M * p.ncoils for p in plans)
    return {
        "workload": "64 uneven 2-D instances: N in {128,192,256,320}, U[32,400] golden-angle spokes x 2N, "
                    "U[1,32] coils (default_rng(4)), rat 2, nspread 3, fwd+adj",
        "coil_samples": coil_samples,
        "plain_ms": res["plain"], "batched_ms": res["batched"], "graph_ms": res["graph"],
        "speedup_batched": res["plain"] / res["batched"], "speedup_graph": res["plain"] / res["graph"],
        "max_rel_vs_plain": worst,
        "launches_per_fwd_adj": {"plain": n * 2 * 3, "batched": info["launches_type2"] + info["launches_type1"]},
        "tiles": [list(info["tiles"]), list(info["gather_tiles"])],
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-fraction", type=int, default=1)
    ap.add_argument("--nspread", type=int, default=3)
    ap.add_argument("--tiles", type=str, default=None, help="spread tile shape, e.g. 4,4,8")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched uneven-instance leg")
    ap.add_argument("--only-batched", action="store_true",
                    help="run only the batched uneven-instance leg (prints its JSON object)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one plain 1 GiB read + 1 GiB write streaming kernel first (HBM counter calibration)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the NUFFT path shards by coil with no exchange, and this "
              "build runs on a single MI355X (requested --gpus %d, WORLD_SIZE %d)." % (args.gpus, world),
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from mripy_amd import data
    from mripy_amd.plan import Plan

    if args.only_batched:
        dev = torch.device("cuda:0")
        torch.cuda.set_device(dev)
        log("batched uneven instances only")
        print(json.dumps(batched_uneven(dev, steps=args.steps, warmup=args.warmup)))
        return

    modes, nc, nspokes, nread = (128, 128, 128), 8, 4096, 256
    log("generating the 3-D UTE workload")
    cfg = data.config_3d_ute(n=128, nc=nc, nspokes=nspokes, nread=nread)
    d, N, G = 3, int(np.prod(modes)), 256 ** 3
    M = cfg["coords"][0].shape[0]

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)
    P = Plan(modes, ncoils=nc, rat=2, nspread=args.nspread, device=dev,
             tiles=[int(t) for t in args.tiles.split(",")] if args.tiles else None)
    coords = [torch.as_tensor(c, device=dev) for c in cfg["coords"]]
    P.set_points(coords)
    img = torch.as_tensor(cfg["image"], device=dev).contiguous()
    ybuf = torch.empty((M, nc), dtype=torch.complex64, device=dev)
    obuf = torch.empty(modes + (nc,), dtype=torch.complex64, device=dev)
    from mripy_amd._lib import GN_TYPE1, GN_TYPE2

    def step():
        P.execute_into(GN_TYPE2, img, ybuf)
        P.execute_into(GN_TYPE1, ybuf, obuf)

    if args.copy_probe:
        # reads exactly COPY_BYTES and writes exactly COPY_BYTES: the reference point for
        # scaling FETCH_SIZE / WRITE_SIZE counter readings to bytes
        a = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        torch.mul(a, 2.0, out=b)
        torch.cuda.synchronize()
        del a, b
    log("set_points done; warm-up")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    P.set_timing(True)
    P.stage_times(reset=True)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = P.stage_times(reset=True)
    P.set_timing(False)
    log("timed %d steps: %.3f ms/step" % (args.steps, ev0.elapsed_time(ev1) / args.steps))
    ms_step = ev0.elapsed_time(ev1) / args.steps

    ab = algorithmic_bytes(d, N, G, M, nc, modes, P.nf)
    stages = {k: {"ms": v[0] / max(1, v[1]), "launches": v[1]} for k, v in st.items()}
    own = ["pad", "gather", "spread", "crop"]
    dom = max(own, key=lambda k: stages[k]["ms"])
    dur = stages[dom]["ms"] * 1e-3
    achieved = ab[dom] / dur
    path_bytes = sum(algorithmic_bytes(d, N, G, M, nc).values())  # the fixed fwd+adj byte model
    for k in stages:
        stages[k]["GBps"] = ab[k] / (stages[k]["ms"] * 1e-3) / 1e9 if stages[k]["ms"] > 0 else None

    out = {
        "metric": "NUFFT fwd+adj samples/s & % HBM roofline; batched/diff vs plain at <=1e-5 rel-L2",
        "value": M * nc / (ms_step * 1e-3),
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
                   "modes": list(modes), "ncoils": nc, "samples": M, "nspread": args.nspread,
                   "tiles": list(P.tiles), "parallelism": "single GPU"},
        "path_roofline": {"algorithmic_bytes": path_bytes, "achieved_GBps": path_bytes / (ms_step * 1e-3) / 1e9,
                          "frac": path_bytes / (ms_step * 1e-3) / HBM_PEAK},
        "roofline": {"kernel": dom, "bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": measured_traffic(dom, list(P.tiles), args.nspread),
                     "bytes_per_launch": ab[dom], "avg_launch_ms": stages[dom]["ms"]},
        "stages": stages,
        "wall_s_timed": wall,
    }
    if not args.no_batched:
        log("batched uneven instances")
        del ybuf, obuf, img, P
        torch.cuda.empty_cache()
        out["batched_uneven"] = batched_uneven(dev, steps=max(3, args.steps // 4), warmup=2)
        log("batched: plain %.2f ms, batched %.2f ms, graph %.2f ms"
            % (out["batched_uneven"]["plain_ms"], out["batched_uneven"]["batched_ms"],
               out["batched_uneven"]["graph_ms"]))
    if not args.no_cpu:
        log("CPU baseline")
        try:
            out["cpu_baseline"] = cpu_baseline(cfg, modes, nc, args.cpu_fraction)
        except Exception as e:  # the baseline is reported, never gating
            out["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
