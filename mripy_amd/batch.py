"""Many NUFFT instances of uneven sizes in one set of launches (include/gaussnufft_batch.h).

The reference runs such workloads one operator call at a time
(``NUFFT3d_cuda.forward/backward``, pics/operators_cuda_class.py:107-120), each call
launching its own kernels and FFT plans.  ``BatchPlan`` keeps one plan per instance
inside one C batch: every stage (gather, spread, split-tile reduce, and the pruned
FFT passes fused with pad/crop, one launch per grid length) is one launch over all
instances' work items, and with ``graph=True`` the whole sequence is replayed from a
captured HIP graph.  Instance i's result equals what ``Plan`` computes for it.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import GN_ADJOINT, GN_COIL_LAST, GN_COIL_MAJOR, GN_TYPE1, GN_TYPE2, GN_TYPE21, check


class BatchPlan:
    """instances: list of dicts with ``modes``, ``coords`` (d arrays of M radians),
    ``ncoils`` and optionally ``dcf``.  All share dim, rat, nspread, iflag, layout, df."""

    def __init__(self, instances, rat=2, nspread=3, iflag=1, layout="coil_last", df=1.0, device=None, graph=False):
        if not instances:
            raise ValueError("a batch needs at least one instance")
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.dim = len(instances[0]["modes"])
        self.layout = GN_COIL_LAST if layout in ("coil_last", GN_COIL_LAST) else GN_COIL_MAJOR
        self.modes, self.ncoils, self.M = [], [], []
        f64 = all(isinstance(c, np.ndarray) and c.dtype == np.float64 for ins in instances for c in ins["coords"])
        dt = torch.float64 if f64 else torch.float32
        keep = []
        arr = (_lib.GNInstance * len(instances))()
        for i, ins in enumerate(instances):
            modes = tuple(int(m) for m in ins["modes"])
            if len(modes) != self.dim:
                raise ValueError("all instances must have the same dimension")
            ts = [torch.as_tensor(c).reshape(-1).to(device=self.device, dtype=dt).contiguous() for c in ins["coords"]]
            if len(ts) != self.dim or any(t.numel() != ts[0].numel() for t in ts):
                raise ValueError("instance %d: need %d coordinate arrays of equal length" % (i, self.dim))
            d = ins.get("dcf")
            if d is not None:
                d = torch.as_tensor(d).reshape(-1).to(device=self.device, dtype=torch.float32).contiguous()
            keep.append((ts, d))
            e = arr[i]
            for a in range(3):
                e.modes[a] = modes[a] if a < self.dim else 1
            e.ncoils = int(ins.get("ncoils", 1))
            e.M = ts[0].numel()
            e.x, e.y = ts[0].data_ptr(), ts[1].data_ptr()
            e.z = ts[2].data_ptr() if self.dim == 3 else None
            e.dcf = d.data_ptr() if d is not None else None
            self.modes.append(modes)
            self.ncoils.append(e.ncoils)
            self.M.append(e.M)
        self.n = len(instances)
        self._h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            check(_lib.load().gn_batch_create(self.dim, self.n, arr, int(rat), int(nspread), int(iflag), self.layout,
                                              1 if f64 else 0, float(df), ctypes.c_void_p(stream),
                                              ctypes.byref(self._h)))
        self.set_graph(graph)
        self._stream = stream

    # ------------------------------------------------------------------ shapes
    def image_shape(self, i):
        if self.layout == GN_COIL_LAST:
            return self.modes[i] + (self.ncoils[i],)
        return (self.ncoils[i],) + self.modes[i]

    def sample_shape(self, i):
        if self.layout == GN_COIL_LAST:
            return (self.M[i], self.ncoils[i])
        return (self.ncoils[i], self.M[i])

    # ------------------------------------------------------------------ control
    def set_graph(self, enable=True):
        check(_lib.load().gn_batch_set_graph(self._h, 1 if enable else 0))
        self.graph = bool(enable)

    def set_tiles(self, tiles=None, gather_tiles=None):
        sp = list(tiles or []) + [0] * (3 - len(tiles or []))
        ga = list(gather_tiles or []) + [0] * (3 - len(gather_tiles or []))
        check(_lib.load().gn_batch_set_tiles(self._h, (ctypes.c_int * 6)(*(sp + ga))))

    def info(self):
        n = ctypes.c_int()
        t = (ctypes.c_int * 6)()
        la = (ctypes.c_int * 2)()
        check(_lib.load().gn_batch_info(self._h, ctypes.byref(n), t, la))
        return {"instances": n.value, "tiles": tuple(t[:self.dim]), "gather_tiles": tuple(t[3:3 + self.dim]),
                "launches_type2": la[0], "launches_type1": la[1]}

    @staticmethod
    def _ptrs(ts):
        return (ctypes.c_void_p * len(ts))(*[t.data_ptr() if t is not None and t.numel() else None for t in ts])

    def _check_stream(self):
        if torch.cuda.current_stream(self.device).cuda_stream != self._stream:
            raise RuntimeError("BatchPlan runs on the stream that was current at creation")

    def execute_into(self, kind, srcs, dsts):
        """Run ``kind`` on lists of contiguous complex64 device tensors (no allocation)."""
        self._check_stream()
        if len(srcs) != self.n or len(dsts) != self.n:
            raise ValueError("need one input and one output per instance")
        check(_lib.load().gn_batch_execute(self._h, kind, self._ptrs(srcs), self._ptrs(dsts)))
        return dsts

    def autotune(self, kind, srcs, dsts, reps=3):
        """Time every candidate tile shape on ``kind`` and keep the fastest; returns
        ((spread tiles, gather tiles), ms per batched execution)."""
        self._check_stream()
        t = (ctypes.c_int * 6)()
        ms = ctypes.c_double()
        check(_lib.load().gn_batch_autotune(self._h, kind, self._ptrs(srcs), self._ptrs(dsts), int(reps), t,
                                            ctypes.byref(ms)))
        return (tuple(t[:self.dim]), tuple(t[3:3 + self.dim])), ms.value

    def _run(self, kind, xs, in_shape, out_shape):
        if len(xs) != self.n:
            raise ValueError("need one array per instance")
        ins = [torch.as_tensor(x).to(device=self.device, dtype=torch.complex64).reshape(in_shape(i)).contiguous()
               for i, x in enumerate(xs)]
        outs = [torch.empty(out_shape(i), dtype=torch.complex64, device=self.device) for i in range(self.n)]
        self.execute_into(kind, ins, outs)
        return outs

    def type2(self, images):
        return self._run(GN_TYPE2, images, self.image_shape, self.sample_shape)

    def type1(self, samples):
        return self._run(GN_TYPE1, samples, self.sample_shape, self.image_shape)

    def adjoint(self, samples):
        return self._run(GN_ADJOINT, samples, self.sample_shape, self.image_shape)

    def type21(self, images):
        return self._run(GN_TYPE21, images, self.image_shape, self.image_shape)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().gn_batch_destroy(h)
            except Exception:
                pass
            self._h = ctypes.c_void_p()
