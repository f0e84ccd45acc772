"""Plan: the Python owner of one C-ABI gaussnufft plan, working on torch tensors.

A plan fixes the image size, coil count, kernel (rat, nspread), sign and layout;
``set_points`` bins a trajectory; ``type2``/``type1``/``adjoint``/``type21`` run
on the current torch stream of the plan's device.  Inputs may be torch tensors
(device-resident, returned as device tensors, asynchronous) or numpy arrays
(copied to the device; the result is copied back, synchronously), mirroring the
numpy-in/numpy-out behaviour of the reference functions.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import (GN_ADJOINT, GN_COIL_LAST, GN_COIL_MAJOR, GN_GRID_GATHER, GN_GRID_SPREAD, GN_SENSE_ADJOINT,
                   GN_SENSE_TYPE1, GN_SENSE_TYPE2, GN_SENSE_TYPE21, GN_TYPE1, GN_TYPE2, GN_TYPE21, check)


def grid_params(modes, eps, max_nspread=None):
    """(rat, nspread, nf, tau) from eps by the reference rule (fft/nufft_func.py:198-249)."""
    L = _lib.load()
    dim = len(modes)
    m = (ctypes.c_int64 * 3)(*[int(v) for v in modes] + [1] * (3 - dim))
    rat, ns = ctypes.c_int(), ctypes.c_int()
    nf = (ctypes.c_int64 * 3)()
    tau = ctypes.c_double()
    check(L.gn_grid_params(dim, m, float(eps), int(max_nspread or 0), ctypes.byref(rat), ctypes.byref(ns), nf,
                           ctypes.byref(tau)))
    return rat.value, ns.value, tuple(nf[:dim]), tau.value


def _device(device):
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)


class Plan:
    """One NUFFT plan (C handle) bound to a device and its current stream."""

    def __init__(self, modes, ncoils=1, eps=None, rat=None, nspread=None, iflag=1, layout="coil_last",
                 device=None, max_nspread=None, tiles=None, gather_tiles=None):
        L = _lib.load()
        self.modes = tuple(int(m) for m in modes)
        self.dim = len(self.modes)
        if self.dim not in (2, 3):
            raise ValueError("only 2-D and 3-D transforms are supported")
        if eps is not None:
            if max_nspread is None and self.dim == 3:
                max_nspread = 6  # every 3-D reference driver passes max_nspread=6
            r, n, _, _ = grid_params(self.modes, eps, max_nspread)
            rat = r if rat is None else rat
            nspread = n if nspread is None else nspread
        if rat is None or nspread is None:
            raise ValueError("give eps or both rat and nspread")
        self.rat, self.nspread = int(rat), int(nspread)
        self.ncoils = int(ncoils)
        self.iflag = int(iflag)
        self.layout = GN_COIL_LAST if layout in ("coil_last", GN_COIL_LAST) else GN_COIL_MAJOR
        self.device = _device(device)
        self._h = ctypes.c_void_p()
        m = (ctypes.c_int64 * 3)(*list(self.modes) + [1] * (3 - self.dim))
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream(self.device).cuda_stream
            check(L.gn_plan_create(self.dim, m, self.rat, self.nspread, self.iflag, self.ncoils, self.layout,
                                   ctypes.c_void_p(stream), ctypes.byref(self._h)))
        if tiles is not None or gather_tiles is not None:
            self.set_tiles(tiles, gather_tiles)
        nf = (ctypes.c_int64 * 3)()
        ns = ctypes.c_int()
        tau = ctypes.c_double()
        check(L.gn_plan_info(self._h, nf, ctypes.byref(ns), ctypes.byref(tau), None, None, None))
        self.nf = tuple(nf[:self.dim])
        self.tau = tau.value
        self.M = None
        self._dcf = None
        self.df = 1.0

    # ------------------------------------------------------------------ setup
    def set_tiles(self, tiles=None, gather_tiles=None):
        """Spread (output) and gather (home) tile shapes in cells; None/0 keeps the default.
        Requires a new set_points afterwards."""
        sp = list(tiles or []) + [0] * (3 - len(tiles or []))
        ga = list(gather_tiles or []) + [0] * (3 - len(gather_tiles or []))
        t = (ctypes.c_int * 6)(*(sp + ga))
        check(_lib.load().gn_plan_set_tiles(self._h, t))
        self.M = None

    def _tile_info(self):
        t = (ctypes.c_int * 6)()
        check(_lib.load().gn_plan_info(self._h, None, None, None, None, None, t))
        return tuple(t[:self.dim]), tuple(t[3:3 + self.dim])

    @property
    def tiles(self):
        return self._tile_info()[0]

    @property
    def gather_tiles(self):
        return self._tile_info()[1]

    def _bind_stream(self):
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(_lib.load().gn_plan_set_stream(self._h, ctypes.c_void_p(stream)))

    def set_points(self, coords, df=1.0, dcf=None):
        """Bin sample coordinates (radians; d arrays of M).  float64 numpy inputs stay
        float64 for the window index; torch float32 tensors are used as they are."""
        if len(coords) != self.dim:
            raise ValueError("need %d coordinate arrays" % self.dim)
        ts = []
        f64 = all(isinstance(c, np.ndarray) and c.dtype == np.float64 for c in coords) or \
            all(torch.is_tensor(c) and c.dtype == torch.float64 for c in coords)
        dt = torch.float64 if f64 else torch.float32
        for c in coords:
            t = torch.as_tensor(c).reshape(-1).to(device=self.device, dtype=dt).contiguous()
            ts.append(t)
        M = ts[0].numel()
        if any(t.numel() != M for t in ts):
            raise ValueError("coordinate arrays differ in length")
        d = None
        if dcf is not None:
            d = torch.as_tensor(dcf).reshape(-1).to(device=self.device, dtype=torch.float32).contiguous()
            if d.numel() != M:
                raise ValueError("dcf must have one entry per sample")
        with torch.cuda.device(self.device):
            self._bind_stream()
            z = ts[2].data_ptr() if self.dim == 3 else None
            check(_lib.load().gn_set_points(self._h, M, ts[0].data_ptr(), ts[1].data_ptr(), z, 1 if f64 else 0,
                                            float(df), d.data_ptr() if d is not None else None))
        self.M = M
        self._dcf = d
        self.df = float(df)
        self._bound_coords = None     # the autograd functions' record of what is binned
        self._points = [c for c in coords]
        self._bound = (ts, f64, float(df), d)   # device copies, for re-binning by autotune
        return self

    # ------------------------------------------------------------------ shapes
    def image_shape(self):
        if self.layout == GN_COIL_LAST:
            return self.modes + (self.ncoils,)
        return (self.ncoils,) + self.modes

    def sample_shape(self):
        if self.layout == GN_COIL_LAST:
            return (self.M, self.ncoils)
        return (self.ncoils, self.M)

    # ------------------------------------------------------------------ execution
    def _run(self, kind, x, in_shape, out_shape):
        if self.M is None:
            raise RuntimeError("set_points must be called first")
        is_np = not torch.is_tensor(x)
        t = torch.as_tensor(x)
        n_in = int(np.prod(in_shape))
        if t.numel() != n_in:
            raise ValueError("input has %d elements, expected %s" % (t.numel(), in_shape))
        t = t.to(device=self.device, dtype=torch.complex64).reshape(in_shape).contiguous()
        out = torch.empty(out_shape, dtype=torch.complex64, device=self.device)
        with torch.cuda.device(self.device):
            self._bind_stream()
            check(_lib.load().gn_execute(self._h, kind, t.data_ptr() if t.numel() else None,
                                         out.data_ptr() if out.numel() else None))
        if is_np:
            return out.cpu().numpy()
        return out

    def type2(self, image):
        """Forward (reference type 2): image -> samples."""
        return self._run(GN_TYPE2, image, self.image_shape(), self.sample_shape())

    def type1(self, samples):
        """Reference type 1 with density weighting (if a dcf was set): samples -> image."""
        return self._run(GN_TYPE1, samples, self.sample_shape(), self.image_shape())

    def adjoint(self, samples):
        """Exact adjoint of type2 (no 1/M, no dcf)."""
        return self._run(GN_ADJOINT, samples, self.sample_shape(), self.image_shape())

    def type21(self, image):
        """type1(dcf * type2(image)): the normal operator of the reference."""
        return self._run(GN_TYPE21, image, self.image_shape(), self.image_shape())

    # ------------------------------------------------------------------ gridding alone
    def grid_spread(self, samples):
        """The gridding step of type 1 (build_grid_*1_fast, fft/nufft_func.py:550-582,
        863-909): samples -> oversampled grid (nc, *nf), no dcf, no scale."""
        return self._run(GN_GRID_SPREAD, samples, self.sample_shape(), (self.ncoils,) + self.nf)

    def grid_gather(self, grid):
        """The gridding step of type 2 (build_grid_*2_fast, fft/nufft_func.py:607-639,
        985-1032): oversampled grid (nc, *nf) -> samples, divided by G."""
        return self._run(GN_GRID_GATHER, grid, (self.ncoils,) + self.nf, self.sample_shape())

    # ------------------------------------------------------------------ SENSE
    def set_sense(self, maps):
        """Attach coil sensitivities (the plan's image layout with nc coils; copied into
        the plan) for the sense_* transforms; None detaches them."""
        if maps is None:
            check(_lib.load().gn_plan_set_sense(self._h, None))
            self._sense = None
            return self
        m = torch.as_tensor(maps).to(device=self.device, dtype=torch.complex64).reshape(self.image_shape())
        m = m.contiguous()
        with torch.cuda.device(self.device):
            self._bind_stream()
            check(_lib.load().gn_plan_set_sense(self._h, m.data_ptr()))
        self._sense = m      # kept until the stream has copied it
        return self

    def sense_type2(self, image):
        """type2(image * S): coil-less image (modes) -> samples of every coil."""
        return self._run(GN_SENSE_TYPE2, image, self.modes, self.sample_shape())

    def sense_type1(self, samples):
        """sum_c conj(S_c) type1(samples)_c: samples -> coil-less image (modes)."""
        return self._run(GN_SENSE_TYPE1, samples, self.sample_shape(), self.modes)

    def sense_adjoint(self, samples):
        """Exact adjoint of sense_type2."""
        return self._run(GN_SENSE_ADJOINT, samples, self.sample_shape(), self.modes)

    def sense_type21(self, image):
        """sense_type1(dcf * sense_type2(image))."""
        return self._run(GN_SENSE_TYPE21, image, self.modes, self.modes)

    def execute_into(self, kind, src, dst):
        """Run on preallocated contiguous complex64 device tensors (no allocation, no checks of shape)."""
        check(_lib.load().gn_execute(self._h, kind, src.data_ptr(), dst.data_ptr()))
        return dst

    def autotune(self, kind, src, dst, reps=3):
        """Time every candidate tile shape on ``kind`` (src/dst: contiguous complex64 device
        tensors) and keep the fastest; returns ((spread tiles, gather tiles), ms per run)."""
        if self.M is None:
            raise RuntimeError("set_points must be called first")
        ts, f64, df, d = self._bound
        t = (ctypes.c_int * 6)()
        ms = ctypes.c_double()
        with torch.cuda.device(self.device):
            self._bind_stream()
            z = ts[2].data_ptr() if self.dim == 3 else None
            check(_lib.load().gn_plan_autotune(self._h, kind, self.M, ts[0].data_ptr(), ts[1].data_ptr(), z,
                                               1 if f64 else 0, df, d.data_ptr() if d is not None else None,
                                               src.data_ptr(), dst.data_ptr(), int(reps), t, ctypes.byref(ms)))
        return (tuple(t[:self.dim]), tuple(t[3:3 + self.dim])), ms.value

    def traj_grad(self, image, weights, reuse_grid=False):
        """(dim, M) float32: sum_c Re(conj(w) d(type2 image)/dx) per sample and axis (analytic).

        reuse_grid: skip the FFT passes when the plan's last operation was type2 of this
        same (live, unmodified) image tensor, whose type-2 grid the plan still holds; the
        result is the same either way."""
        if self.M is None:
            raise RuntimeError("set_points must be called first")
        img = torch.as_tensor(image).to(device=self.device, dtype=torch.complex64).reshape(self.image_shape())
        w = torch.as_tensor(weights).to(device=self.device, dtype=torch.complex64).reshape(self.sample_shape())
        img = img.contiguous()
        w = w.contiguous()
        out = torch.zeros((self.dim, self.M), dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            self._bind_stream()
            check(_lib.load().gn_traj_grad_ex(self._h, img.data_ptr(), w.data_ptr(), out.data_ptr(),
                                              1 if reuse_grid else 0))
        return out

    def sample_order(self):
        """Original sample index of every sorted position (the counting-sort result)."""
        out = np.zeros(self.M or 0, np.int32)
        if self.M:
            check(_lib.load().gn_plan_sample_order(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return out

    def set_fft_engine(self, engine="auto"):
        """"auto": pruned FFT passes fused with pad/crop where every nf is a product of
        2, 3 and 5; "rocfft": separate pad, rocFFT, crop."""
        check(_lib.load().gn_plan_set_fft_engine(self._h, _lib.GN_FFT_ROCFFT if engine == "rocfft"
                                                 else _lib.GN_FFT_AUTO))

    def set_timing(self, enable=True):
        """Enable per-stage HIP event timers inside gn_execute."""
        check(_lib.load().gn_plan_set_timing(self._h, 1 if enable else 0))

    def stage_times(self, reset=True):
        """{stage: (total ms, launches)} accumulated since the last reset (synchronizes)."""
        ms = (ctypes.c_double * 6)()
        n = (ctypes.c_int64 * 6)()
        check(_lib.load().gn_plan_stage_times(self._h, ms, n, 1 if reset else 0))
        return {name: (ms[i], n[i]) for i, name in enumerate(_lib.STAGES)}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().gn_plan_destroy(h)
            except Exception:
                pass
            self._h = ctypes.c_void_p()
