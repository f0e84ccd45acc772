"""Operator classes with mripy's names and methods, running the NUFFT on the HIP path.

* ``NUFFT3d_hip`` / ``NUFFT2d_hip``: the reference ``NUFFT3d_cuda`` operator
  (pics/operators_cuda_class.py:64-127) and its CPU twin ``NUFFT3d``
  (pics/operators_class.py:302-364): same constructor arguments and methods
  (``normalize_set_ktraj``, ``set_ktraj``, ``density_weighting``, ``forward``,
  ``backward``, ``forward_backward``), plus ``adjoint`` (exact A^H, no 1/M, no
  dcf).  ``NUFFT3d_cuda`` and ``NUFFT3d`` are aliases of ``NUFFT3d_hip``.
* ``espirit`` (pics/operators_class.py:437-472), ``None_opt`` (:477-485),
  ``joint2operators`` (:496-514), ``joint3operators`` (:520-539) and
  ``dim_match`` (utilities/utilities_func.py:393-404).

Images are coil-last ``(ms, mt[, mu][, nc])`` and k-space data ``(M[, nc])`` as
in the reference.  numpy inputs give numpy outputs (host-synchronous, like the
reference); torch device tensors stay on the device.
"""
from __future__ import annotations

import numpy as np
import torch

from .plan import Plan, grid_params


def dim_match(A_shape, B_shape):
    """Pad the shorter shape with trailing 1s (utilities/utilities_func.py:393-404)."""
    A_out, B_out = tuple(A_shape), tuple(B_shape)
    if len(A_out) < len(B_out):
        A_out = A_out + (1,) * (len(B_out) - len(A_out))
    elif len(A_out) > len(B_out):
        B_out = B_out + (1,) * (len(A_out) - len(B_out))
    return A_out, B_out


def _xp(a):
    return torch if torch.is_tensor(a) else np


def _flat(a):
    if a is None:
        return None
    return a.reshape(-1) if torch.is_tensor(a) else np.asarray(a).flatten()


def _fingerprint(a):
    """Identity of an input's current contents.  Torch tensors: storage pointer and
    version counter (every in-place update bumps it, also through views).  numpy
    arrays: a 64-bit hash of the bytes, so in-place edits of the trajectory or the
    density weights are seen."""
    if a is None:
        return None
    if torch.is_tensor(a):
        return ("t", a.data_ptr(), a._version, tuple(a.shape), str(a.dtype), str(a.device))
    a = np.ascontiguousarray(a)
    try:
        import xxhash
        h = xxhash.xxh3_64_intdigest(memoryview(a).cast("B"))
    except ImportError:
        import hashlib
        h = hashlib.blake2b(memoryview(a).cast("B"), digest_size=8).hexdigest()
    return ("n", a.shape, str(a.dtype), h)


class _NUFFTnd:
    """Shared implementation of the d-dimensional operator."""

    dim = 3

    def __init__(self, im_shape=None, dcf=None, ktrajx=None, ktrajy=None, ktrajz=None, axes=None, eps=1e-5,
                 rat=None, nspread=None, iflag=1, df=1.0, device=None, tiles=None, gather_tiles=None):
        self.axes = axes if axes is not None else tuple(range(self.dim))
        self.ktrajx = _flat(ktrajx)
        self.ktrajy = _flat(ktrajy)
        self.ktrajz = _flat(ktrajz)
        self.im_shape = im_shape
        self.dcf = _flat(dcf)
        self.eps = eps
        self.iflag = iflag
        self.df = df
        self.device = device
        self.tiles = tiles
        self.gather_tiles = gather_tiles
        if rat is None or nspread is None:
            mx = 6 if self.dim == 3 else None   # 3-D drivers clamp nspread to [3, 6]
            r, n, _, _ = grid_params(tuple(im_shape[:self.dim]) if im_shape is not None else (1,) * self.dim,
                                     eps, mx)
            rat = r if rat is None else rat
            nspread = n if nspread is None else nspread
        self.rat, self.nspread = rat, nspread
        self._plan = None
        self._plan_key = None
        self._points_key = None

    # ---- trajectory ---------------------------------------------------------
    def normalize_set_ktraj(self, ktraj):
        """pi * k / max(k) (pics/operators_cuda_class.py:85-90)."""
        xp = _xp(ktraj)
        ktraj = np.pi * ktraj / ktraj.reshape(-1).max()
        self.ktrajx = _flat(ktraj[0])
        self.ktrajy = _flat(ktraj[1])
        if self.dim == 3:
            self.ktrajz = _flat(ktraj[2])
            return self.ktrajx, self.ktrajy, self.ktrajz
        del xp
        return self.ktrajx, self.ktrajy

    def set_ktraj(self, ktrajx, ktrajy, ktrajz=None, dcf=None):
        """(pics/operators_cuda_class.py:92-96)."""
        self.ktrajx = _flat(ktrajx)
        self.ktrajy = _flat(ktrajy)
        if self.dim == 3:
            self.ktrajz = _flat(ktrajz)
        elif ktrajz is not None and dcf is None:
            dcf = ktrajz  # 2-D call set_ktraj(x, y, dcf)
        self.dcf = _flat(dcf)
        self._points_key = None

    def set_params(self, rat=None, nspread=None):
        """Oversampling ratio and kernel half-width used from the next call on (the
        reference fixes them through eps=1e-5 in every driver call,
        pics/operators_cuda_class.py:107-127).  ``None`` keeps the current value.  The
        plan is rebuilt, and the samples re-binned, at the next operator call; the
        same limits as the library apply (rat >= 2; nspread 2-16 in 2-D, 2-8 in 3-D)."""
        rat = self.rat if rat is None else int(rat)
        nspread = self.nspread if nspread is None else int(nspread)
        if rat < 2:
            raise ValueError("rat must be >= 2, got %r" % (rat,))
        if not 2 <= nspread <= (8 if self.dim == 3 else 16):
            raise ValueError("nspread %d outside the supported range for %d-D" % (nspread, self.dim))
        if (rat, nspread) != (self.rat, self.nspread):
            self.rat, self.nspread = rat, nspread
            self._plan = None
            self._plan_key = None
            self._points_key = None
        return self

    def density_weighting(self, kdata, dcf):
        """kdata * dcf broadcast over trailing dims, squeezed (pics/operators_cuda_class.py:102-105)."""
        ks, ds = dim_match(tuple(kdata.shape), tuple(dcf.shape))
        out = kdata.reshape(ks) * dcf.reshape(ds)
        return out.squeeze()

    # ---- plan management -------------------------------------------------------
    def _coords(self):
        c = [self.ktrajx, self.ktrajy] + ([self.ktrajz] if self.dim == 3 else [])
        if any(v is None for v in c):
            raise ValueError("trajectory not set")
        return c

    def _ensure(self, ncoils, device):
        modes = tuple(int(m) for m in self.im_shape[:self.dim])
        key = (modes, int(ncoils), str(device))
        if self._plan is None or self._plan_key != key:
            self._plan = Plan(modes, ncoils=ncoils, rat=self.rat, nspread=self.nspread, iflag=self.iflag,
                              device=device, tiles=self.tiles, gather_tiles=self.gather_tiles)
            self._plan_key = key
            self._points_key = None
        pk = tuple(_fingerprint(v) for v in self._coords()) + (_fingerprint(self.dcf), self.df)
        if self._points_key != pk:
            self._plan.set_points(self._coords(), df=self.df, dcf=self.dcf)
            self._points_key = pk
        return self._plan

    @staticmethod
    def _device_of(*xs):
        for x in xs:
            if torch.is_tensor(x) and x.is_cuda:
                return x.device
        return torch.device("cuda", torch.cuda.current_device())

    def _split_image(self, im):
        shape = tuple(im.shape)
        trail = shape[self.dim:]
        return trail, int(np.prod(trail)) if trail else 1

    # ---- operators ----------------------------------------------------------------
    def forward(self, im):
        """Type-2 NUFFT: image (ms,mt[,mu][,nc]) -> k-space (M[,nc]) (operators_cuda_class.py:107-111)."""
        trail, nc = self._split_image(im)
        P = self._ensure(nc, self._device_of(im))
        out = P.type2(torch.as_tensor(im).reshape(P.image_shape()))
        return self._out(out, im, (P.M,) + trail)

    def backward(self, kdata):
        """Density weighting then type-1 NUFFT, (1/M) A^H D y (operators_cuda_class.py:113-120)."""
        trail = tuple(kdata.shape)[1:]
        nc = int(np.prod(trail)) if trail else 1
        P = self._ensure(nc, self._device_of(kdata))
        out = P.type1(torch.as_tensor(kdata).reshape(P.sample_shape()))
        return self._out(out, kdata, tuple(self.im_shape[:self.dim]) + trail)

    def forward_backward(self, im):
        """Type 21: (1/M) A^H D A (operators_cuda_class.py:122-127)."""
        trail, nc = self._split_image(im)
        P = self._ensure(nc, self._device_of(im))
        out = P.type21(torch.as_tensor(im).reshape(P.image_shape()))
        return self._out(out, im, tuple(self.im_shape[:self.dim]) + trail)

    def adjoint(self, kdata):
        """Exact adjoint of forward: A^H y (no 1/M, no density weighting)."""
        trail = tuple(kdata.shape)[1:]
        nc = int(np.prod(trail)) if trail else 1
        P = self._ensure(nc, self._device_of(kdata))
        out = P.adjoint(torch.as_tensor(kdata).reshape(P.sample_shape()))
        return self._out(out, kdata, tuple(self.im_shape[:self.dim]) + trail)

    def _sense_apply(self, kind, maps, x):
        """Fused SENSE transforms for joint2operators(espirit(maps), self).

        type2:  x (ms,mt[,mu][,1]) -> k-space (M, nc) = forward(x * maps)
        type1:  k-space (M, nc) -> image (ms,mt[,mu],1) = sum_c conj(maps_c) backward(y)_c
        type21: image -> image = the composition with forward_backward."""
        nc = int(maps.shape[-1])
        dev = self._device_of(x, maps)
        P = self._ensure(nc, dev)
        fp = _fingerprint(maps)
        if getattr(P, "_sense_key", None) != fp:
            P.set_sense(maps)
            P._sense_key = fp
        img_shape = tuple(self.im_shape[:self.dim]) + (1,)
        t = torch.as_tensor(x)
        if kind == "type2":
            out = P.sense_type2(t.reshape(P.modes))
            return self._out(out, x, (P.M, nc))
        if kind == "type1":
            out = P.sense_type1(t.reshape(P.sample_shape()))
            return self._out(out, x, img_shape)
        out = P.sense_type21(t.reshape(P.modes))
        return self._out(out, x, img_shape)

    @staticmethod
    def _out(out, like, shape):
        out = out.reshape(shape)
        if torch.is_tensor(like):
            return out
        o = out.cpu().numpy()
        if np.asarray(like).dtype == np.complex128:
            o = o.astype(np.complex128)
        return o


class NUFFT3d_hip(_NUFFTnd):
    """3-D Gaussian-gridding NUFFT operator on multi-coil data (pics/operators_cuda_class.py:64-127)."""

    dim = 3

    def __init__(self, im_shape=None, dcf=None, ktrajx=None, ktrajy=None, ktrajz=None, axes=(0, 1, 2), **kw):
        super().__init__(im_shape, dcf, ktrajx, ktrajy, ktrajz, axes, **kw)


class NUFFT2d_hip(_NUFFTnd):
    """2-D counterpart of NUFFT3d_hip (same methods; trajectory (x, y))."""

    dim = 2

    def __init__(self, im_shape=None, dcf=None, ktrajx=None, ktrajy=None, axes=(0, 1), **kw):
        super().__init__(im_shape, dcf, ktrajx, ktrajy, None, axes, **kw)


# names of the reference operators
NUFFT3d_cuda = NUFFT3d_hip
NUFFT3d = NUFFT3d_hip
NUFFT2d = NUFFT2d_hip


class espirit:
    """Coil-sensitivity operator (pics/operators_class.py:437-472).

    forward(im) = im * S (broadcast over the coil axis); backward(x) =
    sum_c conj(S) x, keepdims."""

    def __init__(self, sensitivity=None, coil_axis=None):
        self.sens = sensitivity
        if coil_axis is None and self.sens is not None:
            self.coil_axis = len(sensitivity.shape) - 1
        else:
            self.coil_axis = coil_axis

    def backward(self, im_coils):
        ss, ims = dim_match(tuple(self.sens.shape), tuple(im_coils.shape))
        xp = _xp(im_coils)
        conj = torch.conj(self.sens) if torch.is_tensor(self.sens) else np.conj(self.sens)
        prod = im_coils.reshape(ims) * conj.reshape(ss)
        if xp is torch:
            return prod.sum(dim=self.coil_axis, keepdim=True)
        return np.sum(prod, axis=self.coil_axis, keepdims=True)

    def forward(self, im_sos):
        ss, ims = dim_match(tuple(self.sens.shape), tuple(im_sos.shape))
        return im_sos.reshape(ims) * self.sens.reshape(ss)


class None_opt:
    """Identity operator (pics/operators_class.py:477-485)."""

    def forward(self, xin):
        return xin

    def backward(self, xin):
        return xin


class joint2operators:
    """B(A(x)) composition (pics/operators_class.py:496-514).

    When A is ``espirit`` with coil-last maps of the operator's image shape and B a
    NUFFT operator of this module, the three methods run the fused SENSE transforms
    of the plan (the expansion x * S_c inside the first FFT pass, the combine
    sum_c conj(S_c) inside the last one) instead of separate broadcasts; the results
    equal the composition up to FP32 rounding order."""

    def __init__(self, Aopt, Bopt, fused=True):
        self.Aopt = Aopt
        self.Bopt = Bopt
        self.fused = fused

    def _sense(self, xin):
        """The NUFFT operator and the maps when the fused path applies, else None."""
        A, B = self.Aopt, self.Bopt
        if not (self.fused and isinstance(A, espirit) and isinstance(B, _NUFFTnd) and A.sens is not None):
            return None
        S = A.sens
        d = B.dim
        if B.im_shape is None or A.coil_axis != len(S.shape) - 1 or tuple(S.shape[:d]) != tuple(B.im_shape[:d]) \
                or len(S.shape) != d + 1:
            return None
        return B

    def forward(self, xin):
        B = self._sense(xin)
        if B is not None:
            return B._sense_apply("type2", self.Aopt.sens, xin)
        return self.Bopt.forward(self.Aopt.forward(xin))

    def backward(self, xin):
        B = self._sense(xin)
        if B is not None:
            return B._sense_apply("type1", self.Aopt.sens, xin)
        return self.Aopt.backward(self.Bopt.backward(xin))

    def forward_backward(self, xin):
        B = self._sense(xin)
        if B is not None:
            return B._sense_apply("type21", self.Aopt.sens, xin)
        if "forward_backward" in dir(self.Bopt):
            return self.Aopt.backward(self.Bopt.forward_backward(self.Aopt.forward(xin)))
        return self.Aopt.backward(self.Bopt.backward(self.Bopt.forward(self.Aopt.forward(xin))))


class joint3operators:
    """C(B(A(x))) composition (pics/operators_class.py:520-539)."""

    def __init__(self, Aopt, Bopt, Copt):
        self.Aopt = Aopt
        self.Bopt = Bopt
        self.Copt = Copt

    def forward(self, xin):
        return self.Copt.forward(self.Bopt.forward(self.Aopt.forward(xin)))

    def backward(self, xin):
        return self.Aopt.backward(self.Bopt.backward(self.Copt.backward(xin)))

    def forward_backward(self, xin):
        if "forward_backward" in dir(self.Copt):
            return self.Aopt.backward(self.Bopt.backward(self.Copt.forward_backward(self.Bopt.forward(
                self.Aopt.forward(xin)))))
        return self.backward(self.forward(xin))
