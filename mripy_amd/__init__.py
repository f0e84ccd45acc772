"""mripy_amd: MI355X-native Gaussian-gridding NUFFT behind mripy's operator API.

The compute path is the C-ABI library ``libgaussnufft.so`` (hand-written HIP
kernels for gfx950, including the pruned oversampled FFTs; rocFFT only for grid
lengths those do not cover), driven from Python through ctypes on torch
device tensors.  Submodules load the library lazily; nothing here falls back to
a CPU implementation.
"""
__version__ = "0.1.0"


def __getattr__(name):
    if name in ("Plan", "grid_params"):
        from . import plan
        return getattr(plan, name)
    raise AttributeError(name)
