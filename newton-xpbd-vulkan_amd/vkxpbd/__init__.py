"""vkxpbd -- XPBD particle solver for AMD Instinct MI355X (gfx950).

Same Python surface as the reference's `vkxpbd` pybind11 module
(src/bindings.cpp:8-301): `SolverConfig`, `XpbdSolver`, `enumerate_devices`.
The native module wraps the C-ABI of libvkxpbd.so (include/xpbd_capi.h),
whose kernels are hand-written HIP for gfx950.  Importing fails loudly when
the native library has not been built; there is no CPU fallback.

PyTorch, when installed, is imported first: it ships its own HIP runtime, and
a process can only use one.  Loading torch's first makes libvkxpbd.so bind to
it, so solver kernels and torch tensors (vkxpbd.autograd) share one runtime;
without torch the library uses the system ROCm runtime.
"""
try:  # noqa: SIM105
    import torch  # noqa: F401
except ImportError:
    pass

from ._vkxpbd import SolverConfig, XpbdSolver, enumerate_devices  # noqa: E402,F401

__all__ = ["SolverConfig", "XpbdSolver", "enumerate_devices"]
