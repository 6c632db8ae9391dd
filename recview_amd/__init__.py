"""recview_amd -- MI355X filtered back-projection engine with RecView's C ABI.

Modules:
  _lib     ctypes binding of librecview_hip.so (include/recview_hip.h)
  reconst  RecView caller mirror over the drop-in entry points (CudaReconstHost)
  fbp      native batched plans (rvPlan*): multi-slice, LINEAR / LUT16, FP32/FP16/BF16
  phantom  seeded synthetic phantoms and sinograms
  build    in-tree hipcc build for gfx950
"""
from ._lib import LIB_PATH, lib  # noqa: F401

__all__ = ["lib", "LIB_PATH"]
