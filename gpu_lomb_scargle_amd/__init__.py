"""MI355X-native Lomb-Scargle period search.

* ``gpu_lomb_scargle_amd.lombscarglegpu`` -- the reference-compatible Python
  interface (``lombscargle(...)``) over the drop-in C libraries.
* ``gpu_lomb_scargle_amd._native`` -- ctypes bindings of the C ABI
  (``include/lombscargle_amd.h``).
* ``gpu_lomb_scargle_amd.device`` -- the sweep on device-resident torch
  tensors (PyTorch-ROCm provides only memory and streams), and device
  segmentation of object runs.
* ``gpu_lomb_scargle_amd.grad`` -- the differentiable power at arbitrary
  frequencies (``torch.autograd.Function`` over the HIP backward kernel) and
  Newton refinement of periodogram peaks.
"""
from . import _native  # noqa: F401

__all__ = ["lombscarglegpu", "_native", "device", "grad", "synth"]
