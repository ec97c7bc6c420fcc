"""MI355X-native TTI finite-difference hot path (forward, adjoint and gradient passes) for gfx950.

The compute path is the HIP library libtti_hip.so (hand-written gfx950 kernels behind the C ABI in include/).
PyTorch is used only for device memory and streams in the Python-side wrappers and tests.
"""
from .grid import Grid

__all__ = ["Grid"]
