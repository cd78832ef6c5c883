"""MI355X batched FDTD engine: drop-in ``BatchedFDTD2D`` / ``BatchedFDTD3D``.

The solvers run in hand-written HIP kernels for gfx950 behind the C ABI of
``include/ceep_fdtd.h`` (``libceep_fdtd.so``, loaded with ctypes).
"""
from .batched_2d import BatchedFDTD2D
from .batched_3d import BatchedFDTD3D
from .sparams import MultistaticSParameters
from . import _native, imaging, phantoms

__all__ = ["BatchedFDTD2D", "BatchedFDTD3D", "MultistaticSParameters"]
__version__ = "0.1.0"
