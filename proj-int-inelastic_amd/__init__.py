"""MI355X-native fast spectral inelastic Boltzmann collision operator and step loop.

Drop-in for the reference's collision class (src/collision) and solver step
API (src/solver.py, src/boltzmann.py).  Compute runs in hand-written gfx950
HIP kernels behind the C ABI of include/inelastic.h, called through ctypes;
torch-ROCm tensors serve as device buffers.
"""
__version__ = "0.1.0"

from . import _lib  # noqa: F401
