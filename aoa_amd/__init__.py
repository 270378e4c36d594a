"""MI355X-native CKKS evaluation core (HIP kernels for gfx950 behind a C-ABI).

The Python layer is plumbing for tests and benchmarks: device buffers come
from torch, every computation runs in libaoa.so.
"""
from ._lib import AoaError, check, load  # noqa: F401
from .core import ADD, FORWARD, INVERSE, MUL, MULADD, NEGATE, SUB, Batch, Params, Server  # noqa: F401
