"""MI355X-native CKKS evaluation core (HIP kernels for gfx950 behind a C-ABI).

The Python layer is plumbing for tests and benchmarks: device buffers come
from torch, every computation runs in libaoa.so.
"""
from ._lib import AoaError, check, load  # noqa: F401
from .core import (ADD, FORWARD, INVERSE, MUL, MULADD, NEGATE, SAMPLE_GAUSSIAN, SAMPLE_HW,  # noqa: F401
                   SAMPLE_NARROW, SAMPLE_UNIFORM, SAMPLE_ZO, SUB, Batch, Params, Server)
