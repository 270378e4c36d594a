"""Builds and loads tests/native/modarith_check.cpp (host checks of the
product's modular arithmetic and parameter tables) with g++."""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "native", "modarith_check.cpp")
OUT = os.path.join(HERE, "native", "_build", "libchk.so")
DEPS = [SRC, os.path.join(ROOT, "aoa_amd", "csrc", "params.cpp"), os.path.join(ROOT, "aoa_amd", "csrc", "modarith.hpp"),
        os.path.join(ROOT, "aoa_amd", "csrc", "params.hpp")]
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    stale = not os.path.exists(OUT) or any(os.path.getmtime(d) > os.path.getmtime(OUT) for d in DEPS)
    if stale:
        os.makedirs(os.path.dirname(OUT), exist_ok=True)
        subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-I" + os.path.join(ROOT, "aoa_amd", "csrc"),
                        SRC, os.path.join(ROOT, "aoa_amd", "csrc", "params.cpp"), "-o", OUT + ".tmp"], check=True)
        os.replace(OUT + ".tmp", OUT)
    lib = ctypes.CDLL(OUT)
    lib.chk_ref_barrett.restype = ctypes.c_uint64
    lib.chk_ref_barrett.argtypes = [ctypes.c_uint64] * 4 + [ctypes.c_int]
    lib.chk_mul_barrett.restype = ctypes.c_uint64
    lib.chk_mul_barrett.argtypes = [ctypes.c_uint64] * 4 + [ctypes.c_int]
    lib.chk_exact_mults.restype = ctypes.c_int64
    lib.chk_exact_mults.argtypes = [ctypes.c_int] * 4 + [ctypes.c_int64, ctypes.c_uint64]
    lib.chk_moddown_terms.restype = ctypes.c_int
    lib.chk_moddown_terms.argtypes = [ctypes.c_int] * 4 + [ctypes.c_int64, ctypes.c_uint64,
                                                           ctypes.POINTER(ctypes.c_int64)]
    lib.chk_barrett_delta.restype = ctypes.c_int64
    lib.chk_barrett_delta.argtypes = [ctypes.c_int] * 4 + [ctypes.c_int64, ctypes.c_uint64]
    lib.chk_moddown_screen_words.restype = ctypes.c_int64
    lib.chk_moddown_screen_words.argtypes = [ctypes.c_int64, ctypes.c_uint64]
    lib.chk_moddown_accumulate.restype = ctypes.c_int64
    lib.chk_moddown_accumulate.argtypes = [ctypes.c_int] * 4 + [ctypes.c_int64, ctypes.c_uint64]
    lib.chk_moddown_reduce.restype = ctypes.c_int64
    lib.chk_moddown_reduce.argtypes = [ctypes.c_int] * 4 + [ctypes.c_int64, ctypes.c_uint64]
    lib.chk_moddown_finish.restype = ctypes.c_int64
    lib.chk_moddown_finish.argtypes = [ctypes.c_int] * 4 + [ctypes.c_int64, ctypes.c_uint64]
    lib.chk_modup_sums.restype = ctypes.c_int64
    lib.chk_modup_sums.argtypes = [ctypes.c_int] * 4 + [ctypes.c_int64, ctypes.c_uint64]
    lib.chk_moddown_wide.restype = ctypes.c_int64
    lib.chk_moddown_wide.argtypes = [ctypes.c_int] * 4 + [ctypes.c_int64, ctypes.c_uint64]
    lib.chk_row_twiddles.restype = ctypes.c_int64
    lib.chk_row_twiddles.argtypes = [ctypes.c_int] * 4
    lib.chk_table.restype = ctypes.c_int64
    lib.chk_table.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_uint64)]
    _lib = lib
    return lib


def table(N, k, kl, prec, which):
    import numpy as np
    lib = load()
    n = lib.chk_table(N, k, kl, prec, which, None)
    if n < 0:
        raise ValueError("bad table request")
    out = np.zeros(n, dtype=np.uint64)
    lib.chk_table(N, k, kl, prec, which, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return out
