"""ctypes binding of libaoa.so (the C-ABI declared in include/aoa.h).

Loading fails loudly when the library has not been built: there is no
fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaoa.so")

c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_vp = ctypes.c_void_p

# name: (restype, argtypes)
SIGNATURES = {
    "aoa_last_error": (ctypes.c_char_p, []),
    "aoa_params_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(c_vp)]),
    "aoa_params_destroy": (None, [c_vp]),
    "aoa_params_info": (ctypes.c_int, [c_vp] + [ctypes.POINTER(ctypes.c_int)] * 4),
    "aoa_params_primes": (ctypes.c_int, [c_vp, c_u64p, ctypes.c_int]),
    "aoa_params_table": (ctypes.c_int64, [c_vp, ctypes.c_int, c_u64p, ctypes.c_int64]),
    "aoa_ntt": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "aoa_eltwise": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int, c_vp, c_vp, c_vp, c_vp, ctypes.c_int,
                                    ctypes.c_int]),
    "aoa_rescale": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int]),
    "aoa_modup": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_int]),
    "aoa_moddown": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int]),
    "aoa_automorphism": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int, ctypes.c_int]),
    "aoa_workspace_bytes": (ctypes.c_size_t, [c_vp, ctypes.c_int]),
    "aoa_keyswitch": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_int, c_vp,
                                      ctypes.c_size_t]),
    "aoa_hommul_relin_rescale": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                                 ctypes.c_int, c_vp, ctypes.c_size_t]),
    "aoa_rotate_left": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_int,
                                        ctypes.c_int, c_vp, ctypes.c_size_t]),
}

_lib = None


def load():
    """Return the loaded libaoa.so, raising if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libaoa.so is not built (%s); run python -m aoa_amd.build" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class AoaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("aoa error %d: %s" % (code, msg))
        self.code = code


def check(rc):
    if rc != 0:
        raise AoaError(rc, load().aoa_last_error().decode())
    return rc
