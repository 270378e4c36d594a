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
    "aoa_eltwise_scalar": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int, c_vp, c_vp, ctypes.c_uint64, ctypes.c_int,
                                          ctypes.c_int]),
    "aoa_conjugate": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int]),
    "aoa_special_fft": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_int, ctypes.c_int]),
    "aoa_encode": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_int]),
    "aoa_decode": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int, ctypes.c_uint64]),
    "aoa_sample": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                  ctypes.c_uint64]),
    "aoa_rescale": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int]),
    "aoa_encrypt": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_uint64, ctypes.c_uint64,
                                   c_vp, ctypes.c_size_t]),
    "aoa_decrypt": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_int]),
    "aoa_modup": (ctypes.c_int, [c_vp, c_vp, c_vp, ctypes.c_int]),
    "aoa_moddown": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int]),
    "aoa_automorphism": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int, ctypes.c_int]),
    "aoa_workspace_bytes": (ctypes.c_size_t, [c_vp, ctypes.c_int]),
    "aoa_keyswitch": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_int, c_vp,
                                      ctypes.c_size_t]),
    "aoa_hommul_relin_rescale": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                                 ctypes.c_int, c_vp, ctypes.c_size_t]),
    "aoa_hommul_relin_rescale_noalloc": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                                         c_vp, ctypes.c_int]),
    "aoa_noalloc_scratch_bytes": (ctypes.c_size_t, []),
    "aoa_rotate_left": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_int,
                                        ctypes.c_int, c_vp, ctypes.c_size_t]),
    "aoa_rotate_right": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_int,
                                         ctypes.c_int, c_vp, ctypes.c_size_t]),
}



class MulItem(ctypes.Structure):
    _fields_ = [("c1", c_vp * 2), ("c2", c_vp * 2), ("out", c_vp * 2), ("level", ctypes.c_int)]


class RotItem(ctypes.Structure):
    _fields_ = [("inp", c_vp * 2), ("out", c_vp * 2), ("rtk_b", c_vp), ("rtk_a", c_vp),
                ("rot_slots", ctypes.c_int), ("level", ctypes.c_int), ("accumulate", ctypes.c_int)]


class NttItem(ctypes.Structure):
    _fields_ = [("poly", c_vp * 2), ("level", ctypes.c_int), ("dir", ctypes.c_int)]


class Op(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("rot_slots", ctypes.c_int)]


SIGNATURES.update({
    "aoa_batch_add_ntt": (ctypes.c_int, [c_vp, c_vp, ctypes.POINTER(NttItem), ctypes.c_int]),
    "aoa_server_create": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(c_vp)]),
    "aoa_server_register": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(c_vp), ctypes.POINTER(c_vp)]),
    "aoa_server_submit": (ctypes.c_int, [c_vp, ctypes.c_int, c_vp, c_vp, ctypes.c_int, ctypes.c_int, c_vp, c_vp,
                                          ctypes.POINTER(Op), ctypes.c_int, c_vp, c_vp,
                                          ctypes.POINTER(ctypes.c_int64)]),
    "aoa_server_step": (ctypes.c_int, [c_vp, c_vp]),
    "aoa_server_start": (ctypes.c_int, [c_vp]),
    "aoa_server_stop": (ctypes.c_int, [c_vp]),
    "aoa_server_result": (ctypes.c_int, [c_vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]),
    "aoa_server_stats": (ctypes.c_int, [c_vp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                         ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "aoa_server_release": (ctypes.c_int, [c_vp, ctypes.c_int64]),
    "aoa_server_destroy": (None, [c_vp]),
})

SIGNATURES.update({
    "aoa_params_set_ntt_split": (ctypes.c_int, [c_vp, ctypes.c_int]),
    "aoa_params_get_ntt_split": (ctypes.c_int, [c_vp]),
    "aoa_params_set_moddown_exact": (ctypes.c_int, [c_vp, ctypes.c_int]),
    "aoa_params_get_moddown_exact": (ctypes.c_int, [c_vp]),
    "aoa_hommul_relin": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                         ctypes.c_int, c_vp, ctypes.c_size_t]),
    "aoa_rho": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "aoa_modup_from": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.c_int]),
    "aoa_dot": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_int]),
    "aoa_dr2": (ctypes.c_int, [c_vp] * 9),
    "aoa_rns_primes": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, c_u64p, ctypes.c_int]),
    "aoa_autotune_ntt": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_double), ctypes.c_int]),
    "aoa_batch_create": (ctypes.c_int, [ctypes.POINTER(c_vp)]),
    "aoa_batch_add_mul": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(MulItem), ctypes.c_int]),
    "aoa_batch_add_rotate": (ctypes.c_int, [c_vp, c_vp, ctypes.POINTER(RotItem), ctypes.c_int]),
    "aoa_batch_barrier": (ctypes.c_int, [c_vp]),
    "aoa_batch_build": (ctypes.c_int, [c_vp]),
    "aoa_batch_workspace_bytes": (ctypes.c_size_t, [c_vp]),
    "aoa_batch_launch": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int]),
    "aoa_batch_destroy": (None, [c_vp]),
    "aoa_batch_stage_times": (ctypes.c_int, [c_vp, c_vp, ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_int), ctypes.c_int]),
})

class BatchOptions(ctypes.Structure):
    _fields_ = [("ntt_split", ctypes.c_int), ("prime_major", ctypes.c_int), ("xcd_chunks", ctypes.c_int),
                ("strict_split", ctypes.c_int)]


class TuneResult(ctypes.Structure):
    _fields_ = [("opt", BatchOptions), ("ms", ctypes.c_double), ("checksum", ctypes.c_uint64),
                ("exact", ctypes.c_int)]


SIGNATURES.update({
    "aoa_batch_default_options": (None, [ctypes.POINTER(BatchOptions)]),
    "aoa_batch_set_options": (ctypes.c_int, [c_vp, ctypes.POINTER(BatchOptions)]),
    "aoa_batch_get_options": (ctypes.c_int, [c_vp, ctypes.POINTER(BatchOptions)]),
    "aoa_batch_autotune": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int, ctypes.POINTER(TuneResult), ctypes.c_int]),
})

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
