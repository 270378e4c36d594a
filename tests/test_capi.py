"""The C-ABI library loads and exports every entry point include/aoa.h
declares (no GPU calls); host-side argument checks that need no device."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "aoa.h")
LIB = os.path.join(ROOT, "aoa_amd", "libaoa.so")
NEWCKKS = os.path.join(ROOT, "aoa_amd", "libnewckks.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aoa_[a-z0-9_]+)\s*\(", text)))


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.strip()}


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 30
    have = exported(LIB)
    missing = [n for n in names if n not in have]
    assert not missing, missing


def test_library_loads_and_python_binding_matches():
    import aoa_amd
    from aoa_amd import _lib
    lib = aoa_amd.load()
    for name in _lib.SIGNATURES:
        assert hasattr(lib, name)
    assert set(_lib.SIGNATURES) <= set(declared())


def test_host_side_argument_errors():
    lib = ctypes.CDLL(LIB)
    lib.aoa_params_create.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_void_p)]
    lib.aoa_last_error.restype = ctypes.c_char_p
    h = ctypes.c_void_p()
    # unsupported precision / ring dimension fail before any device allocation
    assert lib.aoa_params_create(4096, 4, 5, 50, ctypes.byref(h)) == 4
    assert b"precision" in lib.aoa_last_error()
    assert lib.aoa_params_create(1000, 4, 5, 48, ctypes.byref(h)) == 4
    assert lib.aoa_params_create(4096, 80, 5, 48, ctypes.byref(h)) == 4
    lib.aoa_batch_create.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    b = ctypes.c_void_p()
    assert lib.aoa_batch_create(ctypes.byref(b)) == 0
    lib.aoa_batch_destroy.argtypes = [ctypes.c_void_p]
    lib.aoa_batch_destroy(b)


def test_newckks_api_exported():
    have = exported(NEWCKKS)
    demangled = subprocess.run(["nm", "-DC", "--defined-only", NEWCKKS], capture_output=True, text=True,
                               check=True).stdout
    for fn in ["ckks_encrypt", "ckks_decrypt", "ckks_mul", "ckks_mul_without_rescale", "ckks_rescale",
               "ckks_add", "ckks_rotate_left", "ckks_rotate_right", "ckks_sumslots", "ckks_keygen",
               "ckks_new_sk", "ckks_new_pk", "ckks_new_evk", "ckks_new_mtk", "ckks_new_rtk_left",
               "CUDAManager::init", "CKKSContext::CKKSContext", "cipher_init", "poly_mul", "poly_modup",
               "poly_moddown", "poly_rescale", "fftSpecial", "COMMONEngine::execute"]:
        assert fn in demangled, fn
    assert have
