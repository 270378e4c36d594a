"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline use it, always as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p
_lib = None


def _p(a):
    return a.ctypes.data_as(_u64p) if a is not None else None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)
    lib = ctypes.CDLL(LIB)
    sig = {
        "orc_create": (_vp, [ctypes.c_int] * 4),
        "orc_destroy": (None, [_vp]),
        "orc_set_threads": (ctypes.c_int, [ctypes.c_int]),
        "orc_primes": (ctypes.c_int, [_vp, _u64p]),
        "orc_roots": (ctypes.c_int, [_vp, _u64p]),
        "orc_set_moddown_exact": (None, [_vp, ctypes.c_int]),
        "orc_fft_special": (None, [ctypes.POINTER(ctypes.c_double), ctypes.c_long, ctypes.c_long, ctypes.c_int]),
        "orc_factor": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
        "orc_table": (ctypes.c_int64, [_vp, ctypes.c_int, _u64p]),
        "orc_mod128": (ctypes.c_uint64, [_vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]),
        "orc_mulmod": (ctypes.c_uint64, [_vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]),
        "orc_ntt": (None, [_vp, _u64p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "orc_eltwise": (None, [_vp, ctypes.c_int, _u64p, _u64p, _u64p, ctypes.c_int]),
        "orc_eltwise_scalar": (None, [_vp, ctypes.c_int, _u64p, _u64p, ctypes.c_uint64, ctypes.c_int]),
        "orc_conjugate": (None, [_vp, _u64p, _u64p, ctypes.c_int]),
        "orc_rescale": (None, [_vp, _u64p, _u64p, ctypes.c_int]),
        "orc_modup": (None, [_vp, _u64p, ctypes.c_int]),
        "orc_moddown": (None, [_vp, _u64p, _u64p, ctypes.c_int]),
        "orc_automorphism": (None, [_vp, _u64p, _u64p, ctypes.c_int, ctypes.c_int]),
        "orc_hommul": (None, [_vp] + [_u64p] * 6 + [ctypes.c_int, _u64p, _u64p, ctypes.c_int, ctypes.c_int]),
        "orc_rotate_left": (None, [_vp] + [_u64p] * 6 + [ctypes.c_int, ctypes.c_int]),
        "orc_rotate_right": (None, [_vp] + [_u64p] * 6 + [ctypes.c_int, ctypes.c_int]),
        "orc_dot": (None, [_vp, _u64p, ctypes.POINTER(_u64p), ctypes.POINTER(_u64p), ctypes.c_int]),
        "orc_dr2": (None, [_vp] + [_u64p] * 7),
        "orc_encode": (None, [_vp, _u64p, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_uint64,
                              ctypes.c_int]),
        "orc_decode": (None, [_vp, ctypes.POINTER(ctypes.c_double), _u64p, ctypes.c_int, ctypes.c_uint64]),
        "orc_philox4x32_10": (None, [ctypes.POINTER(ctypes.c_uint32)] * 3),
        "orc_gauss_cdt": (None, [_u64p]),
        "orc_c1_keygen": (_vp, [_vp, ctypes.c_uint64]),
        "orc_c1_free": (None, [_vp]),
        "orc_c1_key": (ctypes.c_int64, [_vp, ctypes.c_int, _u64p]),
        "orc_c1_run": (None, [_vp, _vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                              ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                              ctypes.POINTER(ctypes.c_double), _u64p]),
        "orc_sample": (None, [_vp, _u64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                              ctypes.c_uint64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def philox4x32_10(ctr, key):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    P = ctypes.POINTER(ctypes.c_uint32)
    load().orc_philox4x32_10(c.ctypes.data_as(P), k.ctypes.data_as(P), out.ctypes.data_as(P))
    return [int(x) for x in out]


def gauss_cdt():
    out = np.zeros(64, dtype=np.uint64)
    load().orc_gauss_cdt(_p(out))
    return out


class Oracle:
    def __init__(self, N, k, kl=0, prec=48):
        self.lib = load()
        self.h = self.lib.orc_create(N, k, kl, prec)
        if not self.h:
            raise ValueError("oracle does not support N=%d k=%d prec=%d" % (N, k, prec))
        self.N, self.k = N, k
        self.kl = kl if kl > 0 else k + 1
        buf = np.zeros(self.k + self.kl, dtype=np.uint64)
        self.lib.orc_primes(self.h, _p(buf))
        self.primes = [int(x) for x in buf]
        self.lib.orc_roots(self.h, _p(buf))
        self.roots = [int(x) for x in buf]
        na, nb = ctypes.c_int(), ctypes.c_int()
        self.lib.orc_factor(self.h, ctypes.byref(na), ctypes.byref(nb))
        self.Na, self.Nb = na.value, nb.value

    def __del__(self):
        try:
            self.lib.orc_destroy(self.h)
        except Exception:
            pass

    @staticmethod
    def threads(n=0):
        return load().orc_set_threads(n)

    def set_moddown_exact(self, exact):
        self.lib.orc_set_moddown_exact(self.h, int(bool(exact)))

    def table(self, which):
        n = self.lib.orc_table(self.h, which, None)
        out = np.zeros(n, dtype=np.uint64)
        self.lib.orc_table(self.h, which, _p(out))
        return out

    def mod128(self, rid, v):
        return int(self.lib.orc_mod128(self.h, rid, (v >> 64) & (2**64 - 1), v & (2**64 - 1)))

    def mulmod(self, rid, a, b):
        return int(self.lib.orc_mulmod(self.h, rid, a, b))

    def ntt(self, d, res0, nres, direction):
        assert d.dtype == np.uint64 and d.flags.c_contiguous
        self.lib.orc_ntt(self.h, _p(d), res0, nres, direction)

    def eltwise(self, op, c, a, b, nres):
        self.lib.orc_eltwise(self.h, op, _p(c), _p(a), _p(b), nres)

    def eltwise_scalar(self, op, c, a, b, nres):
        self.lib.orc_eltwise_scalar(self.h, op, _p(c), _p(a), b, nres)

    def conjugate(self, out, inp, nres):
        self.lib.orc_conjugate(self.h, _p(out), _p(inp), nres)

    def encode(self, vals, scale, nres=None):
        """Residues (nres x N) of the slot values `vals` (complex, length a
        power of two); the reference's ckks_encode with no empty slots."""
        nres = self.k if nres is None else nres
        v = np.ascontiguousarray(np.asarray(vals, dtype=np.complex128)).view(np.float64).copy()
        out = np.zeros(nres * self.N, dtype=np.uint64)
        self.lib.orc_encode(self.h, _p(out), v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(vals), scale,
                            nres)
        return out

    def decode(self, limb0, slots, scale):
        """Slot values from limb 0 (coefficient domain), the reference's ckks_decode."""
        a = np.ascontiguousarray(limb0[:self.N], dtype=np.uint64)
        v = np.zeros(2 * slots, dtype=np.float64)
        self.lib.orc_decode(self.h, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _p(a), slots, scale)
        return v.view(np.complex128)

    def sample(self, kind, seed, nonce, res0=0, nres=None):
        """The product sampler's polynomial (limbs res0..res0+nres-1 of a QB-sized buffer)."""
        nres = self.k + self.kl - res0 if nres is None else nres
        out = np.zeros((self.k + self.kl) * self.N, dtype=np.uint64)
        self.lib.orc_sample(self.h, _p(out), res0, nres, kind, seed, nonce)
        return out

    def c1_keys(self, seed):
        """Keys of the end-to-end pipeline (s on QB limbs, pk on Q, evk on QB; NTT domain)."""
        kh = self.lib.orc_c1_keygen(self.h, seed)
        keys = {}
        for i, name in enumerate(("s", "pkb", "pka", "evkb", "evka")):
            n = self.lib.orc_c1_key(kh, i, None)
            a = np.zeros(n, dtype=np.uint64)
            self.lib.orc_c1_key(kh, i, _p(a))
            keys[name] = a
        keys["_h"] = kh
        return keys

    def c1_run(self, keys, z1, z2, scale, seed, nonce):
        """encode -> encrypt -> HomMul+relin+rescale -> decrypt -> decode; returns (slots, c3)."""
        slots = len(z1)
        P = ctypes.POINTER(ctypes.c_double)
        a = np.ascontiguousarray(z1, dtype=np.complex128).view(np.float64).copy()
        b = np.ascontiguousarray(z2, dtype=np.complex128).view(np.float64).copy()
        out = np.zeros(2 * slots, dtype=np.float64)
        c3 = np.zeros(2 * self.k * self.N, dtype=np.uint64)
        self.lib.orc_c1_run(self.h, keys["_h"], a.ctypes.data_as(P), b.ctypes.data_as(P), slots, scale, seed, nonce,
                            out.ctypes.data_as(P), _p(c3))
        return out.view(np.complex128), c3

    def rescale(self, a, b, level):
        self.lib.orc_rescale(self.h, _p(a), _p(b), level)

    def modup(self, qb, level):
        self.lib.orc_modup(self.h, _p(qb), level)

    def moddown(self, out, qb, level):
        self.lib.orc_moddown(self.h, _p(out), _p(qb), level)

    def automorphism(self, out, inp, nres, rot):
        self.lib.orc_automorphism(self.h, _p(out), _p(inp), nres, rot)

    def hommul(self, c1, c2, evk, level, in_state=2, rescale=1):
        qn = self.k * self.N
        o0 = np.zeros(qn, dtype=np.uint64)
        o1 = np.zeros(qn, dtype=np.uint64)
        self.lib.orc_hommul(self.h, _p(o0), _p(o1), _p(c1[0]), _p(c1[1]), _p(c2[0]), _p(c2[1]), in_state,
                            _p(evk[0]), _p(evk[1]), level, rescale)
        return o0, o1

    def rotate_left(self, c, rtk, rot, level):
        qn = self.k * self.N
        o0 = np.zeros(qn, dtype=np.uint64)
        o1 = np.zeros(qn, dtype=np.uint64)
        self.lib.orc_rotate_left(self.h, _p(o0), _p(o1), _p(c[0]), _p(c[1]), _p(rtk[0]), _p(rtk[1]), rot,
                                 level)
        return o0, o1

    def dot(self, a, b):
        """sum_i a[i] b[i] over every QB residue, the reference's dot_prod."""
        k = len(a)
        pa = (_u64p * max(1, k))(*[_p(x) for x in a])
        pb = (_u64p * max(1, k))(*[_p(x) for x in b])
        c = np.zeros((self.k + self.kl) * self.N, dtype=np.uint64)
        self.lib.orc_dot(self.h, _p(c), pa, pb, k)
        return c

    def dr2(self, a01, a02, a11, a12):
        n = (self.k + self.kl) * self.N
        c = [np.zeros(n, dtype=np.uint64) for _ in range(3)]
        self.lib.orc_dr2(self.h, _p(c[0]), _p(c[1]), _p(c[2]), _p(a01), _p(a02), _p(a11), _p(a12))
        return c

    def rotate_right(self, c, rtk, rot, level):
        qn = self.k * self.N
        o0 = np.zeros(qn, dtype=np.uint64)
        o1 = np.zeros(qn, dtype=np.uint64)
        self.lib.orc_rotate_right(self.h, _p(o0), _p(o1), _p(c[0]), _p(c[1]), _p(rtk[0]), _p(rtk[1]), rot,
                                  level)
        return o0, o1

    # ---- input helpers ------------------------------------------------------
    def random_poly(self, rng, nlimbs, first=0):
        """Uniform residues in [0, q_r) for limbs first..first+nlimbs-1."""
        out = np.empty(nlimbs * self.N, dtype=np.uint64)
        for r in range(nlimbs):
            p = self.primes[first + r]
            out[r * self.N:(r + 1) * self.N] = rng.integers(0, p, size=self.N, dtype=np.uint64)
        return out
