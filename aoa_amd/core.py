"""Thin host wrappers over the C-ABI for tests and benchmarks.

Device polynomials are torch int64 tensors holding uint64 residues
(limb-major, limb r at offset r*N); the kernels see raw pointers.
"""
import ctypes

import numpy as np

from ._lib import BatchOptions, MulItem, Op, RotItem, TuneResult, check, load

FORWARD, INVERSE = 0, 1
ADD, SUB, MUL, NEGATE, MULADD = 0, 1, 2, 3, 5
SAMPLE_UNIFORM, SAMPLE_GAUSSIAN, SAMPLE_ZO, SAMPLE_HW, SAMPLE_NARROW = 0, 1, 2, 3, 4


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class Params:
    """aoa_params handle: one (N, k, kl, prec) parameter set resident in HBM."""

    def __init__(self, N, k, kl=0, prec=48):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.aoa_params_create(N, k, kl, prec, ctypes.byref(h)))
        self.h = h
        n, kk, kkl, pr = (ctypes.c_int() for _ in range(4))
        check(self.lib.aoa_params_info(h, ctypes.byref(n), ctypes.byref(kk), ctypes.byref(kkl),
                                       ctypes.byref(pr)))
        self.N, self.k, self.kl, self.prec = n.value, kk.value, kkl.value, pr.value
        buf = (ctypes.c_uint64 * (self.k + self.kl))()
        check(self.lib.aoa_params_primes(h, buf, self.k + self.kl))
        self.primes = [int(x) for x in buf]

    def close(self):
        if self.h:
            self.lib.aoa_params_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def table(self, which):
        n = self.lib.aoa_params_table(self.h, which, None, 0)
        if n < 0:
            raise ValueError("unknown table %d" % which)
        out = np.zeros(n, dtype=np.uint64)
        self.lib.aoa_params_table(self.h, which, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), n)
        return out

    # ---- device helpers -------------------------------------------------
    def zeros(self, nlimbs, device="cuda"):
        import torch
        return torch.zeros(nlimbs * self.N, dtype=torch.int64, device=device)

    def to_device(self, arr):
        import torch
        a = np.ascontiguousarray(arr, dtype=np.uint64).view(np.int64)
        return torch.from_numpy(a.copy()).cuda()

    @staticmethod
    def to_host(t):
        return t.detach().cpu().numpy().view(np.uint64).copy()

    def workspace(self, level):
        import torch
        nbytes = self.lib.aoa_workspace_bytes(self.h, level)
        return torch.empty(nbytes // 8, dtype=torch.int64, device="cuda"), nbytes

    def ntt(self, d, res0, nres, direction, stream=None):
        check(self.lib.aoa_ntt(self.h, _stream(stream), _ptr(d), res0, nres, direction))

    def eltwise(self, op, out, a, b=None, c=None, res0=0, nres=None, stream=None):
        nres = nres if nres is not None else self.k
        check(self.lib.aoa_eltwise(self.h, _stream(stream), op, _ptr(out), _ptr(a), _ptr(b), _ptr(c), res0,
                                   nres))

    def eltwise_scalar(self, op, out, a, b, res0=0, nres=None, stream=None):
        nres = nres if nres is not None else self.k
        check(self.lib.aoa_eltwise_scalar(self.h, _stream(stream), op, _ptr(out), _ptr(a), int(b), res0, nres))

    def conjugate(self, out, inp, nres, stream=None):
        check(self.lib.aoa_conjugate(self.h, _stream(stream), _ptr(out), _ptr(inp), nres))

    def special_fft(self, vals, n, direction, stream=None):
        """In-place encoder FFT of n complex values (float64 device buffer of
        2n values, re/im interleaved); INVERSE is the encoding direction."""
        check(self.lib.aoa_special_fft(self.h, _stream(stream), _ptr(vals), n, direction))

    def encode(self, out, vals, slots, scale, nres=None, stream=None):
        """out (nres x N residues) <- encoding of `slots` complex values held
        in vals (float64 device buffer, transformed in place)."""
        nres = self.k if nres is None else nres
        check(self.lib.aoa_encode(self.h, _stream(stream), _ptr(out), _ptr(vals), slots, int(scale), nres))

    def decode(self, vals, inp, slots, scale, stream=None):
        """vals (float64 device buffer, 2*slots) <- slot values of limb 0 of inp (coefficient domain)."""
        check(self.lib.aoa_decode(self.h, _stream(stream), _ptr(vals), _ptr(inp), slots, int(scale)))

    def encrypt(self, c0, c1, m, pk_b, pk_a, seed, nonce, ws, stream=None):
        """Public-key encryption of the plaintext m (coefficient domain, Q
        limbs; left in the NTT domain): u (ZO), e0, e1 (Gaussian) drawn at
        nonces nonce, nonce+1, nonce+2; c0 = pk_b u + e0 + m, c1 = pk_a u + e1
        in the NTT domain.  ws: at least 3 k N words."""
        w = _ptr(ws)
        check(self.lib.aoa_encrypt(self.h, _stream(stream), _ptr(c0), _ptr(c1), _ptr(m), _ptr(pk_b), _ptr(pk_a),
                                   int(seed), int(nonce), w, ws.numel() * ws.element_size()))

    def decrypt(self, m, c0, c1, s, level, stream=None):
        """m = INTT(c1 s + c0) on limbs 0..level, with c0 and c1 transformed
        to the NTT domain in place; s is the NTT-domain secret key."""
        check(self.lib.aoa_decrypt(self.h, _stream(stream), _ptr(m), _ptr(c0), _ptr(c1), _ptr(s), level))

    def sample(self, out, kind, seed, nonce, res0=0, nres=None, stream=None):
        """Limbs res0..res0+nres-1 of out drawn from distribution `kind`
        (SAMPLE_UNIFORM, _GAUSSIAN, _ZO, _HW, _NARROW) keyed by (seed, nonce)."""
        nres = self.k + self.kl - res0 if nres is None else nres
        check(self.lib.aoa_sample(self.h, _stream(stream), _ptr(out), res0, nres, kind, int(seed), int(nonce)))

    def rescale(self, c0, c1, level, stream=None):
        check(self.lib.aoa_rescale(self.h, _stream(stream), _ptr(c0), _ptr(c1), level))

    def modup(self, qb, level, stream=None):
        check(self.lib.aoa_modup(self.h, _stream(stream), _ptr(qb), level))

    def moddown(self, out, qb, level, stream=None):
        check(self.lib.aoa_moddown(self.h, _stream(stream), _ptr(out), _ptr(qb), level))

    def modup_from(self, qb, q, level, stream=None):
        """Special limbs of the QB polynomial qb from Q limbs 0..level of q."""
        check(self.lib.aoa_modup_from(self.h, _stream(stream), _ptr(qb), _ptr(q), level))

    def dot(self, c, a, b, stream=None):
        """c = sum_i a[i] b[i] over every QB residue (a, b: lists of QB tensors)."""
        import torch
        pa = torch.tensor([t.data_ptr() for t in a], dtype=torch.int64, device="cuda")
        pb = torch.tensor([t.data_ptr() for t in b], dtype=torch.int64, device="cuda")
        check(self.lib.aoa_dot(self.h, _stream(stream), _ptr(c), _ptr(pa), _ptr(pb), len(a)))
        return pa, pb  # keep the pointer arrays alive until the stream has used them

    def dr2(self, c21, c22, c23, a01, a02, a11, a12, stream=None):
        check(self.lib.aoa_dr2(self.h, _stream(stream), _ptr(c21), _ptr(c22), _ptr(c23), _ptr(a01), _ptr(a02),
                               _ptr(a11), _ptr(a12)))

    def automorphism(self, out, inp, nres, rot_slots, stream=None):
        check(self.lib.aoa_automorphism(self.h, _stream(stream), _ptr(out), _ptr(inp), nres, rot_slots))

    def keyswitch(self, out0, out1, d, key_b, key_a, level, ws, add0=None, stream=None):
        w, nb = ws
        check(self.lib.aoa_keyswitch(self.h, _stream(stream), _ptr(out0), _ptr(out1), _ptr(add0), _ptr(d),
                                     _ptr(key_b), _ptr(key_a), level, _ptr(w), nb))

    def hommul(self, c3, c1, c2, evk, level, ws, stream=None):
        w, nb = ws
        check(self.lib.aoa_hommul_relin_rescale(
            self.h, _stream(stream), _ptr(c3[0]), _ptr(c3[1]), _ptr(c1[0]), _ptr(c1[1]), _ptr(c2[0]),
            _ptr(c2[1]), _ptr(evk[0]), _ptr(evk[1]), level, _ptr(w), nb))

    def hommul_noalloc(self, c3, c1, c2, evk, level, stream=None):
        """HomMul+relin+rescale with the library's static device workspace
        (no allocation, no workspace argument)."""
        check(self.lib.aoa_hommul_relin_rescale_noalloc(
            self.h, _stream(stream), _ptr(c3[0]), _ptr(c3[1]), _ptr(c1[0]), _ptr(c1[1]), _ptr(c2[0]),
            _ptr(c2[1]), _ptr(evk[0]), _ptr(evk[1]), level))

    def set_moddown_exact(self, exact):
        """False (default): the reference's ModDown reduction, bit-exact residues;
        True: exact reduction."""
        check(self.lib.aoa_params_set_moddown_exact(self.h, int(bool(exact))))

    def hommul_norescale(self, c3, c1, c2, evk, level, ws, stream=None):
        w, nb = ws
        check(self.lib.aoa_hommul_relin(
            self.h, _stream(stream), _ptr(c3[0]), _ptr(c3[1]), _ptr(c1[0]), _ptr(c1[1]), _ptr(c2[0]),
            _ptr(c2[1]), _ptr(evk[0]), _ptr(evk[1]), level, _ptr(w), nb))

    def rho(self, out, inp, stream=None):
        check(self.lib.aoa_rho(self.h, _stream(stream), _ptr(out), _ptr(inp)))

    def set_ntt_split(self, n1):
        check(self.lib.aoa_params_set_ntt_split(self.h, n1))

    def ntt_split(self):
        return self.lib.aoa_params_get_ntt_split(self.h)

    def autotune_ntt(self, nlimbs=0, reps=10, stream=None):
        """Times every NTT split on the device, keeps the fastest; returns
        {n1: microseconds}."""
        n1 = (ctypes.c_int * 8)()
        us = (ctypes.c_double * 8)()
        n = self.lib.aoa_autotune_ntt(self.h, _stream(stream), nlimbs, reps, n1, us, 8)
        if n < 0:
            check(-n)
        return {int(n1[i]): float(us[i]) for i in range(min(n, 8))}

    def rotate_left(self, out, c, rtk, rot_slots, level, ws, stream=None):
        w, nb = ws
        check(self.lib.aoa_rotate_left(self.h, _stream(stream), _ptr(out[0]), _ptr(out[1]), _ptr(c[0]),
                                       _ptr(c[1]), _ptr(rtk[0]), _ptr(rtk[1]), rot_slots, level, _ptr(w),
                                       nb))

    def rotate_right(self, out, c, rtk, rot_slots, level, ws, stream=None):
        w, nb = ws
        check(self.lib.aoa_rotate_right(self.h, _stream(stream), _ptr(out[0]), _ptr(out[1]), _ptr(c[0]),
                                        _ptr(c[1]), _ptr(rtk[0]), _ptr(rtk[1]), rot_slots, level, _ptr(w),
                                        nb))


class Batch:
    """aoa_batch handle: stage-fused evaluation of many ciphertexts, replayed
    as a hipGraph.  Keeps references to every tensor it was given."""

    def __init__(self):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.aoa_batch_create(ctypes.byref(h)))
        self.h = h
        self._keep = []

    def add_mul(self, params, evk, items):
        """items: list of (c1, c2, out, level) with c1/c2/out pairs of tensors."""
        arr = (MulItem * len(items))()
        for i, (c1, c2, out, level) in enumerate(items):
            arr[i].c1[0], arr[i].c1[1] = c1[0].data_ptr(), c1[1].data_ptr()
            arr[i].c2[0], arr[i].c2[1] = c2[0].data_ptr(), c2[1].data_ptr()
            arr[i].out[0], arr[i].out[1] = out[0].data_ptr(), out[1].data_ptr()
            arr[i].level = level
            self._keep += [c1, c2, out]
        self._keep += [evk, params]
        check(self.lib.aoa_batch_add_mul(self.h, params.h, _ptr(evk[0]), _ptr(evk[1]), arr, len(items)))

    def add_rotate(self, params, items):
        """items: list of (inp, out, rtk, rot_slots, level, accumulate)."""
        arr = (RotItem * len(items))()
        for i, (inp, out, rtk, rot, level, acc) in enumerate(items):
            arr[i].inp[0], arr[i].inp[1] = inp[0].data_ptr(), inp[1].data_ptr()
            arr[i].out[0], arr[i].out[1] = out[0].data_ptr(), out[1].data_ptr()
            arr[i].rtk_b, arr[i].rtk_a = rtk[0].data_ptr(), rtk[1].data_ptr()
            arr[i].rot_slots, arr[i].level, arr[i].accumulate = rot, level, int(acc)
            self._keep += [inp, out, rtk]
        self._keep.append(params)
        check(self.lib.aoa_batch_add_rotate(self.h, params.h, arr, len(items)))

    def barrier(self):
        check(self.lib.aoa_batch_barrier(self.h))

    def build(self):
        check(self.lib.aoa_batch_build(self.h))
        return self.lib.aoa_batch_workspace_bytes(self.h)

    def launch(self, stream=None, graph=True):
        check(self.lib.aoa_batch_launch(self.h, _stream(stream), 1 if graph else 0))

    OPTION_FIELDS = ("ntt_split", "prime_major", "xcd_chunks", "strict_split")

    def options(self):
        """The batch's launch options as a dict."""
        o = BatchOptions()
        check(self.lib.aoa_batch_get_options(self.h, ctypes.byref(o)))
        return {f: getattr(o, f) for f in self.OPTION_FIELDS}

    def set_options(self, **kw):
        """Sets launch options (missing ones keep their current values)."""
        cur = self.options()
        cur.update(kw)
        o = BatchOptions(*[int(cur[f]) for f in self.OPTION_FIELDS])
        check(self.lib.aoa_batch_set_options(self.h, ctypes.byref(o)))

    def autotune(self, stream=None, reps=5, max_results=64):
        """Times every launch-option candidate on this batch and keeps the
        fastest exact one; returns the candidates as dicts (options, ms,
        checksum, exact)."""
        res = (TuneResult * max_results)()
        n = self.lib.aoa_batch_autotune(self.h, _stream(stream), reps, res, max_results)
        if n < 0:
            check(-n)
        out = []
        for r in res[:min(n, max_results)]:
            d = {f: getattr(r.opt, f) for f in self.OPTION_FIELDS}
            d.update(ms=r.ms, checksum=int(r.checksum), exact=bool(r.exact))
            out.append(d)
        return out

    STAGES = ("pre_fwd", "pre_inv", "tensor", "automorphism", "intt", "modup", "ntt", "key_intt", "moddown")

    def stage_times(self, stream=None):
        """One eager run with HIP events around every stage launch; returns
        {stage: (milliseconds summed over launches, launches)}."""
        n = len(self.STAGES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int * n)()
        check(self.lib.aoa_batch_stage_times(self.h, _stream(stream), ms, cnt, n))
        return {s: (ms[i], cnt[i]) for i, s in enumerate(self.STAGES) if cnt[i]}

    def close(self):
        if self.h:
            self.lib.aoa_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """aoa_server handle: continuous batching of encrypted-op requests."""

    MUL, ROTATE = 0, 1
    PENDING, ACTIVE, DONE, FAILED = 0, 1, 2, 3

    def __init__(self, pool_bytes, step_ws_bytes, max_active=64):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.aoa_server_create(pool_bytes, step_ws_bytes, max_active, ctypes.byref(h)))
        self.h = h
        self._keep = []

    def register(self, params, evk, rtks=None):
        rtks = rtks or {}
        steps = sorted(rtks)
        n = len(steps)
        st = (ctypes.c_int * max(1, n))(*steps)
        kb = (ctypes.c_void_p * max(1, n))(*[rtks[s][0].data_ptr() for s in steps])
        ka = (ctypes.c_void_p * max(1, n))(*[rtks[s][1].data_ptr() for s in steps])
        self._keep += [params, evk, rtks]
        rc = self.lib.aoa_server_register(self.h, params.h, _ptr(evk[0]), _ptr(evk[1]), n, st, kb, ka)
        if rc < 0:
            check(-rc)
        return rc

    def submit(self, keyset, c, level, in_ntt, m, ops, out):
        """ops: list of ("mul",) / ("rot", slots)."""
        arr = (Op * max(1, len(ops)))()
        for i, op in enumerate(ops):
            arr[i].kind = self.MUL if op[0] == "mul" else self.ROTATE
            arr[i].rot_slots = op[1] if op[0] == "rot" else 0
        rid = ctypes.c_int64()
        check(self.lib.aoa_server_submit(self.h, keyset, _ptr(c[0]), _ptr(c[1]), level, int(in_ntt),
                                         _ptr(m[0]) if m else None, _ptr(m[1]) if m else None, arr, len(ops),
                                         _ptr(out[0]), _ptr(out[1]), ctypes.byref(rid)))
        self._keep += [c, m, out]
        return rid.value

    def step(self, stream=None):
        rc = self.lib.aoa_server_step(self.h, _stream(stream) if stream is not None else None)
        if rc < 0:
            check(-rc)
        return rc

    def start(self):
        check(self.lib.aoa_server_start(self.h))

    def stop(self):
        check(self.lib.aoa_server_stop(self.h))

    def result(self, rid):
        st, lv, lat = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(self.lib.aoa_server_result(self.h, rid, ctypes.byref(st), ctypes.byref(lv), ctypes.byref(lat)))
        return st.value, lv.value, lat.value

    def release(self, rid):
        check(self.lib.aoa_server_release(self.h, rid))

    def stats(self):
        c, o, s = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        span, used, keys = ctypes.c_double(), ctypes.c_size_t(), ctypes.c_size_t()
        check(self.lib.aoa_server_stats(self.h, ctypes.byref(c), ctypes.byref(o), ctypes.byref(s),
                                        ctypes.byref(span), ctypes.byref(used), ctypes.byref(keys)))
        return {"completed": c.value, "ops": o.value, "steps": s.value, "span_ms": span.value,
                "pool_used": used.value, "key_bytes": keys.value}

    def close(self):
        if self.h:
            self.lib.aoa_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
