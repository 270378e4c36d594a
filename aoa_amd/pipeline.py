"""One plaintext product end to end on the device, over the C-ABI: key
generation, then encode -> encrypt -> HomMul+relin+rescale -> decrypt ->
decode.

The sequence is the one the newckks API runs (ckks_keygen, ckks_encrypt,
ckks_mul, ckks_decrypt: aoa-ntt/src/ckks/ckkskeys.cpp:61-131,
aoa-ntt/src/ckks/ckks.cpp:87-132, :273-358), written out call by call (keys
from aoa_sample / aoa_ntt / aoa_eltwise, then aoa_encode, aoa_encrypt,
aoa_hommul_relin_rescale, aoa_decrypt, aoa_decode) so that it can be
replayed as one HIP graph.  Keys live in the NTT domain: the secret
s on the QB limbs, the public key (b = e - a s) on the Q limbs, the
relinearisation key (b = rho(s^2) + e - a s) on the QB limbs.  aoa_sample
nonces: s 0, public a 1, public e 2, relinearisation a 3 and e 4; an
encryption at nonce n draws u (ZO) at n, e0 at n+1 and e1 at n+2.
"""
from .core import ADD, FORWARD, MUL, SAMPLE_GAUSSIAN, SAMPLE_HW, SAMPLE_UNIFORM, SUB


class Pipeline:
    def __init__(self, params, seed, slots):
        import torch
        P = params
        k, R = P.k, P.k + P.kl
        if k < 2:
            raise ValueError("the pipeline needs k >= 2 (one rescale)")
        if slots < 1 or slots > P.N // 2 or slots & (slots - 1):
            raise ValueError("slots must be a power of two <= N/2")
        self.P, self.seed, self.slots = P, seed, slots
        z = lambda n: P.zeros(n)
        self.s = z(R)
        P.sample(self.s, SAMPLE_HW, seed, 0, 0, R)
        P.ntt(self.s, 0, R, FORWARD)
        e, t = z(R), z(R)
        self.pka = z(k)
        P.sample(self.pka, SAMPLE_UNIFORM, seed, 1, 0, k)
        P.sample(e, SAMPLE_GAUSSIAN, seed, 2, 0, k)
        P.ntt(e, 0, k, FORWARD)
        P.eltwise(MUL, t, self.pka, self.s, res0=0, nres=k)
        self.pkb = z(k)
        P.eltwise(SUB, self.pkb, e, t, res0=0, nres=k)
        self.evka = z(R)
        P.sample(self.evka, SAMPLE_UNIFORM, seed, 3, 0, R)
        P.sample(e, SAMPLE_GAUSSIAN, seed, 4, 0, R)
        P.ntt(e, 0, R, FORWARD)
        s2, rho = z(R), z(R)
        P.eltwise(MUL, s2, self.s, self.s, res0=0, nres=R)
        P.rho(rho, s2)
        P.eltwise(MUL, t, self.evka, self.s, res0=0, nres=R)
        self.evkb = z(R)
        P.eltwise(ADD, self.evkb, rho, e, res0=0, nres=R)
        P.eltwise(SUB, self.evkb, self.evkb, t, res0=0, nres=R)
        # per-run buffers
        f64 = lambda: torch.zeros(2 * slots, dtype=torch.float64, device="cuda")
        self.z = [f64(), f64()]      # inputs (complex, re/im interleaved)
        self.v = [f64(), f64()]      # encoder work buffers
        self.out = f64()             # decoded product
        self.m, self.t = z(k), z(k)
        self.ews = z(3 * k)          # encryption's u, e0, e1
        self.c = [[z(k), z(k)], [z(k), z(k)]]
        self.c3 = [z(k), z(k)]
        self.ws = P.workspace(k - 1)
        self.scale = None

    def set_inputs(self, z1, z2, scale):
        import numpy as np
        import torch
        for d, zz in zip(self.z, (z1, z2)):
            a = np.ascontiguousarray(zz, dtype=np.complex128).view(np.float64)
            d.copy_(torch.from_numpy(a.copy()))
        self.scale = int(scale)
        q = self.P.primes[self.P.k - 1]
        self.scale_out = self.scale * self.scale // q

    def _encrypt(self, j, nonce):
        P, k = self.P, self.P.k
        self.v[j].copy_(self.z[j])
        P.encode(self.m, self.v[j], self.slots, self.scale, nres=k)
        # u (ZO), e0, e1 (Gaussian) at nonce, nonce+1, nonce+2; c0 = pk_b u + e0 + m, c1 = pk_a u + e1
        P.encrypt(self.c[j][0], self.c[j][1], self.m, self.pkb, self.pka, self.seed, nonce, self.ews)

    def run(self, nonce):
        """Encrypt both inputs (nonces `nonce` and `nonce + 3`), multiply,
        decrypt and decode into self.out; stream-ordered, no host sync."""
        P, k = self.P, self.P.k
        L = k - 1
        self._encrypt(0, nonce)
        self._encrypt(1, nonce + 3)
        P.hommul(self.c3, self.c[0], self.c[1], (self.evkb, self.evka), k - 1, self.ws)
        P.decrypt(self.t, self.c3[0], self.c3[1], self.s, L - 1)
        P.decode(self.out, self.t, self.slots, self.scale_out)

    def result(self):
        import numpy as np
        return self.out.cpu().numpy().view(np.complex128).copy()
