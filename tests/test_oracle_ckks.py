"""Scheme-level pins of the oracle (CPU): a HomMul+relin+rescale of two
secret-key encryptions decrypts to round(m1*m2 / q_l) within a small noise
bound; a rotation decrypts to the automorphism of the plaintext.  They fix
the key-switching math (a dropped term, a wrong sign or a transposed operand
breaks decryption) independently of the oracle's own formulas.

They also document the reference's ModDown reduction: in the prec-48 layout
(55-bit special primes, wider than q_i) its Barrett reduction returns
exact - (2^48 - q_i) for some cross-base products, and once such a product
reaches the rescaled limbs the decryption is garbage; the exact reduction
decrypts correctly."""
import numpy as np
import pytest

import ckks_ref as C
from orc import Oracle


def _mul_roundtrip(N, k, kl, prec, exact, seed=5):
    o = Oracle(N, k, kl, prec)
    o.set_moddown_exact(exact)
    rng = np.random.default_rng(seed)
    s = C.ternary_secret(rng, N)
    evk = C.switching_key(o, rng, C.negacyclic(list(s), list(s), N), s)
    D = 1 << 40
    p1 = np.zeros(N, dtype=object)
    p2 = np.zeros(N, dtype=object)
    for i in rng.choice(N, 8, replace=False):
        p1[i] = int(rng.integers(-3, 4))
    for i in rng.choice(N, 8, replace=False):
        p2[i] = int(rng.integers(-3, 4))
    ct1 = C.encrypt(o, rng, p1 * D, s)
    ct2 = C.encrypt(o, rng, p2 * D, s)
    level = k - 1
    c30, c31 = o.hommul(ct1, ct2, evk, level)
    m = C.decrypt_limb0(o, c30, c31, s, in_ntt=False)
    ql = o.primes[level]
    exp = [round(v * D * D / ql) for v in C.negacyclic(list(p1), list(p2), N)]
    return max(abs(int(a) - int(b)) for a, b in zip(m, exp))


@pytest.mark.parametrize("shape", [(4096, 5, 0, 55), (4096, 4, 5, 52), (4096, 5, 0, 45)])
def test_hommul_decrypts(shape):
    assert _mul_roundtrip(*shape, exact=False) < 64


def test_reference_moddown_breaks_prec48_and_exact_mode_fixes_it():
    ref_err = _mul_roundtrip(4096, 30, 30, 48, exact=False)
    exact_err = _mul_roundtrip(4096, 30, 30, 48, exact=True)
    assert exact_err < 64
    assert ref_err > 1 << 40  # reproduced reference behaviour, not a target


def test_rotation_decrypts():
    N, k = 4096, 4
    o = Oracle(N, k, 0, 55)
    rng = np.random.default_rng(8)
    s = C.ternary_secret(rng, N)
    rot = 3
    g = pow(5, rot, 2 * N)
    s_rot = [0] * N  # sigma(s): X^c -> X^(c g mod 2N), negated on wrap
    for c, v in enumerate(s):
        e = c * g % (2 * N)
        if e >= N:
            s_rot[e - N] = -int(v)
        else:
            s_rot[e] = int(v)
    rtk = C.switching_key(o, rng, s_rot, s)
    D = 1 << 30
    m = np.zeros(N, dtype=object)
    for i in rng.choice(N, 10, replace=False):
        m[i] = int(rng.integers(-5, 6)) * D
    c0, c1 = C.encrypt(o, rng, m, s)
    c0 = C.ntt(o, c0, k, 1)
    c1 = C.ntt(o, c1, k, 1)
    r0, r1 = o.rotate_left((c0, c1), rtk, rot, k - 1)
    dec = C.decrypt_limb0(o, r0, r1, s, in_ntt=False)
    exp = [0] * N
    for c, v in enumerate(m):
        e = c * g % (2 * N)
        if e >= N:
            exp[e - N] = -v
        else:
            exp[e] = v
    assert max(abs(int(a) - int(b)) for a, b in zip(dec, exp)) < 1 << 12


def test_c1_pipeline_decrypts():
    """The oracle's end-to-end sequence (keys from the sampler restatement,
    encode, public-key encryption, HomMul+relin+rescale, decryption, decode)
    gives the FP64 product within the reference's bound (tests.cpp:34) for two
    encryption nonces; the keys satisfy their defining relations: pk.b + pk.a s
    and evk.b + evk.a s - rho(s^2) are small errors (|e| <= 32)."""
    import numpy as np
    from orc import Oracle
    o = Oracle(4096, 4, 5, 52)
    keys = o.c1_keys(42)
    rng = np.random.default_rng(1)
    z1 = rng.uniform(-1, 1, 2048) + 1j * rng.uniform(-1, 1, 2048)
    z2 = rng.uniform(-1, 1, 2048) + 1j * rng.uniform(-1, 1, 2048)
    for nonce in (5, 100):
        out, _ = o.c1_run(keys, z1, z2, 1 << 52, 42, nonce)
        assert np.max(np.abs(out - z1 * z2)) < 0.005
    N, k, R = o.N, o.k, o.k + o.kl

    def small(x, nres):
        y = x.copy()
        o.ntt(y, 0, nres, 1)
        for r in range(nres):
            p = o.primes[r]
            v = y[r * N:(r + 1) * N].astype(object)
            c = np.array([int(a) - p if int(a) > p // 2 else int(a) for a in v])
            assert np.abs(c).max() <= 32, r
        return y

    t = np.zeros(R * N, dtype=np.uint64)
    o.eltwise(2, t, keys["pka"], keys["s"], k)
    u = np.zeros(R * N, dtype=np.uint64)
    o.eltwise(0, u, keys["pkb"], t, k)
    small(u, k)
    s2 = np.zeros(R * N, dtype=np.uint64)
    o.eltwise(2, s2, keys["s"], keys["s"], R)
    bq = o.table(7)
    rho = np.zeros(R * N, dtype=np.uint64)
    for r in range(k):
        rho[r * N:(r + 1) * N] = [(int(a) * int(bq[r])) % o.primes[r] for a in s2[r * N:(r + 1) * N]]
    o.eltwise(2, t, keys["evka"], keys["s"], R)
    o.eltwise(0, u, keys["evkb"], t, R)
    o.eltwise(1, u, u, rho, R)
    small(u, R)
