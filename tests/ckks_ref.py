"""Small CKKS helpers over the oracle for CPU pins: secret-key encryption,
switching keys and decryption in residue form (no FFT; plaintexts are integer
polynomials).  Test infrastructure only."""
import numpy as np


def residues(o, ints, nlimbs, first=0):
    """Signed integer coefficients -> residue-major uint64 array."""
    N = o.N
    out = np.zeros(nlimbs * N, dtype=np.uint64)
    for r in range(nlimbs):
        p = o.primes[first + r]
        out[r * N:(r + 1) * N] = np.array([int(v) % p for v in ints], dtype=np.uint64)
    return out


def ntt(o, x, nres, direction):
    y = x.copy()
    o.ntt(y, 0, nres, direction)
    return y


def mul(o, a, b, nres):
    c = np.zeros_like(a)
    o.eltwise(2, c, a, b, nres)
    return c


def add(o, a, b, nres):
    c = np.zeros_like(a)
    o.eltwise(0, c, a, b, nres)
    return c


def sub(o, a, b, nres):
    c = np.zeros_like(a)
    o.eltwise(1, c, a, b, nres)
    return c


def ternary_secret(rng, N, weight=64):
    s = np.zeros(N, dtype=np.int64)
    idx = rng.choice(N, weight, replace=False)
    s[idx] = rng.choice([-1, 1], size=weight)
    return s


def gaussian(rng, N, sigma=3.2):
    return np.rint(rng.normal(0, sigma, size=N)).astype(np.int64)


def switching_key(o, rng, s1_ints, s_ints):
    """QB key (b, a) in the NTT domain: b = rho(s1) + e - a s with rho = (B mod q_i)
    on Q limbs and 0 on special limbs (aoa-ntt/src/ckks/ckkskeys.cpp:101-131)."""
    R = o.k + o.kl
    N = o.N
    B = 1
    for b in o.primes[o.k:]:
        B *= b
    a = o.random_poly(rng, R)
    s = ntt(o, residues(o, s_ints, R), R, 0)
    s1 = residues(o, s1_ints, R)
    rho = np.zeros_like(s1)
    for r in range(o.k):
        p = o.primes[r]
        rho[r * N:(r + 1) * N] = np.array([(int(v) * (B % p)) % p for v in s1[r * N:(r + 1) * N]], dtype=np.uint64)
    e = residues(o, gaussian(rng, N), R)
    b_coef = add(o, rho, e, R)
    b = sub(o, ntt(o, b_coef, R, 0), mul(o, a, s, R), R)
    return b, a


def encrypt(o, rng, m_ints, s_ints):
    """Secret-key encryption in the NTT domain: (c0, c1) = (-a s + m + e, a)."""
    k, N = o.k, o.N
    a = o.random_poly(rng, k)
    s = ntt(o, residues(o, s_ints, k), k, 0)
    me = residues(o, np.asarray(m_ints, dtype=object) + gaussian(rng, N).astype(object), k)
    c0 = sub(o, ntt(o, me, k, 0), mul(o, a, s, k), k)
    return c0, a


def decrypt_limb0(o, c0, c1, s_ints, in_ntt):
    """Centered limb-0 coefficients of c0 + c1 s (the reference decodes residue 0)."""
    k, N = o.k, o.N
    if not in_ntt:
        c0, c1 = ntt(o, c0, k, 0), ntt(o, c1, k, 0)
    s = ntt(o, residues(o, s_ints, k), k, 0)
    m = ntt(o, add(o, c0, mul(o, c1, s, k), k), k, 1)
    p = o.primes[0]
    return np.array([int(v) - p if int(v) > p // 2 else int(v) for v in m[:N]], dtype=object)


def negacyclic(a, b, N):
    c = [0] * N
    for i, x in enumerate(a):
        if x == 0:
            continue
        for j, y in enumerate(b):
            if y == 0:
                continue
            if i + j < N:
                c[i + j] += x * y
            else:
                c[i + j - N] -= x * y
    return c
