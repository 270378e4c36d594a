"""Inputs aimed at the reference's ModDown deviation window (test helpers over
the oracle; no GPU).

ModDown (aoa-ntt/src/cuda/manager.cu:312-341) reduces each cross-base product
a_j c_ij, a_j = [x_bj (B/b_j)^-1]_bj < b_j and c_ij = (B/b_j) mod q_i, with the
reference's Barrett `mod` (aoa-ntt/src/cuda/htrans/common.cu:257-291), which
returns r - eps instead of the exact remainder r when eps <= r < delta eps
(eps = 2^nb - q_i, delta the quotient error of that product).  Random inputs
land there about once in 10^4..10^6 products; these helpers put many products
there on purpose.
"""
import numpy as np


def window_bounds(O, level):
    """Per limb i <= level: (q, eps, thr, wide, cinv) with thr = delta_max eps
    (delta_max = floor(X1max rho / (q 2^64)) + 2), a wider bound from b_max q
    alone, and c_ij^-1 mod q_i for every j."""
    k, kl = O.k, O.kl
    c = O.table(5).reshape(k, kl)
    Q, B = O.primes[:k], O.primes[k:]
    bmax = max(B)
    out = []
    for i in range(level + 1):
        q = Q[i]
        nb = q.bit_length()
        eps = (1 << nb) - q
        x1 = (bmax - 1) * (q - 1) >> (2 * nb - 64)
        thr = (x1 * ((1 << (2 * nb)) % q) // (q << 64) + 2) * eps
        wide = ((bmax * q >> (2 * nb)) + 3) * eps
        cinv = [pow(int(c[i, j]), -1, q) if int(c[i, j]) else 0 for j in range(kl)]
        out.append((q, eps, thr, wide, cinv))
    return out


def window_aimed_qb(O, rng, level, frac=4):
    """A QB polynomial (coefficient domain) whose special residues put ModDown
    products at remainders inside the deviation window [eps, delta_max eps) of
    q_i, at its edges, or below the wider bound: for one coefficient in
    `frac`, every j aims at a random limb i <= level."""
    N, k, kl = O.N, O.k, O.kl
    inv_bhat = [int(v) for v in O.table(4)]
    B = O.primes[k:]
    bhat = [pow(inv_bhat[j], -1, B[j]) for j in range(kl)]
    win = window_bounds(O, level)
    qb = O.random_poly(rng, k + kl).reshape(k + kl, N)
    for cid in range(0, N, frac):
        for j in range(kl):
            q, eps, thr, wide, cinv = win[int(rng.integers(0, level + 1))]
            if wide >= q or cinv[j] == 0:
                continue
            pick = int(rng.integers(0, 6))
            target = [eps, thr - 1, thr, eps - 1, int(rng.integers(eps, thr)),
                      int(rng.integers(eps, wide))][pick] % q
            a = target * cinv[j] % q
            m = (B[j] - 1 - a) // q
            a += int(rng.integers(0, m + 1)) * q if m > 0 else 0
            a %= B[j]
            qb[k + j, cid] = a * bhat[j] % B[j]
    return qb.reshape(-1)


def aimed_key(O, rng, a1, b1, level, targets):
    """A switching key (key_b, key_a), NTT domain QB, for HomMul of inputs whose
    second polynomials are a1, b1 (NTT domain, Q limbs) at `level`: the
    special limbs of each key product INTT(ModUp(d2) key_y), d2 = a1 b1, equal
    the special limbs of targets[y] (coefficient domain QB), so the ModDown
    inside HomMul sees the targets' products.  Q limbs of the key are uniform.

    X_B = NTT(ModUp(INTT(d2))) on the special limbs, as the reference's
    poly_modup + poly_mul compute it (aoa-ntt/src/arithmetic/poly_t.cpp:345-361),
    and key_B = NTT(target_B) / X_B pointwise mod b_j."""
    N, k, kl = O.N, O.k, O.kl
    d2 = np.zeros(k * N, dtype=np.uint64)
    O.eltwise(2, d2, a1, b1, k)
    qb = np.zeros((k + kl) * N, dtype=np.uint64)
    qb[:k * N] = d2
    O.ntt(qb, 0, k, 1)
    O.modup(qb, level)
    O.ntt(qb, k, kl, 0)
    # X^-1 = X^(b_j - 2) on the special limbs, square-and-multiply with the
    # oracle's element-wise product (limb r uses prime r)
    inv = np.zeros_like(qb)
    inv[k * N:] = 1
    sq = qb.copy()
    for t in range(64):
        bits = np.array([0] * k + [(O.primes[k + j] - 2) >> t & 1 for j in range(kl)], dtype=bool)
        if not bits.any() and all((O.primes[k + j] - 2) >> t == 0 for j in range(kl)):
            break
        f = np.where(np.repeat(bits, N), sq, np.uint64(1))
        O.eltwise(2, inv, inv, f, k + kl)
        O.eltwise(2, sq, sq, sq, k + kl)
    keys = []
    for t in targets:
        tb = t.copy()
        O.ntt(tb, k, kl, 0)
        key = O.random_poly(rng, k + kl)
        prod = np.zeros_like(key)
        O.eltwise(2, prod, tb, inv, k + kl)
        nz = qb[k * N:] != 0
        key[k * N:] = np.where(nz, prod[k * N:], key[k * N:])
        keys.append(key)
    return keys
