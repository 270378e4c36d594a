"""The device sampler's restatement in the oracle (Philox4x32-10 and the
mapping of its output to each distribution), pinned by the generator's
published known-answer vectors, by an independent normal CDF (scipy) for
the Gaussian table, and by the distributions' frequencies and moments.
The reference's own streams (cuRAND Sobol64 / XORWOW, host rand()) cannot be
reproduced; its distributions are what is kept (aoa-ntt/src/cuda/sampler.cu:21-225)."""
import math

import numpy as np
import pytest
from scipy.special import ndtr

import native_chk
import orc

UNIFORM, GAUSSIAN, ZO, HW, NARROW = 0, 1, 2, 3, 4


@pytest.mark.parametrize("ctr,key,expect", [
    # Random123 kat_vectors, philox4x32 10 rounds
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, expect):
    assert orc.philox4x32_10(ctr, key) == expect


def test_gauss_table():
    """Entry j is 2^64 P(round(3.2 z) <= j - 32) = 2^64 Phi((j - 31.5) / 3.2),
    against scipy's normal CDF (relative 1e-13 on the smaller of P and 1 - P);
    the product's table (built by params.cpp) is the same, word for word."""
    t = orc.gauss_cdt()
    assert np.array_equal(t, native_chk.table(2048, 2, 3, 55, 13))
    assert all(int(t[j]) <= int(t[j + 1]) for j in range(63))
    for j in range(64):
        x = (j - 32 + 0.5) / 3.2
        if x < 0:
            want = ndtr(x) * 2.0 ** 64
            assert abs(int(t[j]) - want) <= max(1.0, want * 1e-13), j
        else:
            tail = ndtr(-x) * 2.0 ** 64
            got_tail = 2 ** 64 - int(t[j])
            assert abs(got_tail - tail) <= max(1.0, tail * 1e-13), j


def _signed(col, q):
    c = col.astype(object)
    return np.array([int(v) - q if int(v) > q // 2 else int(v) for v in c])


@pytest.fixture(scope="module")
def ora():
    return orc.Oracle(65536, 3, 2, 48)


def test_uniform_range_and_chi2(ora):
    N, R = ora.N, ora.k + ora.kl
    out = ora.sample(UNIFORM, 7, 1).reshape(R, N)
    for l, q in enumerate(ora.primes):
        col = out[l].astype(np.float64)
        assert int(out[l].max()) < q
        counts = np.bincount(np.floor(col / q * 64).astype(int), minlength=64)
        chi2 = float(np.sum((counts - N / 64) ** 2 / (N / 64)))
        assert chi2 < 120, (l, chi2)  # 63 degrees of freedom; P(chi2 > 120) < 1e-5
    # limbs are drawn independently
    assert not np.array_equal(out[1], out[2])


def test_gaussian_moments_and_pmf(ora):
    N, R = ora.N, ora.k + ora.kl
    out = ora.sample(GAUSSIAN, 11, 2).reshape(R, N)
    e = _signed(out[0], ora.primes[0])
    for l in range(1, R):
        assert np.array_equal(_signed(out[l], ora.primes[l]), e)
    assert abs(e.mean()) < 5 * 3.2 / math.sqrt(N)
    var = 3.2 ** 2 + 1.0 / 12  # variance of a rounded N(0, 3.2^2)
    assert abs(e.var() - var) < 0.25
    assert np.abs(e).max() <= 32
    ks = np.arange(-12, 13)
    pmf = ndtr((ks + 0.5) / 3.2) - ndtr((ks - 0.5) / 3.2)
    counts = np.array([(e == k).sum() for k in ks])
    chi2 = float(np.sum((counts - N * pmf) ** 2 / (N * pmf)))
    assert chi2 < 70, chi2  # 24 degrees of freedom


def _zo_feistel(seed, nonce, logN, i):
    """The oracle's ZO permutation restated from Philox4x32-10 outputs: six
    rounds on (H, Lo), H the high ceil(logN/2) bits of i, Lo the rest."""
    hb, lb = (logN + 1) // 2, logN // 2
    h, lo = i >> lb, i & ((1 << lb) - 1)
    for r in range(6):
        tag = (ZO << 24) | (0xFFFF00 + r)
        if r & 1:
            h ^= orc.philox4x32_10([lo, tag, nonce, 0], [seed, 0])[0] & ((1 << hb) - 1)
        else:
            lo ^= orc.philox4x32_10([h, tag, nonce, 0], [seed, 0])[0] & ((1 << lb) - 1)
    return (h << lb) | lo


def test_zo_fixed_weight(ora):
    """ZO as sample_Z0 draws it (aoa-ntt/src/cuda/sampler.cu:119-150: N/2
    values +-1 shuffled over the N positions): exactly N/2 nonzeros, the same
    signed value in every limb, signs balanced, nonzeros spread evenly over
    the positions."""
    N, R = ora.N, ora.k + ora.kl
    for nonce in range(3):
        out = ora.sample(ZO, 5, nonce).reshape(R, N)
        v = _signed(out[0], ora.primes[0])
        assert int((v != 0).sum()) == N // 2
        assert set(np.unique(v)) == {-1, 0, 1}
        for l in range(1, R):
            assert np.array_equal(_signed(out[l], ora.primes[l]), v)
        nz = v[v != 0]
        assert abs(nz.mean()) < 5 / math.sqrt(N // 2)
        # nonzeros per block of 1024 positions: hypergeometric, sd about 16
        blocks = (v != 0).reshape(-1, 1024).sum(axis=1)
        assert np.abs(blocks - 512).max() < 6 * 16, blocks


@pytest.mark.parametrize("N", [2048, 8192])
def test_zo_permutation_restated(N):
    """The positions and signs at N = 2^11 and 2^13 (odd and even index
    widths) against a Python restatement of the Feistel permutation over
    Philox outputs, which is a bijection of [0, N)."""
    o = orc.Oracle(N, 2, 1, 55)
    seed, nonce = 9, 4
    logN = N.bit_length() - 1
    perm = [_zo_feistel(seed, nonce, logN, i) for i in range(N)]
    assert sorted(perm) == list(range(N))
    want = np.zeros(N, dtype=np.int64)
    for i, p in enumerate(perm):
        if p < N // 2:
            sign = orc.philox4x32_10([i, (ZO << 24) | 0xFFFFFF, nonce, 0], [seed, 0])[1] & 1
            want[i] = -1 if sign else 1
    v = _signed(o.sample(ZO, seed, nonce).reshape(3, N)[0], o.primes[0])
    assert np.array_equal(v, want)


@pytest.mark.parametrize("kind,probs", [(NARROW, {0: 1 / 3, 1: 1 / 3, -1: 1 / 3})])
def test_ternary_frequencies(ora, kind, probs):
    N, R = ora.N, ora.k + ora.kl
    out = ora.sample(kind, 5, 3).reshape(R, N)
    v = _signed(out[0], ora.primes[0])
    assert set(np.unique(v)) <= set(probs)
    for l in range(1, R):
        assert np.array_equal(_signed(out[l], ora.primes[l]), v)
    for val, p in probs.items():
        f = (v == val).mean()
        assert abs(f - p) < 5 * math.sqrt(p * (1 - p) / N), (val, f)


def test_hamming_weight(ora):
    N, R = ora.N, ora.k + ora.kl
    for nonce in range(4):
        out = ora.sample(HW, 9, nonce).reshape(R, N)
        v = _signed(out[0], ora.primes[0])
        assert int((v != 0).sum()) == 64
        assert set(np.unique(v)) == {-1, 0, 1}
        for l in range(1, R):
            assert np.array_equal(_signed(out[l], ora.primes[l]), v)


def test_counter_layout(ora):
    """Each value depends on (seed, nonce, kind, limb, coefficient) only: a
    sub-range of limbs equals the same limbs of a full draw; a new nonce or
    seed changes the draw; the same arguments repeat it."""
    N, R = ora.N, ora.k + ora.kl
    for kind in (UNIFORM, GAUSSIAN, ZO, NARROW, HW):
        full = ora.sample(kind, 3, 4).reshape(R, N)
        part = ora.sample(kind, 3, 4, res0=2, nres=2).reshape(R, N)
        assert np.array_equal(part[2:4], full[2:4])
        assert np.array_equal(ora.sample(kind, 3, 4).reshape(R, N), full)
        assert not np.array_equal(ora.sample(kind, 3, 5).reshape(R, N), full)
        assert not np.array_equal(ora.sample(kind, 4, 4).reshape(R, N), full)
