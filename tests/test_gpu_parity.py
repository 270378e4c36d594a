"""Bit-exact parity of the HIP kernels (through the C-ABI) with the CPU oracle.

Every integer output (NTT, element-wise, rescale, ModUp, ModDown, key switch,
HomMul+relin+rescale, rotation) must equal the oracle's residues exactly on
the same seeded uniform inputs; for Q polynomials at level l the contract is
limbs 0..l (the reference leaves stale values above the level).
"""
import numpy as np
import pytest

from aim import window_aimed_qb
from orc import Oracle

pytestmark = pytest.mark.gpu

# (N, k, kl, prec): N=4096 k=4 prec 52, small/medium shapes, and the reference's test params
# (N=2048 and 4096 run the single-workgroup transform, larger N the two-pass one).
SMALL = [(4096, 4, 5, 52), (8192, 6, 0, 48), (16384, 5, 5, 55), (32768, 4, 0, 48), (4096, 3, 3, 45),
         (2048, 2, 2, 55)]


def _params(N, k, kl, prec):
    from aoa_amd import Params
    return Params(N, k, kl, prec), Oracle(N, k, kl, prec)


def test_ntt_single_and_two_pass_agree(gpu):
    """At N=4096 the single-workgroup transform (split 4096) and the two-pass
    one (split 64) give the oracle's residues."""
    P, O = _params(4096, 4, 5, 52)
    rng = np.random.default_rng(12)
    R = O.k + O.kl
    x = O.random_poly(rng, R)
    ref = x.copy()
    O.ntt(ref, 0, R, 0)
    for n1 in (64, 4096):
        P.set_ntt_split(n1)
        d = P.to_device(x)
        P.ntt(d, 0, R, 0)
        assert np.array_equal(P.to_host(d), ref), "split %d" % n1
        P.ntt(d, 0, R, 1)
        assert np.array_equal(P.to_host(d), x), "split %d inverse" % n1


@pytest.mark.parametrize("shape", SMALL + [(65536, 3, 0, 48), (2048, 20, 20, 55)])
def test_ntt_parity(gpu, shape):
    P, O = _params(*shape)
    rng = np.random.default_rng(11)
    R = O.k + O.kl
    x = O.random_poly(rng, R)
    d = P.to_device(x)
    P.ntt(d, 0, R, 0)
    ref = x.copy()
    O.ntt(ref, 0, R, 0)
    got = P.to_host(d)
    assert np.array_equal(got, ref), "forward NTT mismatch %s" % (shape,)
    P.ntt(d, 0, R, 1)
    assert np.array_equal(P.to_host(d), x), "inverse NTT is not the exact inverse"
    # a sub-range of limbs (special limbs only) uses the right primes
    y = O.random_poly(rng, O.kl, first=O.k)
    full = np.zeros(R * O.N, dtype=np.uint64)
    full[O.k * O.N:] = y
    dd = P.to_device(full)
    P.ntt(dd, O.k, O.kl, 0)
    ref2 = full.copy()
    O.ntt(ref2, O.k, O.kl, 0)
    assert np.array_equal(P.to_host(dd), ref2)


@pytest.mark.parametrize("op", [0, 1, 2, 3])
def test_eltwise_parity(gpu, op):
    P, O = _params(8192, 5, 0, 48)
    rng = np.random.default_rng(3 + op)
    R = O.k + O.kl
    a = O.random_poly(rng, R)
    b = O.random_poly(rng, R)
    a[:7] = 0  # zeros exercise the NEGATE special case
    out = P.zeros(R)
    P.eltwise(op, out, P.to_device(a), P.to_device(b), res0=0, nres=R)
    ref = np.zeros_like(a)
    O.eltwise(op, ref, a, b, R)
    assert np.array_equal(P.to_host(out), ref)


@pytest.mark.parametrize("op", [0, 1, 2])
def test_eltwise_scalar_parity(gpu, op):
    """Scalar operand used as given (ntt.cu:531-558), including scalars above
    the primes and sums that wrap in 64 bits."""
    P, O = _params(8192, 5, 0, 48)
    rng = np.random.default_rng(30 + op)
    R = O.k + O.kl
    a = O.random_poly(rng, R)
    a[:5] = 0
    for b in (0, 7, O.primes[1] - 1, O.primes[1] + 3, (1 << 52) + 9, (1 << 60) + 1, (1 << 64) - 5):
        out = P.zeros(R)
        P.eltwise_scalar(op, out, P.to_device(a), b, res0=0, nres=R)
        ref = np.zeros_like(a)
        O.eltwise_scalar(op, ref, a, b, R)
        assert np.array_equal(P.to_host(out), ref), "op %d b %d" % (op, b)
    # a limb range starting inside the special limbs
    out = P.zeros(R)
    P.eltwise_scalar(op, out, P.to_device(a), 12345, res0=O.k + 1, nres=O.kl - 1)
    ref = np.zeros_like(a)
    O.eltwise_scalar(op, ref, a, 12345, R)  # the oracle numbers residues from 0: compare the tail
    assert np.array_equal(P.to_host(out)[(O.k + 1) * O.N:], ref[(O.k + 1) * O.N:])


def test_conjugate_parity(gpu):
    P, O = _params(16384, 4, 0, 52)
    rng = np.random.default_rng(35)
    R = O.k + O.kl
    a = O.random_poly(rng, R)
    out = P.zeros(R)
    P.conjugate(out, P.to_device(a), R)
    ref = np.zeros_like(a)
    O.conjugate(ref, a, R)
    assert np.array_equal(P.to_host(out), ref)


@pytest.mark.parametrize("shape", SMALL)
def test_rescale_modup_moddown_parity(gpu, shape):
    P, O = _params(*shape)
    rng = np.random.default_rng(5)
    k, kl, N = O.k, O.kl, O.N
    for level in sorted({1, k - 1}):
        # rescale
        a = O.random_poly(rng, k)
        b = O.random_poly(rng, k)
        da, db = P.to_device(a), P.to_device(b)
        P.rescale(da, db, level)
        O.rescale(a, b, level)
        assert np.array_equal(P.to_host(da)[:level * N], a[:level * N])
        assert np.array_equal(P.to_host(db)[:level * N], b[:level * N])
        # ModUp: Q limbs 0..level -> special limbs
        qb = O.random_poly(rng, k + kl)
        dqb = P.to_device(qb)
        P.modup(dqb, level)
        O.modup(qb, level)
        assert np.array_equal(P.to_host(dqb)[k * N:], qb[k * N:]), "ModUp level %d" % level
        # ModDown: QB -> Q limbs 0..level
        qb2 = O.random_poly(rng, k + kl)
        out = P.zeros(k)
        P.moddown(out, P.to_device(qb2), level)
        ref = np.zeros(k * N, dtype=np.uint64)
        O.moddown(ref, qb2, level)
        assert np.array_equal(P.to_host(out)[:(level + 1) * N], ref[:(level + 1) * N]), "ModDown level %d" % level


@pytest.mark.parametrize("shape", [(4096, 4, 5, 52), (16384, 8, 0, 48), (65536, 30, 0, 48)])
def test_moddown_window_products_parity(gpu, shape):
    """ModDown on inputs whose cross-base products sit in or at the edges of
    the reference's Barrett deviation window: the screened fast path and the
    per-term literal recomputation give the oracle's residues, and the exact
    mode gives the oracle's exact mode.  (The fused rescale and the batched
    kernel see such products through test_gpu_shapes.py's aimed keys.)"""
    P, O = _params(*shape)
    rng = np.random.default_rng(41)
    k, N = O.k, O.N
    for level in sorted({1, k - 1}):
        qb = window_aimed_qb(O, rng, level)
        for exact in (False, True):
            P.set_moddown_exact(exact)
            O.set_moddown_exact(exact)
            out = P.zeros(k)
            P.moddown(out, P.to_device(qb), level)
            ref = np.zeros(k * N, dtype=np.uint64)
            O.moddown(ref, qb, level)
            assert np.array_equal(P.to_host(out)[:(level + 1) * N], ref[:(level + 1) * N]), \
                "ModDown level %d exact %d" % (level, exact)
    P.set_moddown_exact(False)
    O.set_moddown_exact(False)


@pytest.mark.parametrize("shape", SMALL)
def test_hommul_parity(gpu, shape):
    P, O = _params(*shape)
    rng = np.random.default_rng(7)
    k, kl, N = O.k, O.kl, O.N
    for level in sorted({1, k - 1}):
        c1 = [O.random_poly(rng, k) for _ in range(2)]
        c2 = [O.random_poly(rng, k) for _ in range(2)]
        evk = [O.random_poly(rng, k + kl) for _ in range(2)]
        ref0, ref1 = O.hommul(c1, c2, evk, level)
        d = [P.to_device(x) for x in c1 + c2 + evk]
        out = [P.zeros(k), P.zeros(k)]
        ws = P.workspace(level)
        P.hommul(out, (d[0], d[1]), (d[2], d[3]), (d[4], d[5]), level, ws)
        n = level * N
        assert np.array_equal(P.to_host(out[0])[:n], ref0[:n]), "c0 level %d" % level
        assert np.array_equal(P.to_host(out[1])[:n], ref1[:n]), "c1 level %d" % level


@pytest.mark.parametrize("n1", [64, 4096])
def test_hommul_split_parity(gpu, n1):
    """HomMul at N=4096 with the two-pass split (the additions of d0 and d1
    join the key products before the inverse NTT) and with the
    single-workgroup transform (the additions stay in ModDown): both give the
    oracle's residues, at the top level and at level 1."""
    P, O = _params(4096, 4, 5, 52)
    P.set_ntt_split(n1)
    rng = np.random.default_rng(17)
    k, kl, N = O.k, O.kl, O.N
    for level in (1, k - 1):
        c1 = [O.random_poly(rng, k) for _ in range(2)]
        c2 = [O.random_poly(rng, k) for _ in range(2)]
        evk = [O.random_poly(rng, k + kl) for _ in range(2)]
        ref0, ref1 = O.hommul(c1, c2, evk, level)
        d = [P.to_device(x) for x in c1 + c2 + evk]
        out = [P.zeros(k), P.zeros(k)]
        P.hommul(out, (d[0], d[1]), (d[2], d[3]), (d[4], d[5]), level, P.workspace(level))
        n = level * N
        assert np.array_equal(P.to_host(out[0])[:n], ref0[:n]), "c0 split %d level %d" % (n1, level)
        assert np.array_equal(P.to_host(out[1])[:n], ref1[:n]), "c1 split %d level %d" % (n1, level)


@pytest.mark.parametrize("shape", SMALL[:3])
def test_rotate_parity(gpu, shape):
    P, O = _params(*shape)
    rng = np.random.default_rng(9)
    k, kl, N = O.k, O.kl, O.N
    level = k - 1
    for rot in (1, 4, N // 4 - 1):
        c = [O.random_poly(rng, k) for _ in range(2)]
        rtk = [O.random_poly(rng, k + kl) for _ in range(2)]
        ref0, ref1 = O.rotate_left(c, rtk, rot, level)
        d = [P.to_device(x) for x in c + rtk]
        out = [P.zeros(k), P.zeros(k)]
        P.rotate_left(out, (d[0], d[1]), (d[2], d[3]), rot, level, P.workspace(level))
        n = (level + 1) * N
        assert np.array_equal(P.to_host(out[0])[:n], ref0[:n]), "rot %d c0" % rot
        assert np.array_equal(P.to_host(out[1])[:n], ref1[:n]), "rot %d c1" % rot


def test_hommul_c2_parity(gpu):
    """N=2^16, 30 Q limbs (prec 48), 31 special limbs, top level."""
    P, O = _params(65536, 30, 0, 48)
    rng = np.random.default_rng(2)
    k, kl, N = O.k, O.kl, O.N
    level = k - 1
    c1 = [O.random_poly(rng, k) for _ in range(2)]
    c2 = [O.random_poly(rng, k) for _ in range(2)]
    evk = [O.random_poly(rng, k + kl) for _ in range(2)]
    ref0, ref1 = O.hommul(c1, c2, evk, level)
    d = [P.to_device(x) for x in c1 + c2 + evk]
    out = [P.zeros(k), P.zeros(k)]
    P.hommul(out, (d[0], d[1]), (d[2], d[3]), (d[4], d[5]), level, P.workspace(level))
    n = level * N
    assert np.array_equal(P.to_host(out[0])[:n], ref0[:n])
    assert np.array_equal(P.to_host(out[1])[:n], ref1[:n])


@pytest.mark.parametrize("shape", [(8192, 6, 0, 48), (65536, 30, 0, 48)])
def test_hommul_exact_moddown_parity(gpu, shape):
    """Exact ModDown mode (not in the reference) against the oracle's exact mode;
    the default mode stays bit-exact with the reference's reduction."""
    P, O = _params(*shape)
    P.set_moddown_exact(True)
    O.set_moddown_exact(True)
    rng = np.random.default_rng(71)
    k, kl, N = O.k, O.kl, O.N
    level = k - 1
    c1 = [O.random_poly(rng, k) for _ in range(2)]
    c2 = [O.random_poly(rng, k) for _ in range(2)]
    evk = [O.random_poly(rng, k + kl) for _ in range(2)]
    ref0, ref1 = O.hommul(c1, c2, evk, level)
    d = [P.to_device(x) for x in c1 + c2 + evk]
    out = [P.zeros(k), P.zeros(k)]
    P.hommul(out, (d[0], d[1]), (d[2], d[3]), (d[4], d[5]), level, P.workspace(level))
    n = level * N
    assert np.array_equal(P.to_host(out[0])[:n], ref0[:n])
    assert np.array_equal(P.to_host(out[1])[:n], ref1[:n])
    # the two modes differ somewhere at this shape (the reference's reduction deviates)
    O.set_moddown_exact(False)
    r0, _ = O.hommul(c1, c2, evk, level)
    if shape[1] >= 30:
        assert not np.array_equal(r0[:n], ref0[:n])


@pytest.mark.parametrize("shape", [(2048, 2, 2, 55), (8192, 6, 0, 48), (65536, 30, 0, 48)])
def test_manager_entry_points_parity(gpu, shape):
    """CUDAManager's execute_dot (1, 3 and 8 terms; sums of 63-bit products
    beyond Barrett's range included), execute_dr2 and execute_modup from a
    separate input, against the oracle."""
    P, O = _params(*shape)
    rng = np.random.default_rng(23)
    k, kl, N = O.k, O.kl, O.N
    R = k + kl
    for kk in (1, 3, 8):
        a = [O.random_poly(rng, R) for _ in range(kk)]
        b = [O.random_poly(rng, R) for _ in range(kk)]
        out = P.zeros(R)
        keep = P.dot(out, [P.to_device(x) for x in a], [P.to_device(x) for x in b])
        assert np.array_equal(P.to_host(out), O.dot(a, b)), "dot %d terms" % kk
        del keep
    x = [O.random_poly(rng, R) for _ in range(4)]
    c = [P.zeros(R) for _ in range(3)]
    P.dr2(*c, *[P.to_device(v) for v in x])
    for got, ref in zip(c, O.dr2(*x)):
        assert np.array_equal(P.to_host(got), ref), "dr2"
    for level in sorted({0, k - 1}):
        q = O.random_poly(rng, k)
        qb = np.zeros(R * N, dtype=np.uint64)
        qb[:k * N] = q
        O.modup(qb, level)
        dq = P.to_device(q)
        dqb = P.zeros(R)
        P.modup_from(dqb, dq, level)
        assert np.array_equal(P.to_host(dqb)[k * N:], qb[k * N:]), "ModUp from level %d" % level
