"""Parity at the full workload shapes: the exact batch the throughput figure is
measured on, rotations at N=2^16 (left and right) against the oracle, and
key products aimed into the reference's ModDown deviation window reaching the
ModDown of HomMul (rescale fused) in the plain and the batched (descriptor)
kernels."""
import numpy as np
import pytest

from aim import aimed_key, window_aimed_qb
from orc import Oracle

pytestmark = pytest.mark.gpu


def _sampled(P, nlimbs, nonce, seed=77):
    """Uniform residues drawn on the device (limb r uses prime r)."""
    from aoa_amd import SAMPLE_UNIFORM
    t = P.zeros(nlimbs)
    P.sample(t, SAMPLE_UNIFORM, seed, nonce, res0=0, nres=nlimbs)
    return t


def test_batch_headline_equals_plain_and_oracle(gpu):
    """32 HomMul+relin+rescale at N=2^16, k=30 (prec 48), K=31, level 29 in one
    graph, the throughput workload: every item of two graph replays equals the
    plain call (compared on the device) and the first item equals the oracle."""
    import torch
    from aoa_amd import Batch, Params
    N, k = 65536, 30
    P = Params(N, k, 0, 48)
    level = k - 1
    R = k + P.kl
    print("headline batch: 32 x N=%d k=%d level %d" % (N, k, level), flush=True)
    evk = [_sampled(P, R, 1), _sampled(P, R, 2)]
    items = []
    for i in range(32):
        c1 = [_sampled(P, k, 10 + 4 * i), _sampled(P, k, 11 + 4 * i)]
        c2 = [_sampled(P, k, 12 + 4 * i), _sampled(P, k, 13 + 4 * i)]
        items.append((c1, c2, [P.zeros(k), P.zeros(k)], level))
    B = Batch()
    B.add_mul(P, evk, items)
    B.build()
    n = level * N
    ws = P.workspace(level)
    ref = [P.zeros(k), P.zeros(k)]
    for rep in range(2):
        for it in items:
            it[2][0].zero_()
            it[2][1].zero_()
        B.launch(graph=True)
        torch.cuda.synchronize()
        for i, (c1, c2, out, lv) in enumerate(items):
            P.hommul(ref, c1, c2, evk, lv, ws)
            assert torch.equal(out[0][:n], ref[0][:n]) and torch.equal(out[1][:n], ref[1][:n]), \
                "replay %d item %d" % (rep, i)
    O = Oracle(N, k, 0, 48)
    c1, c2, out, lv = items[0]
    h = [P.to_host(t) for t in c1 + c2 + evk]
    o0, o1 = O.hommul((h[0], h[1]), (h[2], h[3]), (h[4], h[5]), lv)
    assert np.array_equal(P.to_host(out[0])[:n], o0[:n])
    assert np.array_equal(P.to_host(out[1])[:n], o1[:n])


@pytest.mark.parametrize("rot", [1, 1 << 14])
def test_rotate_n16_parity(gpu, rot):
    """Rotations at N=2^16, k=24 (prec 48), K=25, top level, left and right
    (ckks.cpp:953-1009), against the oracle."""
    from aoa_amd import Params
    N, k = 65536, 24
    P, O = Params(N, k, 0, 48), Oracle(N, k, 0, 48)
    rng = np.random.default_rng(90 + rot % 7)
    level = k - 1
    print("rotation N=%d k=%d step %d" % (N, k, rot), flush=True)
    c = [O.random_poly(rng, k) for _ in range(2)]
    rtk = [O.random_poly(rng, k + O.kl) for _ in range(2)]
    d = [P.to_device(x) for x in c + rtk]
    ws = P.workspace(level)
    n = (level + 1) * N
    for side in ("left", "right"):
        ref0, ref1 = getattr(O, "rotate_" + side)(c, rtk, rot, level)
        out = [P.zeros(k), P.zeros(k)]
        getattr(P, "rotate_" + side)(out, (d[0], d[1]), (d[2], d[3]), rot, level, ws)
        assert np.array_equal(P.to_host(out[0])[:n], ref0[:n]), "%s c0" % side
        assert np.array_equal(P.to_host(out[1])[:n], ref1[:n]), "%s c1" % side


@pytest.mark.parametrize("shape", [(4096, 4, 5, 52), (16384, 8, 0, 48), (65536, 30, 0, 48)])
def test_hommul_window_aimed_keys(gpu, shape):
    """HomMul+relin+rescale with keys chosen so that the special limbs of both
    key products put many ModDown products in the reference's deviation
    window: the plain form (k_moddown, rescale fused) and a batch of the same
    item at two levels (k_moddown_d) give the oracle's residues, and the
    window is really hit (the oracle's exact mode differs)."""
    import torch
    from aoa_amd import Batch, Params
    N, k, kl, prec = shape
    P, O = Params(*shape), Oracle(*shape)
    rng = np.random.default_rng(61)
    frac = 4 if N <= 16384 else 16
    print("aimed keys %s" % (shape,), flush=True)
    cases = []
    for level in sorted({k - 1, max(1, k // 2)}):
        c1 = [O.random_poly(rng, k) for _ in range(2)]
        c2 = [O.random_poly(rng, k) for _ in range(2)]
        tg = [window_aimed_qb(O, rng, level, frac) for _ in range(2)]
        evk = aimed_key(O, rng, c1[1], c2[1], level, tg)
        ref = O.hommul(c1, c2, evk, level)
        O.set_moddown_exact(True)
        ex = O.hommul(c1, c2, evk, level)
        O.set_moddown_exact(False)
        n = level * N
        assert not np.array_equal(ref[0][:n], ex[0][:n]), "the aimed products never deviate"
        cases.append((level, [P.to_device(x) for x in c1 + c2 + evk], ref))
    for level, d, ref in cases:
        n = level * N
        out = [P.zeros(k), P.zeros(k)]
        P.hommul(out, (d[0], d[1]), (d[2], d[3]), (d[4], d[5]), level, P.workspace(level))
        assert np.array_equal(P.to_host(out[0])[:n], ref[0][:n]), "plain c0 level %d" % level
        assert np.array_equal(P.to_host(out[1])[:n], ref[1][:n]), "plain c1 level %d" % level
    # batched: each case is its own group (its own key), one launch per stage
    B = Batch()
    outs = []
    for level, d, ref in cases:
        out = [P.zeros(k), P.zeros(k)]
        outs.append(out)
        B.add_mul(P, (d[4], d[5]), [((d[0], d[1]), (d[2], d[3]), out, level)])
    B.launch(graph=True)
    torch.cuda.synchronize()
    for (level, d, ref), out in zip(cases, outs):
        n = level * N
        assert np.array_equal(P.to_host(out[0])[:n], ref[0][:n]), "batched c0 level %d" % level
        assert np.array_equal(P.to_host(out[1])[:n], ref[1][:n]), "batched c1 level %d" % level
