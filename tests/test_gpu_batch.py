"""Batched (stage-fused, hipGraph-replayed) form: every result equals the plain
one-ciphertext-per-call result bit for bit, alone and in the mix (mixed levels,
mixed ring dimensions, repeated graph replays), and the plain form equals the
oracle."""
import numpy as np
import pytest

from orc import Oracle

pytestmark = pytest.mark.gpu


def _rand(P, rng, nlimbs):
    out = np.empty(nlimbs * P.N, dtype=np.uint64)
    for r in range(nlimbs):
        out[r * P.N:(r + 1) * P.N] = rng.integers(0, P.primes[r], size=P.N, dtype=np.uint64)
    return out


def _mul_items(P, rng, levels):
    items = []
    for lv in levels:
        c1 = [P.to_device(_rand(P, rng, P.k)) for _ in range(2)]
        c2 = [P.to_device(_rand(P, rng, P.k)) for _ in range(2)]
        out = [P.zeros(P.k), P.zeros(P.k)]
        items.append((c1, c2, out, lv))
    return items


def _plain_mul(P, evk, item):
    c1, c2, _, lv = item
    out = [P.zeros(P.k), P.zeros(P.k)]
    P.hommul(out, c1, c2, evk, lv, P.workspace(lv))
    return [P.to_host(t)[:lv * P.N] for t in out]


def test_batch_mul_mixed_levels_equals_plain(gpu):
    import torch
    from aoa_amd import Batch, Params
    P = Params(32768, 20, 0, 48)  # N=2^15, k=20, K=21
    rng = np.random.default_rng(31)
    evk = [P.to_device(_rand(P, rng, P.k + P.kl)) for _ in range(2)]
    levels = [int(x) for x in rng.integers(1, 20, size=12)]
    items = _mul_items(P, rng, levels)
    ref = [_plain_mul(P, evk, it) for it in items]
    B = Batch()
    B.add_mul(P, evk, items)
    ws = B.build()
    assert ws > 0
    for rep in range(3):  # graph replays with eager work in between
        for it in items:
            it[2][0].zero_()
            it[2][1].zero_()
        B.launch(graph=True)
        torch.cuda.synchronize()
        for it, r in zip(items, ref):
            lv = it[3]
            assert np.array_equal(P.to_host(it[2][0])[:lv * P.N], r[0]), "replay %d level %d" % (rep, lv)
            assert np.array_equal(P.to_host(it[2][1])[:lv * P.N], r[1]), "replay %d level %d" % (rep, lv)
        _ = torch.ones(1000, device="cuda").sum().item()
    # the batch equals the oracle too
    O = Oracle(32768, 20, 0, 48)
    it = items[0]
    h = [P.to_host(t) for t in it[0] + it[1] + evk]
    o0, o1 = O.hommul((h[0], h[1]), (h[2], h[3]), (h[4], h[5]), it[3])
    n = it[3] * P.N
    assert np.array_equal(P.to_host(it[2][0])[:n], o0[:n])
    assert np.array_equal(P.to_host(it[2][1])[:n], o1[:n])


def test_batch_mixed_ring_dimensions(gpu):
    import torch
    from aoa_amd import Batch, Params
    rng = np.random.default_rng(32)
    B = Batch()
    checks = []
    # N <= 4096: tensor kernel + single-workgroup transforms; from 8192 up: the
    # tensor product fused into the two-pass inverse transform
    for N, k, prec in ((2048, 2, 55), (4096, 3, 52), (8192, 4, 48), (16384, 8, 48), (65536, 3, 48)):
        P = Params(N, k, 0, prec)
        evk = [P.to_device(_rand(P, rng, P.k + P.kl)) for _ in range(2)]
        items = _mul_items(P, rng, [k - 1, 1])
        checks.append((P, evk, items, [_plain_mul(P, evk, it) for it in items]))
        B.add_mul(P, evk, items)
    B.launch(graph=True)
    torch.cuda.synchronize()
    for P, evk, items, refs in checks:
        for it, r in zip(items, refs):
            lv = it[3]
            assert np.array_equal(P.to_host(it[2][0])[:lv * P.N], r[0])
            assert np.array_equal(P.to_host(it[2][1])[:lv * P.N], r[1])


def test_batch_rotate_and_sum_equals_plain(gpu):
    """Rotate-and-sum (ckks_sumslots) as layers of batched rotate-accumulate
    steps, against the plain per-ciphertext rotations plus additions."""
    import torch
    from aoa_amd import Batch, Params
    P = Params(16384, 6, 0, 48)
    rng = np.random.default_rng(33)
    level = P.k - 1
    steps = [1, 2, 4, 8]
    rtk = {r: [P.to_device(_rand(P, rng, P.k + P.kl)) for _ in range(2)] for r in steps}
    cts = [[P.to_device(_rand(P, rng, P.k)) for _ in range(2)] for _ in range(3)]
    # plain reference
    refs = []
    ws = P.workspace(level)
    for c in cts:
        cur = [t.clone() for t in c]
        for r in steps:
            rot = [P.zeros(P.k), P.zeros(P.k)]
            P.rotate_left(rot, cur, rtk[r], r, level, ws)
            nxt = [P.zeros(P.k), P.zeros(P.k)]
            for y in range(2):
                P.eltwise(0, nxt[y], cur[y], rot[y], res0=0, nres=level + 1)
            cur = nxt
        refs.append([P.to_host(t)[:(level + 1) * P.N] for t in cur])
    # batched: in place, one layer per step
    work = [[t.clone() for t in c] for c in cts]
    B = Batch()
    for i, r in enumerate(steps):
        if i:
            B.barrier()
        B.add_rotate(P, [(w, w, rtk[r], r, level, 1) for w in work])
    B.launch(graph=True)
    torch.cuda.synchronize()
    for w, ref in zip(work, refs):
        assert np.array_equal(P.to_host(w[0])[:(level + 1) * P.N], ref[0])
        assert np.array_equal(P.to_host(w[1])[:(level + 1) * P.N], ref[1])


def test_autotuner_splits_agree(gpu):
    """Every NTT split the autotuner may pick gives identical residues."""
    from aoa_amd import Params
    P = Params(16384, 4, 0, 48)
    rng = np.random.default_rng(34)
    x = _rand(P, rng, P.k + P.kl)
    outs = []
    for n1 in (64, 128, 256):
        P.set_ntt_split(n1)
        d = P.to_device(x)
        P.ntt(d, 0, P.k + P.kl, 0)
        outs.append(P.to_host(d))
        P.ntt(d, 0, P.k + P.kl, 1)
        assert np.array_equal(P.to_host(d), x)
    assert all(np.array_equal(o, outs[0]) for o in outs)
    times = P.autotune_ntt(reps=3)
    assert set(times) == {64, 128, 256}
    assert P.ntt_split() in times


def test_stage_times_runs_the_batch(gpu):
    """The stage-timing run is an ordinary eager run of the batch: outputs equal
    the plain form, every stage of HomMul is timed once per layer (from
    N = 8192 up the tensor product runs inside the inverse-NTT stage)."""
    import torch
    from aoa_amd import Batch, Params
    P = Params(16384, 6, 0, 48)
    rng = np.random.default_rng(35)
    evk = [P.to_device(_rand(P, rng, P.k + P.kl)) for _ in range(2)]
    items = _mul_items(P, rng, [5, 3, 1])
    ref = [_plain_mul(P, evk, it) for it in items]
    B = Batch()
    B.add_mul(P, evk, items)
    split = B.stage_times()
    torch.cuda.synchronize()
    assert set(split) == {"intt", "modup", "ntt", "key_intt", "moddown"}
    assert all(n == 1 and ms > 0 for ms, n in split.values())
    for it, r in zip(items, ref):
        lv = it[3]
        assert np.array_equal(P.to_host(it[2][0])[:lv * P.N], r[0])
        assert np.array_equal(P.to_host(it[2][1])[:lv * P.N], r[1])


def test_batch_c3_full_size_equals_plain(gpu):
    """The batched workload at its full size: 256 HomMul+relin+rescale at
    N=2^15, k=20, levels uniform in 1..19, one graph; every output equals the
    plain call's bit for bit (compared on the device)."""
    import torch
    from aoa_amd import Batch, Params
    P = Params(32768, 20, 0, 48)
    rng = np.random.default_rng(34)
    evk = [P.to_device(_rand(P, rng, P.k + P.kl)) for _ in range(2)]
    levels = [int(x) for x in rng.integers(1, 20, size=256)]
    print("c3 full size: 256 items", flush=True)
    items = _mul_items(P, rng, levels)
    B = Batch()
    B.add_mul(P, evk, items)
    B.build()
    B.launch(graph=True)
    torch.cuda.synchronize()
    ws = {}
    for it in items:
        c1, c2, out, lv = it
        ref = [P.zeros(P.k), P.zeros(P.k)]
        if lv not in ws:
            ws[lv] = P.workspace(lv)
        P.hommul(ref, c1, c2, evk, lv, ws[lv])
        n = lv * P.N
        assert torch.equal(out[0][:n], ref[0][:n]) and torch.equal(out[1][:n], ref[1][:n]), lv


def test_batch_c4_full_size_equals_plain(gpu):
    """Rotate-and-sum at its full size: 64 ciphertexts at N=2^16, k=24, 15
    layers of rotate-accumulate (steps 2^0..2^14) with resident rotation keys;
    every result equals the plain rotations plus additions bit for bit."""
    import torch
    from aoa_amd import Batch, Params
    P = Params(65536, 24, 0, 48)
    rng = np.random.default_rng(35)
    level = P.k - 1
    steps = [1 << i for i in range(15)]
    print("c4 full size: 64 ciphertexts x 15 rotations", flush=True)
    rtk = {r: [P.to_device(_rand(P, rng, P.k + P.kl)) for _ in range(2)] for r in steps}
    cts = [[P.to_device(_rand(P, rng, P.k)) for _ in range(2)] for _ in range(64)]
    work = [[t.clone() for t in c] for c in cts]
    B = Batch()
    for i, r in enumerate(steps):
        if i:
            B.barrier()
        B.add_rotate(P, [(w, w, rtk[r], r, level, 1) for w in work])
    B.launch(graph=True)
    torch.cuda.synchronize()
    ws = P.workspace(level)
    n = (level + 1) * P.N
    rot = [P.zeros(P.k), P.zeros(P.k)]
    for c, w in zip(cts, work):
        cur = [t.clone() for t in c]
        for r in steps:
            P.rotate_left(rot, cur, rtk[r], r, level, ws)
            for y in range(2):
                P.eltwise(0, cur[y], cur[y], rot[y], res0=0, nres=level + 1)
        assert torch.equal(w[0][:n], cur[0][:n]) and torch.equal(w[1][:n], cur[1][:n])


def test_batch_fused_every_level_equals_oracle(gpu):
    """The batched HomMul at the smallest two-pass size (N=2^13, k=8, prec 48),
    one item at every level 1..k-1: the tensor pass of d2, d0 and d1 folded
    into the key products and ModDown without addends give the oracle's
    residues for every item."""
    import torch
    from aoa_amd import Batch, Params
    N, k = 8192, 8
    P, O = Params(N, k, 0, 48), Oracle(N, k, 0, 48)
    rng = np.random.default_rng(43)
    evk = [P.to_device(_rand(P, rng, P.k + P.kl)) for _ in range(2)]
    levels = list(range(1, k))
    items = _mul_items(P, rng, levels)
    B = Batch()
    B.add_mul(P, evk, items)
    B.build()
    B.launch(graph=False)
    torch.cuda.synchronize()
    he = [P.to_host(t) for t in evk]
    for it in items:
        lv = it[3]
        h = [P.to_host(t) for t in it[0] + it[1]]
        o0, o1 = O.hommul((h[0], h[1]), (h[2], h[3]), (he[0], he[1]), lv)
        n = lv * N
        assert np.array_equal(P.to_host(it[2][0])[:n], o0[:n]), "c0 level %d" % lv
        assert np.array_equal(P.to_host(it[2][1])[:n], o1[:n]), "c1 level %d" % lv
