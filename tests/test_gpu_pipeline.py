"""End to end on the device (aoa_amd.pipeline: keys, encode, encrypt,
HomMul+relin+rescale, decrypt, decode) against the oracle's restatement of the
same sequence: keys, ciphertext residues and decoded slot values are equal bit
for bit, and the product decrypts within the reference's bound (0.005,
aoa-ntt/tests/tests.cpp:34)."""
import numpy as np
import pytest

from orc import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,k,kl,prec,slots,exact", [(4096, 4, 5, 52, 2048, 0), (2048, 3, 3, 55, 1024, 0),
                                                     (8192, 5, 6, 48, 512, 1)])
def test_pipeline_bitexact(gpu, N, k, kl, prec, slots, exact):
    from aoa_amd import Params
    from aoa_amd.pipeline import Pipeline
    P, O = Params(N, k, kl, prec), Oracle(N, k, kl, prec)
    P.set_moddown_exact(exact)
    O.set_moddown_exact(exact)
    seed = 42
    print("pipeline N=%d k=%d prec=%d slots=%d exact=%d" % (N, k, prec, slots, exact), flush=True)
    pipe = Pipeline(P, seed, slots)
    keys = O.c1_keys(seed)
    for name in ("s", "pkb", "pka", "evkb", "evka"):
        assert np.array_equal(P.to_host(getattr(pipe, name)), keys[name]), name
    rng = np.random.default_rng(N + k)
    z1 = rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots)
    z2 = rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots)
    scale = 1 << prec
    pipe.set_inputs(z1, z2, scale)
    for nonce in (5, 11):
        pipe.run(nonce)
        got = pipe.result()
        ref, c3 = O.c1_run(keys, z1, z2, scale, seed, nonce)
        # the device's c3 is in the NTT domain after decryption's transforms
        L = k - 1
        for y in range(2):
            want = c3[y * k * N:(y + 1) * k * N].copy()
            O.ntt(want, 0, L, 0)
            assert np.array_equal(P.to_host(pipe.c3[y])[:L * N], want[:L * N]), (nonce, y)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), nonce
        assert np.max(np.abs(got - z1 * z2)) < 0.005


def test_pipeline_graph_replays(gpu):
    """The pipeline captured as one HIP graph: every replay, with eager work
    and synchronisations between replays, equals the eager result bit for bit."""
    import torch
    from aoa_amd import Params
    from aoa_amd.pipeline import Pipeline
    P = Params(4096, 4, 5, 52)
    pipe = Pipeline(P, 42, 2048)
    rng = np.random.default_rng(3)
    z1 = rng.uniform(-1, 1, 2048) + 1j * rng.uniform(-1, 1, 2048)
    z2 = rng.uniform(-1, 1, 2048) + 1j * rng.uniform(-1, 1, 2048)
    pipe.set_inputs(z1, z2, 1 << 52)
    pipe.run(5)
    eager = pipe.result()
    c3 = [P.to_host(t) for t in pipe.c3]
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        pipe.run(5)
    torch.cuda.current_stream().wait_stream(side)
    with torch.cuda.graph(g, capture_error_mode="thread_local"):
        pipe.run(5)
    for rep in range(3):
        pipe.out.zero_()
        for t in pipe.c3:
            t.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(pipe.result().view(np.uint64), eager.view(np.uint64)), rep
        assert all(np.array_equal(P.to_host(t), c) for t, c in zip(pipe.c3, c3)), rep
        pipe.run(7)  # eager work between replays
        torch.cuda.synchronize()


@pytest.mark.parametrize("N,k,kl,prec", [(4096, 4, 5, 52), (16384, 5, 5, 55)])
def test_encrypt_decrypt_entry_points_match_primitives(gpu, N, k, kl, prec):
    """aoa_encrypt and aoa_decrypt equal the primitive calls they replace
    (aoa_sample, aoa_ntt, aoa_eltwise), bit for bit, with the single-workgroup
    transform (N = 4096) and the two-pass one."""
    import torch
    from aoa_amd import Params
    from aoa_amd.core import ADD, FORWARD, INVERSE, MULADD, SAMPLE_GAUSSIAN, SAMPLE_ZO
    P, O = Params(N, k, kl, prec), Oracle(N, k, kl, prec)
    rng = np.random.default_rng(N)
    seed, nonce = 7, 123
    m = P.to_device(O.random_poly(rng, k))
    pkb, pka = P.to_device(O.random_poly(rng, k)), P.to_device(O.random_poly(rng, k))
    # primitives
    m1 = m.clone()
    u, e0, e1 = P.zeros(k), P.zeros(k), P.zeros(k)
    P.ntt(m1, 0, k, FORWARD)
    P.sample(u, SAMPLE_ZO, seed, nonce, 0, k)
    P.sample(e0, SAMPLE_GAUSSIAN, seed, nonce + 1, 0, k)
    P.sample(e1, SAMPLE_GAUSSIAN, seed, nonce + 2, 0, k)
    for x in (u, e0, e1):
        P.ntt(x, 0, k, FORWARD)
    r0, r1 = P.zeros(k), P.zeros(k)
    P.eltwise(MULADD, r0, pkb, u, e0, res0=0, nres=k)
    P.eltwise(ADD, r0, r0, m1, res0=0, nres=k)
    P.eltwise(MULADD, r1, pka, u, e1, res0=0, nres=k)
    # entry point
    m2 = m.clone()
    c0, c1, ws = P.zeros(k), P.zeros(k), P.zeros(3 * k)
    P.encrypt(c0, c1, m2, pkb, pka, seed, nonce, ws)
    assert torch.equal(m1, m2) and torch.equal(r0, c0) and torch.equal(r1, c1)
    # decryption at level k-2 with a random NTT-domain key
    s = P.to_device(O.random_poly(rng, k))
    level = k - 2
    L = level + 1
    a0, a1 = P.to_device(O.random_poly(rng, k)), P.to_device(O.random_poly(rng, k))
    b0, b1, t = a0.clone(), a1.clone(), P.zeros(k)
    P.ntt(b0, 0, L, FORWARD)
    P.ntt(b1, 0, L, FORWARD)
    P.eltwise(MULADD, t, b1, s, b0, res0=0, nres=L)
    P.ntt(t, 0, L, INVERSE)
    got = P.zeros(k)
    P.decrypt(got, a0, a1, s, level)
    n = L * N
    assert torch.equal(got[:n], t[:n])
    assert torch.equal(a0[:n], b0[:n]) and torch.equal(a1[:n], b1[:n])
