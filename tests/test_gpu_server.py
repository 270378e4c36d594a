"""Serving loop (continuous batching): every request's result is the same
alone and in the mix (mixed ring dimensions, staggered arrivals while other
requests run, admission limited by the HBM pool), and equals the plain
per-call evaluation of its op chain; oversize requests are rejected."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(P, rng, nlimbs):
    out = np.empty(nlimbs * P.N, dtype=np.uint64)
    for r in range(nlimbs):
        out[r * P.N:(r + 1) * P.N] = rng.integers(0, P.primes[r], size=P.N, dtype=np.uint64)
    return out


def _setup(rng, N, k, steps=(1, 2, 4)):
    from aoa_amd import Params
    P = Params(N, k, 0, 48)
    evk = [P.to_device(_rand(P, rng, P.k + P.kl)) for _ in range(2)]
    rtk = {r: [P.to_device(_rand(P, rng, P.k + P.kl)) for _ in range(2)] for r in steps}
    return P, evk, rtk


def _plain_chain(P, evk, rtk, c, m, level, ops):
    """The op chain with the plain C-ABI calls, input in the NTT domain."""
    cur = [t.clone() for t in c]
    ntt = True
    for op in ops:
        ws = P.workspace(level)
        out = [P.zeros(P.k), P.zeros(P.k)]
        if op[0] == "mul":
            if not ntt:
                for t in cur:
                    P.ntt(t, 0, level + 1, 0)
            P.hommul(out, cur, m, evk, level, ws)
            level -= 1
        else:
            if ntt:
                for t in cur:
                    P.ntt(t, 0, level + 1, 1)
            P.rotate_left(out, cur, rtk[op[1]], op[1], level, ws)
        cur, ntt = out, False
    if ntt:
        for t in cur:
            P.ntt(t, 0, level + 1, 1)
    return [P.to_host(t)[:(level + 1) * P.N] for t in cur], level


def _requests(rng, sets, n):
    reqs = []
    for i in range(n):
        s = i % len(sets)
        P, evk, rtk = sets[s]
        level = P.k - 1
        nmul = int(rng.integers(1, min(4, level) + 1))
        nrot = int(rng.integers(0, 3))
        ops = [("mul",)] * nmul + [("rot", int(rng.choice(list(rtk))))] * nrot
        rng.shuffle(ops)
        ops = [tuple(o) for o in ops]
        c = [P.to_device(_rand(P, rng, P.k)) for _ in range(2)]
        m = [P.to_device(_rand(P, rng, P.k)) for _ in range(2)]
        reqs.append((s, c, m, level, ops))
    return reqs


@pytest.mark.parametrize("shapes,nreq", [([(8192, 4), (16384, 8), (32768, 5)], 12),
                                         ([(65536, 30), (32768, 16), (8192, 4)], 8)])
def test_server_alone_vs_mix(gpu, shapes, nreq):
    """Small shapes, and the serving workload's largest ones (N=2^16 with 30
    limbs, N=2^15 with 16) mixed with a small one."""
    import torch
    from aoa_amd import Server
    rng = np.random.default_rng(41)
    print("server shapes %s" % (shapes,), flush=True)
    sets = [_setup(rng, N, k) for N, k in shapes]
    reqs = _requests(rng, sets, nreq)
    refs = [_plain_chain(sets[s][0], sets[s][1], sets[s][2], c, m, lv, ops) for s, c, m, lv, ops in reqs]

    def run(max_active, stagger, burst=False):
        srv = Server(pool_bytes=2 << 30, step_ws_bytes=2 << 30, max_active=max_active)
        ks = [srv.register(P, evk, rtk) for P, evk, rtk in sets]
        outs, ids = [], []
        if not burst:
            srv.start()
        for s, c, m, lv, ops in reqs:
            P = sets[s][0]
            out = [P.zeros(P.k), P.zeros(P.k)]
            outs.append(out)
            ids.append(srv.submit(ks[s], c, lv, True, m, ops, out))
            if stagger:
                time.sleep(0.0005)  # arrivals while earlier requests run
        if burst:
            srv.start()
        deadline = time.time() + 60
        while time.time() < deadline:
            if all(srv.result(i)[0] >= Server.DONE for i in ids):
                break
            time.sleep(0.002)
        srv.stop()
        torch.cuda.synchronize()
        res = [srv.result(i) for i in ids]
        stats = srv.stats()
        srv.close()
        return outs, res, stats

    alone, ra, _ = run(1, False)
    mixed, rm, stats = run(64, True)
    burst, rb, bstats = run(64, False, burst=True)
    assert stats["completed"] == len(reqs) and bstats["completed"] == len(reqs)
    # a burst of queued requests shares steps: far fewer steps than ops
    assert bstats["steps"] <= max(len(r[4]) for r in reqs) + 1
    for (s, c, m, lv, ops), ref, a, b, cc, sa, sb, sc in zip(reqs, refs, alone, mixed, burst, ra, rm, rb):
        P = sets[s][0]
        assert sa[0] == Server.DONE and sb[0] == Server.DONE and sc[0] == Server.DONE
        assert sa[1] == ref[1] and sb[1] == ref[1] and sc[1] == ref[1]
        n = (ref[1] + 1) * P.N
        for y in range(2):
            ga, gb, gc = P.to_host(a[y])[:n], P.to_host(b[y])[:n], P.to_host(cc[y])[:n]
            assert np.array_equal(ga, ref[0][y]), "alone differs from the plain chain"
            assert np.array_equal(gb, ref[0][y]), "staggered mix differs from the plain chain"
            assert np.array_equal(gc, ref[0][y]), "burst mix differs from the plain chain"


def test_server_admission_budget(gpu):
    import torch
    from aoa_amd import AoaError, Server
    rng = np.random.default_rng(42)
    P, evk, rtk = _setup(rng, 16384, 8)
    ct = 2 * P.k * P.N * 8
    # the pool holds two working ciphertext pairs: requests queue behind it
    srv = Server(pool_bytes=2 * (2 * ct + 1024), step_ws_bytes=1 << 28, max_active=64)
    ks = srv.register(P, evk, rtk)
    ids, outs = [], []
    for i in range(6):
        c = [P.to_device(_rand(P, rng, P.k)) for _ in range(2)]
        m = [P.to_device(_rand(P, rng, P.k)) for _ in range(2)]
        out = [P.zeros(P.k), P.zeros(P.k)]
        outs.append((c, m, out))
        ids.append(srv.submit(ks, c, P.k - 1, True, m, [("mul",), ("rot", 2)], out))
    n = 0
    while True:
        left = srv.step()
        n += 1
        assert n < 100
        if left == 0:
            break
    torch.cuda.synchronize()
    assert all(srv.result(i)[0] == Server.DONE for i in ids)
    assert srv.stats()["pool_used"] == 0
    ref, lv = _plain_chain(P, evk, rtk, outs[3][0], outs[3][1], P.k - 1, [("mul",), ("rot", 2)])
    nn = (lv + 1) * P.N
    assert np.array_equal(P.to_host(outs[3][2][0])[:nn], ref[0])
    # a released record is gone; releasing twice is an error
    srv.release(ids[0])
    with pytest.raises(AoaError):
        srv.result(ids[0])
    with pytest.raises(AoaError):
        srv.release(ids[0])
    assert srv.result(ids[1])[0] == Server.DONE
    # a request larger than the whole pool is refused
    big = Server(pool_bytes=1 << 20, step_ws_bytes=1 << 28, max_active=4)
    kb = big.register(P, evk, rtk)
    with pytest.raises(AoaError):
        big.submit(kb, outs[0][0], P.k - 1, True, outs[0][1], [("mul",)], outs[0][2])
