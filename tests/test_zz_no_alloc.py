"""aoa_hommul_relin_rescale_noalloc (static device workspace, no allocation)
against aoa_hommul_relin_rescale (caller workspace), bit for bit, through the
C-ABI: two calls in a row on one device, the second on a smaller input (a
smaller ring and fewer limbs), so that the second call runs over the first
call's leftovers in the shared scratch array."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _sampled(P, nlimbs, nonce, seed=91):
    from aoa_amd import SAMPLE_UNIFORM
    t = P.zeros(nlimbs)
    P.sample(t, SAMPLE_UNIFORM, seed, nonce, res0=0, nres=nlimbs)
    return t


def _case(P, level, nonce):
    k, R = P.k, P.k + P.kl
    c1 = [_sampled(P, k, nonce), _sampled(P, k, nonce + 1)]
    c2 = [_sampled(P, k, nonce + 2), _sampled(P, k, nonce + 3)]
    evk = [_sampled(P, R, nonce + 4), _sampled(P, R, nonce + 5)]
    return c1, c2, evk


def test_noalloc_equals_workspace_entry(gpu):
    from aoa_amd import Params
    import aoa_amd
    lib = aoa_amd.load()
    cap = lib.aoa_noalloc_scratch_bytes()
    assert cap == 144 << 20
    # (N, k, kl, prec, level): the largest shape first (136.5 MiB of workspace at
    # the top level), then a smaller ring at a lower level
    shapes = [(65536, 30, 0, 48, 29), (16384, 8, 0, 48, 3)]
    for n, (N, k, kl, prec, level) in enumerate(shapes):
        print("noalloc call %d: N=%d k=%d level %d" % (n, N, k, level), flush=True)
        P = Params(N, k, kl, prec)
        assert P.lib.aoa_workspace_bytes(P.h, level) <= cap
        c1, c2, evk = _case(P, level, 10 * (n + 1))
        want = [P.zeros(k), P.zeros(k)]
        P.hommul(want, c1, c2, evk, level, P.workspace(level))
        got = [torch.full_like(want[0], -1), torch.full_like(want[1], -1)]
        P.hommul_noalloc(got, c1, c2, evk, level)
        torch.cuda.synchronize()
        m = level * N
        for y in range(2):
            assert torch.equal(got[y][:m], want[y][:m]), "call %d output %d differs" % (n, y)
        # a repeat on the same device gives the same residues
        again = [P.zeros(k), P.zeros(k)]
        P.hommul_noalloc(again, c1, c2, evk, level)
        torch.cuda.synchronize()
        assert torch.equal(again[0][:m], want[0][:m]) and torch.equal(again[1][:m], want[1][:m])
        del P


def test_noalloc_rejects_shapes_beyond_the_scratch(gpu):
    """A shape whose workspace exceeds the static scratch fails on the host,
    before any launch, with AOA_ERR_WORKSPACE."""
    from aoa_amd import AoaError, Params
    P = Params(65536, 30, 40, 48)  # K = 40: (180 + 120) limbs of 512 KiB = 150 MiB
    level = 29
    assert P.lib.aoa_workspace_bytes(P.h, level) > P.lib.aoa_noalloc_scratch_bytes()
    k, R = P.k, P.k + P.kl
    z = [P.zeros(k) for _ in range(6)]
    evk = [P.zeros(R), P.zeros(R)]
    with pytest.raises(AoaError) as e:
        P.hommul_noalloc((z[0], z[1]), (z[2], z[3]), (z[4], z[5]), evk, level)
    assert e.value.code == 5 and "static scratch" in str(e.value)
    # a lower level of the same parameters fits and works
    lv = 3
    assert P.lib.aoa_workspace_bytes(P.h, lv) <= P.lib.aoa_noalloc_scratch_bytes()
    c1, c2, ev = _case(P, lv, 300)
    want = [P.zeros(k), P.zeros(k)]
    P.hommul(want, c1, c2, ev, lv, P.workspace(lv))
    got = [P.zeros(k), P.zeros(k)]
    P.hommul_noalloc(got, c1, c2, ev, lv)
    torch.cuda.synchronize()
    m = lv * P.N
    assert torch.equal(got[0][:m], want[0][:m]) and torch.equal(got[1][:m], want[1][:m])
