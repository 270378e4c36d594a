"""The batch autotuner (aoa_batch_autotune): every launch-option candidate
(NTT split, descriptor order, XCD chunks, strict-split rule) gives the same
residues -- each candidate's output checksum equals the first's -- the chosen
options give the plain form's residues bit for bit, and a table written by
aoa_amd.tuning is applied back to a new batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand(P, rng, nlimbs):
    out = np.empty(nlimbs * P.N, dtype=np.uint64)
    for r in range(nlimbs):
        out[r * P.N:(r + 1) * P.N] = rng.integers(0, P.primes[r], size=P.N, dtype=np.uint64)
    return out


def _mul_batch(P, rng, levels, evk):
    from aoa_amd import Batch
    items = []
    for lv in levels:
        c1 = [P.to_device(_rand(P, rng, P.k)) for _ in range(2)]
        c2 = [P.to_device(_rand(P, rng, P.k)) for _ in range(2)]
        items.append((c1, c2, [P.zeros(P.k), P.zeros(P.k)], lv))
    B = Batch()
    B.add_mul(P, evk, items)
    B.build()
    return B, items


def _check_plain(P, evk, items):
    import torch
    for c1, c2, out, lv in items:
        ref = [P.zeros(P.k), P.zeros(P.k)]
        P.hommul(ref, c1, c2, evk, lv, P.workspace(lv))
        n = lv * P.N
        assert torch.equal(out[0][:n], ref[0][:n]) and torch.equal(out[1][:n], ref[1][:n]), "level %d" % lv


@pytest.mark.parametrize("shape", [(16384, 8, 48, 24), (32768, 12, 48, 16), (65536, 6, 48, 4)])
def test_every_candidate_is_exact(gpu, shape):
    """Mixed-level HomMul batches; at N=2^14 and 2^15 several NTT splits are
    candidates, at N=2^16 one."""
    import torch
    from aoa_amd import Params
    N, k, prec, nitems = shape
    print("autotune: N=%d k=%d items=%d" % (N, k, nitems), flush=True)
    P = Params(N, k, 0, prec)
    rng = np.random.default_rng(N + k)
    evk = [P.to_device(_rand(P, rng, P.k + P.kl)) for _ in range(2)]
    levels = [int(x) for x in rng.integers(1, k, size=nitems)]
    B, items = _mul_batch(P, rng, levels, evk)
    cands = B.autotune(reps=2)
    splits = {c["ntt_split"] for c in cands}
    assert len(cands) == 8 * len(splits)
    if N < 65536:
        assert len(splits) >= 2, splits
    assert all(c["exact"] for c in cands), [c for c in cands if not c["exact"]]
    assert len({c["checksum"] for c in cands}) == 1
    # the chosen options replayed: the plain form's residues
    for it in items:
        it[2][0].zero_()
        it[2][1].zero_()
    B.launch(graph=True)
    torch.cuda.synchronize()
    _check_plain(P, evk, items)
    B.close()


def test_rotation_batch_candidates_exact(gpu):
    """Rotations into separate outputs (reproducible between replays)."""
    import torch
    from aoa_amd import Batch, Params
    P = Params(16384, 6, 0, 48)
    rng = np.random.default_rng(5)
    R = P.k + P.kl
    rtk = [P.to_device(_rand(P, rng, R)) for _ in range(2)]
    items = []
    for i in range(8):
        inp = [P.to_device(_rand(P, rng, P.k)) for _ in range(2)]
        items.append((inp, [P.zeros(P.k), P.zeros(P.k)], rtk, 1 << i, 4, False))
    B = Batch()
    B.add_rotate(P, items)
    B.build()
    cands = B.autotune(reps=2)
    assert all(c["exact"] for c in cands)
    B.launch(graph=True)
    torch.cuda.synchronize()
    ws = P.workspace(4)
    for inp, out, key, rot, lv, _ in items:
        ref = [P.zeros(P.k), P.zeros(P.k)]
        P.rotate_left(ref, inp, key, rot, lv, ws)
        n = (lv + 1) * P.N
        assert torch.equal(out[0][:n], ref[0][:n]) and torch.equal(out[1][:n], ref[1][:n]), rot


def test_table_round_trip(gpu, tmp_path):
    from aoa_amd import Params, tuning
    P = Params(16384, 4, 0, 48)
    rng = np.random.default_rng(6)
    evk = [P.to_device(_rand(P, rng, P.k + P.kl)) for _ in range(2)]
    B, _ = _mul_batch(P, rng, [3, 2, 1, 3], evk)
    path = str(tmp_path / "table.json")
    k = tuning.key(P.N, P.k + P.kl, 4, "mul")
    chosen, cands = tuning.tune(B, k, reps=2, path=path)
    B2, _ = _mul_batch(P, rng, [3, 3, 2, 1], evk)
    assert tuning.apply(B2, k, tuning.load_table(path)) == chosen
    assert B2.options() == chosen
    assert tuning.apply(B2, "no such key", tuning.load_table(path)) is None
