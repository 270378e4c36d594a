"""Known-answer vectors (tests/golden/n2048_prec55.npz, written by
tests/golden/gen_golden.py from the definitions with Python integers) against
the oracle and the product's host tables (CPU), and against the HIP kernels
through the C-ABI (GPU).  N=2048, k=3, kl=3, prec 55; outputs are checked at
every 16th coefficient."""
import os
import sys

import numpy as np
import pytest

import native_chk
from orc import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import gen_golden  # noqa: E402

N, K, KL, PREC = 2048, 3, 3, 55
S = gen_golden.SAMPLE


@pytest.fixture(scope="module")
def gold():
    with np.load(gen_golden.FIXTURE, allow_pickle=False) as f:
        g = {k: f[k] for k in f.files}
    g.update(gen_golden.inputs())
    return g


def test_tables(gold):
    o = Oracle(N, K, KL, PREC)
    assert o.primes == [int(v) for v in gold["primes"]]
    assert o.roots == [int(v) for v in gold["roots"]]
    assert [int(v) for v in native_chk.table(N, K, KL, PREC, 10)] == o.primes
    assert np.array_equal(native_chk.table(N, K, KL, PREC, 9), gold["mu"])
    # psi^brev(1) = psi^(N/2): the product's forward twiddle table, entry 1 of each prime
    tw = native_chk.table(N, K, KL, PREC, 0).reshape(K + KL, N)
    for r, (p, psi) in enumerate(zip(o.primes, gold["psi"])):
        assert int(tw[r, 1]) == pow(int(psi), N // 2, p)


def test_oracle_ntt(gold):
    o = Oracle(N, K, KL, PREC)
    x = np.zeros((K + KL) * N, dtype=np.uint64)
    x[:2 * N] = gold["ntt_in"].reshape(-1)
    o.ntt(x, 0, 2, 0)
    assert np.array_equal(x[:2 * N].reshape(2, N)[:, S], gold["ntt_out"])


def test_oracle_rescale_modup_moddown_automorphism(gold):
    o = Oracle(N, K, KL, PREC)
    a = gold["rescale_in"][0].reshape(-1).copy()
    b = gold["rescale_in"][1].reshape(-1).copy()
    o.rescale(a, b, K - 1)
    assert np.array_equal(a.reshape(K, N)[:K - 1, S], gold["rescale_out"][0])
    assert np.array_equal(b.reshape(K, N)[:K - 1, S], gold["rescale_out"][1])
    qb = np.zeros((K + KL) * N, dtype=np.uint64)
    qb[:K * N] = gold["modup_in"].reshape(-1)
    o.modup(qb, K - 1)
    assert np.array_equal(qb.reshape(K + KL, N)[K:, S], gold["modup_out"])
    out = np.zeros(K * N, dtype=np.uint64)
    o.moddown(out, gold["moddown_in"].reshape(-1).copy(), K - 1)
    assert np.array_equal(out.reshape(K, N)[:, S], gold["moddown_out"])
    y = np.zeros(K * N, dtype=np.uint64)
    y[N:2 * N] = gold["auto_in"]
    z = np.zeros_like(y)
    o.automorphism(z, y, 2, 1)
    assert np.array_equal(z[N:2 * N][S], gold["auto_out"])


@pytest.mark.gpu
def test_gpu_known_answers(gpu, gold):
    from aoa_amd import Params
    P = Params(N, K, KL, PREC)
    x = np.zeros((K + KL) * N, dtype=np.uint64)
    x[:2 * N] = gold["ntt_in"].reshape(-1)
    d = P.to_device(x)
    P.ntt(d, 0, 2, 0)
    assert np.array_equal(P.to_host(d)[:2 * N].reshape(2, N)[:, S], gold["ntt_out"])
    a, b = P.to_device(gold["rescale_in"][0].reshape(-1)), P.to_device(gold["rescale_in"][1].reshape(-1))
    P.rescale(a, b, K - 1)
    assert np.array_equal(P.to_host(a).reshape(K, N)[:K - 1, S], gold["rescale_out"][0])
    assert np.array_equal(P.to_host(b).reshape(K, N)[:K - 1, S], gold["rescale_out"][1])
    qb = np.zeros((K + KL) * N, dtype=np.uint64)
    qb[:K * N] = gold["modup_in"].reshape(-1)
    dq = P.to_device(qb)
    P.modup(dq, K - 1)
    assert np.array_equal(P.to_host(dq).reshape(K + KL, N)[K:, S], gold["modup_out"])
    out = P.zeros(K)
    P.moddown(out, P.to_device(gold["moddown_in"].reshape(-1)), K - 1)
    assert np.array_equal(P.to_host(out).reshape(K, N)[:, S], gold["moddown_out"])
    y = np.zeros(K * N, dtype=np.uint64)
    y[N:2 * N] = gold["auto_in"]
    z = P.zeros(K)
    P.automorphism(z, P.to_device(y), 2, 1)
    assert np.array_equal(P.to_host(z)[N:2 * N][S], gold["auto_out"])
