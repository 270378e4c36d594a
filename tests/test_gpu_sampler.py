"""Device sampler (aoa_sample) against the oracle's restatement, bit for bit,
for every distribution, at N=2048 and N=2^16 with all QB limbs, on full and
partial limb ranges."""
import numpy as np
import pytest

from orc import Oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,k,kl,prec", [(2048, 3, 4, 55), (65536, 30, 31, 48)])
@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4])
def test_sample_parity(gpu, N, k, kl, prec, kind):
    from aoa_amd import Params
    P, O = Params(N, k, kl, prec), Oracle(N, k, kl, prec)
    R = k + kl
    for seed, nonce in ((42, 0), (2 ** 63 + 5, 2 ** 40 + 3)):
        print("sample N=%d kind=%d seed=%d nonce=%d" % (N, kind, seed, nonce), flush=True)
        d = P.zeros(R)
        P.sample(d, kind, seed, nonce)
        assert np.array_equal(P.to_host(d), O.sample(kind, seed, nonce)), (seed, nonce)
    # a partial range touches only its limbs
    d = P.zeros(R)
    P.sample(d, kind, 7, 1, res0=1, nres=2)
    got = P.to_host(d).reshape(R, N)
    ref = O.sample(kind, 7, 1, res0=1, nres=2).reshape(R, N)
    assert np.array_equal(got, ref)
    assert not got[0].any() and not got[3:].any()


def test_sample_argument_errors(gpu):
    from aoa_amd import Params
    from aoa_amd._lib import AoaError
    P = Params(2048, 2, 0, 55)
    d = P.zeros(5)
    with pytest.raises(AoaError):
        P.sample(d, 5, 1, 1)
    with pytest.raises(AoaError):
        P.sample(d, 0, 1, 1, res0=4, nres=2)
