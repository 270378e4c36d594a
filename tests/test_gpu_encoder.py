"""GPU encoder (aoa_special_fft, aoa_encode, aoa_decode) against the oracle's
restatement of the reference's fftSpecial / fftSpecialInv / ckks_encode /
ckks_decode (aoa-ntt/src/ckks/encoder.cu:33-298), bit for bit: the FP64
values are rounded in the same operation order and the residues are integers.
The one intended difference: for a negative value whose magnitude is a
multiple of q_r the reference writes q_r and the device 0 (the same residue)."""
import ctypes

import numpy as np
import pytest

import native_chk
import orc
from orc import Oracle

pytestmark = pytest.mark.gpu


def _dev_c(vals):
    import torch
    return torch.from_numpy(np.ascontiguousarray(vals, dtype=np.complex128).view(np.float64).copy()).cuda()


def _host_c(t):
    return t.cpu().numpy().view(np.complex128).copy()


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.complex128).view(np.uint64)


def _oracle_fft(vals, ringN, inverse):
    buf = np.ascontiguousarray(vals, dtype=np.complex128).view(np.float64).copy()
    orc.load().orc_fft_special(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(vals), ringN, int(inverse))
    return buf.view(np.complex128)


def _prec(N):
    return 55 if N == 2048 else 48


@pytest.mark.parametrize("N,n", [(2048, 1), (2048, 2), (2048, 1024), (4096, 256), (65536, 64), (65536, 32768)])
def test_special_fft_bitexact(gpu, N, n):
    from aoa_amd import FORWARD, INVERSE, Params
    P = Params(N, 2, 0, _prec(N))
    rng = np.random.default_rng(N + n)
    u = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    for direction, inv in ((FORWARD, 0), (INVERSE, 1)):
        print("special_fft N=%d n=%d dir=%d" % (N, n, direction), flush=True)
        t = _dev_c(u)
        P.special_fft(t, n, direction)
        assert np.array_equal(_bits(_host_c(t)), _bits(_oracle_fft(u, N, inv))), "dir %d" % direction


def test_encoder_tables(gpu):
    from aoa_amd import Params
    for N in (2048, 65536):
        P = Params(N, 2, 3, _prec(N))
        for which in (11, 12):
            assert np.array_equal(P.table(which), native_chk.table(N, 2, 3, _prec(N), which))


@pytest.mark.parametrize("N,k,slots,amp", [(2048, 4, 1, 1.0), (2048, 4, 8, 1.0), (2048, 4, 1024, 100.0),
                                           (65536, 3, 32768, 1.0), (65536, 30, 4096, 1000.0)])
def test_encode_decode_bitexact(gpu, N, k, slots, amp):
    from aoa_amd import Params
    prec = _prec(N)
    P, O = Params(N, k, 0, prec), Oracle(N, k, 0, prec)
    scale = 1 << prec
    rng = np.random.default_rng(slots + k)
    z = amp * (rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots))
    print("encode N=%d k=%d slots=%d" % (N, k, slots), flush=True)
    vals = _dev_c(z)
    out = P.zeros(k)
    P.encode(out, vals, slots, scale)
    ref = O.encode(z, scale)
    # the in-place inverse FFT leaves the same values as the reference's
    assert np.array_equal(_bits(_host_c(vals)), _bits(_oracle_fft(z, N, 1)))
    got = P.to_host(out).reshape(k, N)
    ref = ref.reshape(k, N)
    for l in range(k):
        assert np.array_equal(got[l], ref[l] % np.uint64(O.primes[l])), "limb %d" % l
    print("decode N=%d slots=%d" % (N, slots), flush=True)
    dv = _dev_c(np.zeros(slots, dtype=np.complex128))
    P.decode(dv, out, slots, scale)
    back = _host_c(dv)
    assert np.array_equal(_bits(back), _bits(O.decode(ref[0], slots, scale)))
    tol = amp * 4.0 * slots * (1 + O.primes[0] * 2.0 ** -53) / scale + 1e-9
    assert np.max(np.abs(back - z)) < tol


def test_decode_random_residues_bitexact(gpu):
    from aoa_amd import Params
    N, k = 65536, 2
    P, O = Params(N, k, 0, 48), Oracle(N, k, 0, 48)
    rng = np.random.default_rng(99)
    limb0 = O.random_poly(rng, 1)
    d = P.to_device(np.concatenate([limb0, np.zeros(N, dtype=np.uint64)]))
    for slots in (1, 512, 32768):
        dv = _dev_c(np.zeros(slots, dtype=np.complex128))
        P.decode(dv, d, slots, 1 << 48)
        assert np.array_equal(_bits(_host_c(dv)), _bits(O.decode(limb0, slots, 1 << 48))), slots


def test_encode_negative_multiple(gpu):
    """A slot value of -q_1 at scale 1: the reference writes q_1 on limb 1, the device 0."""
    from aoa_amd import Params
    N, k = 2048, 3
    P, O = Params(N, k, 0, 48), Oracle(N, k, 0, 48)
    z = np.array([-float(O.primes[1]) + 0j])
    out = P.zeros(k)
    P.encode(out, _dev_c(z), 1, 1)
    got = P.to_host(out).reshape(k, N)
    ref = O.encode(z, 1).reshape(k, N)
    assert int(ref[1, 0]) == O.primes[1] and int(got[1, 0]) == 0
    assert int(got[0, 0]) == int(ref[0, 0]) and int(got[2, 0]) == int(ref[2, 0])


def test_encoder_argument_errors(gpu):
    from aoa_amd import Params
    from aoa_amd._lib import AoaError
    P = Params(2048, 2, 0, 55)
    v = _dev_c(np.zeros(4, dtype=np.complex128))
    out = P.zeros(2)
    for bad in (3, 0, 2048):
        with pytest.raises(AoaError):
            P.encode(out, v, bad, 1 << 40)
    with pytest.raises(AoaError):
        P.encode(out, v, 4, 0)
    with pytest.raises(AoaError):
        P.encode(out, v, 4, 1 << 40, nres=3)
