"""Encoder (FP64 FFT and the conversions between slot values and residues).

The oracle restates the reference's fftSpecial / fftSpecialInv loops and
ckks_encode / ckks_decode (aoa-ntt/src/ckks/encoder.cu:33-298).  Here it is
pinned against mathematics: the FFT against the definition
z_j = sum_s u_s w^(s 5^j), w = exp(2 pi i / 4n); encoding against the canonical
embedding (the residues, read as an integer polynomial m, satisfy
m(exp(i pi 5^j / N)) = scale * z_j), against agreement of the signed value
across limbs, and decode o encode = id.  The product's encoder tables are
checked against their definitions.  The GPU encoder is compared with the
oracle bit for bit in test_gpu_encoder.py."""
import ctypes
import math

import numpy as np
import pytest

import native_chk
import orc

TOL = 1e-9


def oracle_fft(vals, ringN, inverse):
    buf = np.ascontiguousarray(np.asarray(vals, dtype=np.complex128)).view(np.float64).copy()
    orc.load().orc_fft_special(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(vals), ringN, int(inverse))
    return buf.view(np.complex128)


def definition(u):
    n = len(u)
    M = 4 * n
    pts = [pow(5, j, M) for j in range(n)]
    s = np.arange(n)
    return np.array([np.sum(u * np.exp(2j * np.pi * s * t / M)) for t in pts])


@pytest.mark.parametrize("n,ringN", [(1, 4096), (8, 4096), (32, 2048), (128, 65536)])
def test_oracle_fft_definition_and_roundtrip(n, ringN):
    rng = np.random.default_rng(7 * n)
    u = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    assert np.max(np.abs(oracle_fft(u, ringN, False) - definition(u))) < TOL * n
    assert np.max(np.abs(oracle_fft(oracle_fft(u, ringN, True), ringN, False) - u)) < TOL
    assert np.max(np.abs(oracle_fft(oracle_fft(u, ringN, False), ringN, True) - u)) < TOL


def test_oracle_fft_ring_independent():
    """The twiddles depend on the slot count only: the same values for every ring dimension."""
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, 64) + 1j * rng.uniform(-1, 1, 64)
    for inv in (False, True):
        a = oracle_fft(u, 2048, inv)
        for ringN in (4096, 65536):
            assert np.array_equal(a, oracle_fft(u, ringN, inv))


def test_embedding_is_multiplicative():
    """Slots multiply when plaintext polynomials multiply negacyclically: the
    packing (real parts at s*gap, imaginary at N/2 + s*gap) of the inverse
    FFT's output is the polynomial whose evaluations at exp(i pi 5^j / N) are
    the slots."""
    N, n = 64, 8
    gap = (N // 2) // n
    rng = np.random.default_rng(3)

    def poly_of(z):
        u = oracle_fft(z, N, True)
        m = np.zeros(N)
        for s in range(n):
            m[s * gap] = u[s].real
            m[s * gap + N // 2] = u[s].imag
        return m

    def slots_of(m):
        roots = [np.exp(1j * np.pi * pow(5, j, 2 * N) / N) for j in range(n)]
        return np.array([np.sum(m * r ** np.arange(N)) for r in roots])

    z1 = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    z2 = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    m1, m2 = poly_of(z1), poly_of(z2)
    assert np.max(np.abs(slots_of(m1) - z1)) < 1e-9
    prod = np.zeros(N)
    for i in range(N):
        for j in range(N):
            if i + j < N:
                prod[i + j] += m1[i] * m2[j]
            else:
                prod[i + j - N] -= m1[i] * m2[j]
    assert np.max(np.abs(slots_of(prod) - z1 * z2)) < 1e-9


def _signed(res, p):
    r = res.astype(object)
    return np.array([int(x) - p if int(x) > p // 2 else int(x) for x in r], dtype=object)


@pytest.mark.parametrize("slots", [1, 4, 256, 1024])
def test_oracle_encode_embedding(slots):
    N, k, prec = 2048, 4, 55
    o = orc.Oracle(N, k, 0, prec)
    scale = 1 << 40
    rng = np.random.default_rng(slots)
    z = rng.uniform(-1, 1, slots) + 1j * rng.uniform(-1, 1, slots)
    out = o.encode(z, scale).reshape(k, N)
    m0 = _signed(out[0], o.primes[0])
    for l in range(1, k):  # the same signed integer on every limb
        assert list(_signed(out[l], o.primes[l])) == list(m0)
    gap = (N // 2) // slots
    nz = np.nonzero(np.array([int(x) != 0 for x in m0]))[0]
    assert all(i % gap == 0 for i in nz)
    # truncation to integers moves each of the 2*slots coefficients by < 1/scale;
    # decoding a negative residue rounds (double)x - (double)p with q_0 ~ 2^62
    # (ADJUSTSCALE, encoder.cu:129), up to q_0 2^-53 / scale per coefficient
    tol = 4.0 * slots * (1 + o.primes[0] * 2.0 ** -53) / scale + 1e-10
    m = np.array([float(x) for x in m0]) / scale
    for j in range(min(slots, 16)):
        root = np.exp(1j * np.pi * pow(5, j, 2 * N) / N)
        val = np.sum(m * root ** np.arange(N))
        assert abs(val - z[j]) < tol
    back = o.decode(out[0], slots, scale)
    assert np.max(np.abs(back - z)) < tol


def test_oracle_encode_negative_multiple_quirk():
    """ADJUSTNEG (encoder.cu:128) writes p, not 0, for a negative multiple of p."""
    N, k, prec = 2048, 3, 48
    o = orc.Oracle(N, k, 0, prec)
    p1 = o.primes[1]
    assert p1 < 2 ** 53
    out = o.encode(np.array([-float(p1) + 0j]), 1).reshape(k, N)
    assert int(out[1, 0]) == p1
    assert int(out[0, 0]) == o.primes[0] - p1
    assert int(out[2, 0]) == (o.primes[2] - p1 % o.primes[2]) % o.primes[2]


def test_oracle_decode_centering():
    """Residues up to p/2 decode as x / scale, larger ones as negative values
    (x - p) / scale, the subtraction done after both are rounded to double as
    the reference does: within one ulp of q_0 (2^10 at 63 bits) of the true value."""
    N, k, prec = 2048, 2, 55
    o = orc.Oracle(N, k, 0, prec)
    p = o.primes[0]
    a = np.zeros(N, dtype=np.uint64)
    a[0], a[N // 2] = p - 3, 5
    v = o.decode(a, 1, 1)
    assert v[0].imag == 5.0
    assert abs(v[0].real + 3) <= 2.0 ** 10
    a[0], a[N // 2] = p >> 1, (p >> 1) + 1
    v = o.decode(a, 1, 1)
    assert v[0].real > 0 and v[0].imag < 0
    assert abs(v[0].real - (p >> 1)) <= 2.0 ** 10 and abs(v[0].imag + (p >> 1)) <= 2.0 ** 10


@pytest.mark.parametrize("N", [2048, 65536])
def test_encoder_tables(N):
    """Roots exp(2 pi i t / 2N): close to the definition, and bit for bit glibc's
    sincos of 2 pi t / 2N (what the reference's optimised host code calls)."""
    w = native_chk.table(N, 2, 3, 55 if N == 2048 else 48, 11).view(np.float64).reshape(2 * N, 2)
    t = np.arange(2 * N)
    assert np.max(np.abs(w[:, 0] - np.cos(2 * np.pi * t / (2 * N)))) < 1e-15
    assert np.max(np.abs(w[:, 1] - np.sin(2 * np.pi * t / (2 * N)))) < 1e-15
    libm = ctypes.CDLL("libm.so.6")
    libm.sincos.restype = None
    libm.sincos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    s, c = ctypes.c_double(), ctypes.c_double()
    for tt in range(2 * N):
        libm.sincos(2.0 * math.pi * tt / (2 * N), ctypes.byref(s), ctypes.byref(c))
        assert w[tt, 0] == c.value and w[tt, 1] == s.value, tt
    g = native_chk.table(N, 2, 3, 55 if N == 2048 else 48, 12)
    assert [int(x) for x in g] == [pow(5, j, 2 * N) for j in range(N // 2)]
