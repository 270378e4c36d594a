"""Encoder FFT (FP64, the only floating-point code on the path): the product's
canonical-embedding FFT against the oracle's restatement of the reference's
fftSpecial / fftSpecialInv loops (aoa-ntt/src/ckks/encoder.cu:33-124) and
against the definition z_j = sum_s u_s w^(s 5^j), w = exp(2 pi i / 4n)."""
import ctypes

import numpy as np
import pytest

import native_chk
import orc

TOL = 1e-9


def _call(fn, vals, *args):
    buf = np.ascontiguousarray(vals.astype(np.complex128)).view(np.float64).copy()
    fn(buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), *args)
    return buf.view(np.complex128)


def product_fft(vals, inverse):
    return _call(native_chk.load().chk_fft_special, vals, len(vals), int(inverse))


def oracle_fft(vals, ringN, inverse):
    return _call(orc.load().orc_fft_special, vals, len(vals), ringN, int(inverse))


def definition(u):
    n = len(u)
    M = 4 * n
    pts = [pow(5, j, M) for j in range(n)]
    s = np.arange(n)
    return np.array([np.sum(u * np.exp(2j * np.pi * s * t / M)) for t in pts])


@pytest.mark.parametrize("n,ringN", [(4, 4096), (64, 4096), (2048, 4096), (512, 65536)])
def test_product_fft_matches_reference_loops(n, ringN):
    rng = np.random.default_rng(n)
    v = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    assert np.max(np.abs(product_fft(v, False) - oracle_fft(v, ringN, False))) < TOL * n
    assert np.max(np.abs(product_fft(v, True) - oracle_fft(v, ringN, True))) < TOL


@pytest.mark.parametrize("n", [8, 32, 128])
def test_fft_definition_and_roundtrip(n):
    rng = np.random.default_rng(7 * n)
    u = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    assert np.max(np.abs(product_fft(u, False) - definition(u))) < TOL * n
    assert np.max(np.abs(oracle_fft(u, 4096, False) - definition(u))) < TOL * n
    assert np.max(np.abs(product_fft(product_fft(u, True), False) - u)) < TOL


def test_embedding_is_multiplicative():
    """Slots multiply when plaintext polynomials multiply negacyclically: the
    encoder's packing (real parts at s*gap, imaginary at N/2 + s*gap) turns a
    slot vector into the polynomial whose evaluations at the roots
    exp(i pi 5^j / N) are the slots."""
    N, n = 64, 8
    gap = (N // 2) // n
    rng = np.random.default_rng(3)

    def poly_of(z):
        u = product_fft(z, True)
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
