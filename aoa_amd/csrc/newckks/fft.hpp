// Canonical-embedding FFT of the CKKS encoder (FP64, host).
//
// For a slot vector of n = size values, decoding computes
//   z_j = sum_{s<n} u_s w^(s 5^j),  w = exp(2 pi i / (4n)),  j < n,
// i.e. the plaintext polynomial evaluated at the primitive 2N-th roots in the
// rotation group 5^j (the reference's fftSpecial, aoa-ntt/src/ckks/encoder.cu:33-73);
// encoding is its exact inverse u_s = (1/n) sum_j z_j w^(-s 5^j)
// (fftSpecialInv, encoder.cu:75-124).  Both are computed with one radix-2
// FFT of length 4n over all 4n-th roots of unity, reading or writing only
// the points w^(5^j mod 4n).
#pragma once
#include <cmath>
#include <complex>
#include <vector>

namespace aoa_fft {

inline void fft_radix2(std::vector<std::complex<double>>& a, int sign) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const double ang = sign * 2.0 * M_PI / (double)len;
    std::vector<std::complex<double>> w(len / 2);
    for (size_t k = 0; k < len / 2; ++k) w[k] = std::complex<double>(std::cos(ang * k), std::sin(ang * k));
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; ++k) {
        std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w[k];
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
  }
}

inline std::vector<size_t> rot_points(long size) {
  const size_t M = 4 * (size_t)size;
  std::vector<size_t> pts(size);
  size_t g = 1;
  for (long j = 0; j < size; ++j) {
    pts[j] = g;
    g = (g * 5) % M;
  }
  return pts;
}

// Decoding direction.
inline void special_fft(std::complex<double>* vals, long size) {
  const size_t M = 4 * (size_t)size;
  std::vector<std::complex<double>> a(M, 0.0);
  for (long s = 0; s < size; ++s) a[s] = vals[s];
  fft_radix2(a, +1);  // a[t] = sum_s u_s w^(s t)
  const std::vector<size_t> pts = rot_points(size);
  for (long j = 0; j < size; ++j) vals[j] = a[pts[j]];
}

// Encoding direction (exact inverse of special_fft).
inline void special_fft_inv(std::complex<double>* vals, long size) {
  const size_t M = 4 * (size_t)size;
  std::vector<std::complex<double>> a(M, 0.0);
  const std::vector<size_t> pts = rot_points(size);
  for (long j = 0; j < size; ++j) a[pts[j]] = vals[j];
  fft_radix2(a, -1);  // a[s] = sum_t W_t w^(-s t)
  for (long s = 0; s < size; ++s) vals[s] = a[s] / (double)size;
}

}  // namespace aoa_fft
