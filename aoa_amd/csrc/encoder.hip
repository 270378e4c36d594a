// CKKS encoder on the GPU (FP64): the canonical-embedding FFT and the
// conversions between slot values and residues (aoa-ntt/src/ckks/encoder.cu:33-298).
//
// The FFT follows the reference's loop structure (HEAAN's fftSpecial /
// fftSpecialInv): a bit reversal and log2(n) radix-2 stages whose twiddle
// for butterfly j of a block of length len is exp(+-2 pi i (5^j mod 4 len) /
// (4 len)), read from the parameter set's table of 2N-th roots.  Every
// stage is one launch of n/2 threads (n <= N/2 slot values); the arithmetic
// is written operation by operation with contraction off, so each value is
// rounded exactly as the host loops round it.
#include "device_common.hpp"
#include "kernels.hpp"

#pragma clang fp contract(off)

namespace aoa {

struct alignas(16) cplx {
  double re, im;
};

__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  const double re = a.re * b.re - a.im * b.im;
  const double im = a.re * b.im + a.im * b.re;
  return cplx{re, im};
}

__device__ __forceinline__ cplx enc_root(const DevParams& P, uint32_t t) {
  return cplx{P.enc_w[2 * (size_t)t], P.enc_w[2 * (size_t)t + 1]};
}

// In-place bit reversal: the thread of the smaller index of each pair swaps
// it (arrayBitReverse, encoder.cu:11-23); with inv_n != 0 every value is
// also divided by n, the last step of fftSpecialInv (encoder.cu:117-124).
__global__ __launch_bounds__(256) void k_enc_bitrev(cplx* v, int n, int logn, int div) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int j = logn ? (int)(__brev((unsigned)i) >> (32 - logn)) : 0;
  if (i > j) return;
  cplx a = v[i], b = v[j];
  if (div) {
    const double dn = (double)n;
    a = cplx{a.re / dn, a.im / dn};
    b = cplx{b.re / dn, b.im / dn};
  }
  v[i] = b;
  if (i != j) v[j] = a;
}

// fftSpecial stage (encoder.cu:56-70): u = v[a], t = v[b] * ksi; v[a] = u + t, v[b] = u - t.
__global__ __launch_bounds__(256) void k_enc_stage_fwd(DevParams P, cplx* v, int n, int len) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= n / 2) return;
  const int lenh = len >> 1, lenq = len << 2;
  const int i = (b / lenh) * len, j = b % lenh;
  const uint32_t t = (P.rot_group[j] % (uint32_t)lenq) * (uint32_t)(2 * P.N / lenq);
  const cplx u = v[i + j], w = cmul(v[i + j + lenh], enc_root(P, t));
  v[i + j] = cplx{u.re + w.re, u.im + w.im};
  v[i + j + lenh] = cplx{u.re - w.re, u.im - w.im};
}

// fftSpecialInv stage (encoder.cu:96-110): u = v[a] + v[b], t = (v[a] - v[b]) * ksi^-1.
__global__ __launch_bounds__(256) void k_enc_stage_inv(DevParams P, cplx* v, int n, int len) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= n / 2) return;
  const int lenh = len >> 1, lenq = len << 2;
  const int i = (b / lenh) * len, j = b % lenh;
  const uint32_t t = ((uint32_t)lenq - P.rot_group[j] % (uint32_t)lenq) * (uint32_t)(2 * P.N / lenq);
  const cplx x = v[i + j], y = v[i + j + lenh];
  v[i + j] = cplx{x.re + y.re, x.im + y.im};
  v[i + j + lenh] = cmul(cplx{x.re - y.re, x.im - y.im}, enc_root(P, t));
}

// Residues of the encoded plaintext on limbs 0..k-1 (encoder.cu:166-221):
// value s (already divided by n) times the scale, truncated to int64, at
// coefficient s*gap (real part) and N/2 + s*gap (imaginary part), zero
// elsewhere; negative values map to p - (|v| mod p).
__global__ __launch_bounds__(256) void k_enc_pack(DevParams P, uint64_t* out, const cplx* u, int slots,
                                                  double scale, int k) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)k * P.N) return;
  const int l = (int)(e >> P.logN);
  const int pos = (int)(e & (size_t)(P.N - 1));
  const int half = P.N / 2, gap = half / slots;
  uint64_t r = 0;
  const int off = pos < half ? pos : pos - half;
  if (off % gap == 0) {
    const int s = off / gap;
    const double x = pos < half ? u[s].re : u[s].im;
    const int64_t v = (int64_t)(x * scale);
    const uint64_t p = P.primes[l].p;
    r = v >= 0 ? (uint64_t)v % p : (p - ((uint64_t)(-v) % p)) % p;
  }
  out[e] = r;
}

// Slot values from limb 0 in the coefficient domain (encoder.cu:223-257):
// the residue x <= p/2 gives x / scale, a larger one (x - p) / scale, both
// converted to double before the subtraction as the reference does.
__global__ __launch_bounds__(256) void k_dec_unpack(DevParams P, cplx* out, const uint64_t* in, int slots,
                                                    double scale) {
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= slots) return;
  const int half = P.N / 2, gap = half / slots;
  const uint64_t p = P.primes[0].p;
  const uint64_t a = in[(size_t)s * gap], b = in[(size_t)s * gap + half];
  const double re = a <= (p >> 1) ? (double)a / scale : ((double)a - (double)p) / scale;
  const double im = b <= (p >> 1) ? (double)b / scale : ((double)b - (double)p) / scale;
  out[s] = cplx{re, im};
}

static int ilog2(int n) {
  int g = 0;
  while ((1 << g) < n) ++g;
  return g;
}

static unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

// Stages of length <= B inside aligned blocks of B values, in LDS: one
// workgroup per block (B <= 4096: 64 KiB of values and 8 KiB of the rotation
// group).  Forward: stages len = 2..B; inverse: len = B..2.  With whole != 0
// the block is the whole vector and the bit reversal is folded into the load
// (forward) or, with the division by n, into the store (inverse).  Each
// thread takes the butterflies b = t, t + 1024 of every stage and requests
// the next stage's roots before the barrier, so the table latency overlaps
// it; the arithmetic of each butterfly is unchanged.
constexpr int kFftThreads = 1024;
constexpr int kFftBlock = 4096;
constexpr int kFftPer = kFftBlock / 2 / kFftThreads;  // butterflies per thread and stage

__global__ __launch_bounds__(kFftThreads) void k_enc_fft_block(DevParams P, cplx* v, int B, int logB, int inverse,
                                                               int whole) {
  extern __shared__ cplx buf[];
  uint32_t* rot = reinterpret_cast<uint32_t*>(buf + B);  // rotation group 5^j mod 2N, j < B/2
  const int t = threadIdx.x;
  cplx* blk = v + (size_t)blockIdx.x * B;
  if (whole && !inverse) {
    for (int i = t; i < B; i += kFftThreads) buf[logB ? (__brev((unsigned)i) >> (32 - logB)) : 0] = blk[i];
  } else {
    for (int i = t; i < B; i += kFftThreads) buf[i] = blk[i];
  }
  for (int i = t; i < B / 2; i += kFftThreads) rot[i] = P.rot_group[i];
  __syncthreads();
  // root of butterfly b in the stage of length len
  auto root = [&](int len, int b) {
    const int lenh = len >> 1, lenq = len << 2;
    const uint32_t step = (uint32_t)(2 * P.N / lenq);
    const uint32_t r = rot[b % lenh] % (uint32_t)lenq;
    return enc_root(P, (inverse ? (uint32_t)lenq - r : r) * step);
  };
  cplx w[kFftPer];
  const int first = inverse ? B : 2;
#pragma unroll
  for (int q = 0; q < kFftPer; ++q) {
    const int b = t + q * kFftThreads;
    if (b < B / 2) w[q] = root(first, b);
  }
  if (!inverse) {
    for (int len = 2; len <= B; len <<= 1) {
      const int lenh = len >> 1;
#pragma unroll
      for (int q = 0; q < kFftPer; ++q) {
        const int b = t + q * kFftThreads;
        if (b >= B / 2) continue;
        const int i = (b / lenh) * len, j = b % lenh;
        const cplx u = buf[i + j], x = cmul(buf[i + j + lenh], w[q]);
        buf[i + j] = cplx{u.re + x.re, u.im + x.im};
        buf[i + j + lenh] = cplx{u.re - x.re, u.im - x.im};
      }
      if (len < B) {
#pragma unroll
        for (int q = 0; q < kFftPer; ++q) {
          const int b = t + q * kFftThreads;
          if (b < B / 2) w[q] = root(len << 1, b);
        }
      }
      __syncthreads();
    }
  } else {
    for (int len = B; len >= 2; len >>= 1) {
      const int lenh = len >> 1;
#pragma unroll
      for (int q = 0; q < kFftPer; ++q) {
        const int b = t + q * kFftThreads;
        if (b >= B / 2) continue;
        const int i = (b / lenh) * len, j = b % lenh;
        const cplx x = buf[i + j], y = buf[i + j + lenh];
        buf[i + j] = cplx{x.re + y.re, x.im + y.im};
        buf[i + j + lenh] = cmul(cplx{x.re - y.re, x.im - y.im}, w[q]);
      }
      if (len > 2) {
#pragma unroll
        for (int q = 0; q < kFftPer; ++q) {
          const int b = t + q * kFftThreads;
          if (b < B / 2) w[q] = root(len >> 1, b);
        }
      }
      __syncthreads();
    }
  }
  if (whole && inverse) {
    const double dn = (double)B;
    for (int i = t; i < B; i += kFftThreads)
      blk[logB ? (__brev((unsigned)i) >> (32 - logB)) : 0] = cplx{buf[i].re / dn, buf[i].im / dn};
  } else {
    for (int i = t; i < B; i += kFftThreads) blk[i] = buf[i];
  }
}

// The stages of length <= 4096 run in LDS (one launch); a vector longer than
// 4096 adds one launch per remaining stage and a separate bit reversal.
void launch_special_fft(const DevParams& P, double* vals, int n, int inverse, hipStream_t st) {
  cplx* v = reinterpret_cast<cplx*>(vals);
  const int logn = ilog2(n);
  const int B = n < kFftBlock ? n : kFftBlock;
  const int logB = ilog2(B);
  const size_t lds = (size_t)B * sizeof(cplx) + (size_t)(B / 2) * sizeof(uint32_t);
  const unsigned nblk = (unsigned)(n / B);
  const int whole = n == B;
  if (!inverse) {
    if (!whole) k_enc_bitrev<<<blocks(n), 256, 0, st>>>(v, n, logn, 0);
    k_enc_fft_block<<<nblk, kFftThreads, lds, st>>>(P, v, B, logB, 0, whole);
    for (int len = 2 * B; len <= n; len <<= 1) k_enc_stage_fwd<<<blocks(n / 2), 256, 0, st>>>(P, v, n, len);
  } else {
    for (int len = n; len > B; len >>= 1) k_enc_stage_inv<<<blocks(n / 2), 256, 0, st>>>(P, v, n, len);
    k_enc_fft_block<<<nblk, kFftThreads, lds, st>>>(P, v, B, logB, 1, whole);
    if (!whole) k_enc_bitrev<<<blocks(n), 256, 0, st>>>(v, n, logn, 1);
  }
}

void launch_encode_pack(const DevParams& P, uint64_t* out, const double* vals, int slots, double scale, int k,
                        hipStream_t st) {
  k_enc_pack<<<blocks((size_t)k * P.N), 256, 0, st>>>(P, out, reinterpret_cast<const cplx*>(vals), slots, scale,
                                                       k);
}

void launch_decode_unpack(const DevParams& P, double* vals, const uint64_t* in, int slots, double scale,
                          hipStream_t st) {
  k_dec_unpack<<<blocks(slots), 256, 0, st>>>(P, reinterpret_cast<cplx*>(vals), in, slots, scale);
}

}  // namespace aoa
