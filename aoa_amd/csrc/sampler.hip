// Device-side sampler: the distributions of the reference's Sampler
// (aoa-ntt/src/cuda/sampler.cu:10-225) drawn on the GPU from a counter-based
// generator, so that keys and encryption randomness are born in HBM.
//
// Generator: Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC 2011): ten rounds of two 32x32->64
// multiplies with the multipliers 0xD2511F53 and 0xCD9E8D57, the key bumped by
// the Weyl constants 0x9E3779B9 and 0xBB67AE85 between rounds.  Every output
// is a pure function of (seed, nonce, kind, limb, coefficient), so each
// thread draws its own values with no state and the result does not depend
// on the launch shape.  Counter words: {coefficient or attempt, kind << 24 |
// limb (0xFFFFFF for values shared by all limbs), nonce low, nonce high}; key
// words: {seed low, seed high}.
//
// Distributions:
//  * uniform: 128 bits X per residue, r = floor(X q / 2^128) (statistical
//    distance below q / 2^128 from uniform on [0, q));
//  * Gaussian: e = round(3.2 z), z ~ N(0, 1), as sample_DG (sampler.cu:72-117),
//    by inversion of a 64-entry table of P(e <= m) scaled to 2^64 (integer
//    compares only), one value per coefficient shared by all limbs;
//  * ZO: exactly N/2 coefficients +-1, the others 0, as sample_Z0
//    (sampler.cu:119-150, which shuffles N/2 values +-1 over the N
//    positions): coefficient i is nonzero when a keyed permutation of
//    [0, N) -- six Feistel rounds on the index bits, round function Philox --
//    maps it below N/2, with the sign of its own draw (zo_nonzero);
//  * narrow: uniform on {-1, 0, 1};
//  * HW: exactly 64 coefficients +-1 at distinct positions (sample_hw,
//    sampler.cu:157-192), chosen by one wave in attempt order.
// Negative values v are stored as q + v, as the reference does.
#include "device_common.hpp"
#include "kernels.hpp"

namespace aoa {

struct u32x4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
    const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
    c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
  }
  return c;
}

__device__ __forceinline__ u32x4 draw(uint64_t seed, uint64_t nonce, uint32_t tag, uint32_t idx) {
  return philox4x32_10(u32x4{idx, tag, (uint32_t)nonce, (uint32_t)(nonce >> 32)}, (uint32_t)seed,
                       (uint32_t)(seed >> 32));
}

__device__ __forceinline__ uint64_t lo64(u32x4 c) { return ((uint64_t)c.y << 32) | c.x; }
__device__ __forceinline__ uint64_t hi64(u32x4 c) { return ((uint64_t)c.w << 32) | c.z; }

__device__ __forceinline__ uint64_t small_residue(int64_t v, uint64_t q) { return v >= 0 ? (uint64_t)v : q + v; }

// Whether coefficient i of a ZO draw is nonzero: the index's logN bits split
// into a high part H of ceil(logN/2) bits and a low part Lo of floor(logN/2)
// bits; six rounds alternately replace Lo by Lo ^ F_r(H) and H by
// H ^ F_r(Lo) (the last round changes H, on which the choice depends), F_r
// the first word of the Philox draw (kind ZO, round tag 0xFFFF00 + r) at the
// other part, cut to the part's width.  Each round is invertible, so
// i -> (H, Lo) is a permutation of [0, N) and exactly the N/2 indices whose
// image has H's top bit clear are nonzero.  (The sign is bit 0
// of the second word of i's own draw, tag 0xFFFFFF.)
__device__ __forceinline__ bool zo_nonzero(int logN, uint64_t seed, uint64_t nonce, uint32_t i) {
  const int hb = (logN + 1) / 2, lb = logN / 2;
  uint32_t h = i >> lb, lo = i & ((1u << lb) - 1u);
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    const uint32_t tag = ((uint32_t)SAMPLE_ZO << 24) | (0xFFFF00u + (uint32_t)r);
    if (r & 1)
      h ^= draw(seed, nonce, tag, lo).x & ((1u << hb) - 1u);
    else
      lo ^= draw(seed, nonce, tag, h).x & ((1u << lb) - 1u);
  }
  return (h >> (hb - 1)) == 0;
}

__global__ __launch_bounds__(256) void k_sample_uniform(DevParams P, uint64_t* out, int res0, int nres, uint64_t seed,
                                                        uint64_t nonce) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)nres * P.N) return;
  const int l = res0 + (int)(e >> P.logN);
  const uint32_t i = (uint32_t)(e & (size_t)(P.N - 1));
  const uint64_t q = ld_prime(P.primes, l).p;
  const u32x4 c = draw(seed, nonce, ((uint32_t)SAMPLE_UNIFORM << 24) | (uint32_t)l, i);
  // floor((xh 2^64 + xl) q / 2^128)
  const uint64_t t_hi = __umul64hi(lo64(c), q);
  const uint64_t xh = hi64(c);
  const uint64_t u_lo = xh * q, u_hi = __umul64hi(xh, q);
  const uint64_t s = u_lo + t_hi;
  out[(size_t)l * P.N + i] = u_hi + (s < u_lo ? 1 : 0);
}

// Gaussian, ZO and narrow: one value per coefficient, written to every limb.
__device__ __forceinline__ void sample_small_at(const DevParams& P, uint64_t* out, int res0, int nres, int kind,
                                                uint64_t seed, uint64_t nonce, uint32_t i) {
  const u32x4 c = draw(seed, nonce, ((uint32_t)kind << 24) | 0xFFFFFFu, i);
  int64_t v;
  if (kind == SAMPLE_GAUSSIAN) {
    const uint64_t u = lo64(c);
    const cptr<uint64_t> cdt = as_const(P.gauss_cdt);
    int cnt = 0;
#pragma unroll 8
    for (int j = 0; j < 2 * kGaussHalf; ++j) cnt += u >= cdt[j] ? 1 : 0;
    v = cnt - kGaussHalf;
  } else if (kind == SAMPLE_ZO) {
    v = zo_nonzero(P.logN, seed, nonce, i) ? ((c.y & 1u) ? -1 : 1) : 0;
  } else {  // narrow
    v = (int64_t)__umul64hi(lo64(c), 3) - 1;
  }
  for (int l = res0; l < res0 + nres; ++l)
    out[(size_t)l * P.N + i] = small_residue(v, ld_prime(P.primes, l).p);
}

__global__ __launch_bounds__(256) void k_sample_small(DevParams P, uint64_t* out, int res0, int nres, int kind,
                                                      uint64_t seed, uint64_t nonce) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= (uint32_t)P.N) return;
  sample_small_at(P, out, res0, nres, kind, seed, nonce, i);
}

// An encryption's three draws in one launch (blockIdx.y = 0, 1, 2): u (ZO) at
// nonce, e0 and e1 (Gaussian) at nonce + 1 and nonce + 2, limbs 0..nres-1;
// the same values as three aoa_sample calls.
__global__ __launch_bounds__(256) void k_sample_encrypt(DevParams P, uint64_t* u, uint64_t* e0, uint64_t* e1,
                                                        int nres, uint64_t seed, uint64_t nonce) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= (uint32_t)P.N) return;
  const int y = blockIdx.y;
  sample_small_at(P, y == 0 ? u : (y == 1 ? e0 : e1), 0, nres, y == 0 ? SAMPLE_ZO : SAMPLE_GAUSSIAN, seed,
                  nonce + (uint64_t)y, i);
}

void launch_sample_encrypt(const DevParams& P, uint64_t* u, uint64_t* e0, uint64_t* e1, int nres, uint64_t seed,
                           uint64_t nonce, hipStream_t st) {
  k_sample_encrypt<<<dim3((unsigned)((P.N + 255) / 256), 3), 256, 0, st>>>(P, u, e0, e1, nres, seed, nonce);
}

__global__ __launch_bounds__(256) void k_fill_zero(uint64_t* out, size_t n) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e < n) out[e] = 0;
}

// Hamming-weight secret: one wave.  Lane 0 draws candidate positions in
// attempt order and keeps the first kHammingWeight distinct ones (an LDS
// bitmap of N bits); then lane t writes entry t to every limb.  At most
// kHwMaxAttempts draws (the chance of needing more than 4096 draws for 64
// distinct positions out of N >= 2048 is below 2^-1000).
__global__ __launch_bounds__(64) void k_sample_hw(DevParams P, uint64_t* out, int res0, int nres, uint64_t seed,
                                                  uint64_t nonce) {
  __shared__ uint32_t bitmap[65536 / 32];
  __shared__ uint32_t pos[kHammingWeight];
  __shared__ int32_t sgn[kHammingWeight];
  __shared__ int found;
  const int t = threadIdx.x;
  const int words = P.N / 32;
  for (int w = t; w < words; w += 64) bitmap[w] = 0;
  __syncthreads();
  const int weight = kHammingWeight < P.N ? kHammingWeight : P.N;
  if (t == 0) {
    int n = 0;
    for (uint32_t a = 0; a < (uint32_t)kHwMaxAttempts && n < weight; ++a) {
      const u32x4 c = draw(seed, nonce, ((uint32_t)SAMPLE_HW << 24) | 0xFFFFFFu, a);
      const uint32_t p = c.x & (uint32_t)(P.N - 1);
      const uint32_t bit = 1u << (p & 31);
      if (bitmap[p >> 5] & bit) continue;
      bitmap[p >> 5] |= bit;
      pos[n] = p;
      sgn[n] = (c.y & 1u) ? -1 : 1;
      ++n;
    }
    found = n;
  }
  __syncthreads();
  if (t < found)
    for (int l = res0; l < res0 + nres; ++l)
      out[(size_t)l * P.N + pos[t]] = small_residue(sgn[t], ld_prime(P.primes, l).p);
}

static unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

void launch_sample(const DevParams& P, uint64_t* out, int res0, int nres, int kind, uint64_t seed, uint64_t nonce,
                   hipStream_t st) {
  const size_t n = (size_t)nres * P.N;
  switch (kind) {
    case SAMPLE_UNIFORM:
      k_sample_uniform<<<blocks(n), 256, 0, st>>>(P, out, res0, nres, seed, nonce);
      break;
    case SAMPLE_HW:
      k_fill_zero<<<blocks(n), 256, 0, st>>>(out + (size_t)res0 * P.N, n);
      k_sample_hw<<<1, 64, 0, st>>>(P, out, res0, nres, seed, nonce);
      break;
    default:
      k_sample_small<<<blocks(P.N), 256, 0, st>>>(P, out, res0, nres, kind, seed, nonce);
      break;
  }
}

}  // namespace aoa
