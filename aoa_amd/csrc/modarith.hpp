// Modular arithmetic on uint64 residues, usable from host C++ and gfx950 HIP.
//
// Primes are < 2^63 (the first Q prime is 63-bit, the rest 45..55-bit), so
// every canonical residue fits in 63 bits and a + b never wraps.
//
//  * mul_shoup      constant operand w with precomputed floor(w * 2^64 / p):
//                   twiddles and RNS constants (one mulhi, two mullo).
//  * mul_barrett    variable x variable, exact for a, b < p (HAC 14.42 with a
//                   128-bit final correction so that the 63-bit prime is safe).
//  * ref_barrett    the reference's own Barrett reduction
//                   (aoa-ntt/src/cuda/htrans/common.cu:257-291).  It is exact
//                   for a < p^2 but not for the larger products that ModDown
//                   feeds it (aoa-ntt/src/cuda/manager.cu:331-334), where it
//                   returns exact - (2^nb - p) for a small fraction of terms.
//                   Bit-exact ModDown must therefore use this formula for that
//                   one product.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define AOA_HD __host__ __device__ __forceinline__
#else
#define AOA_HD static inline
#endif

namespace aoa {

AOA_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

AOA_HD uint64_t add_mod(uint64_t a, uint64_t b, uint64_t p) {
  uint64_t s = a + b;
  return s >= p ? s - p : s;
}

AOA_HD uint64_t sub_mod(uint64_t a, uint64_t b, uint64_t p) {
  return a >= b ? a - b : a + (p - b);
}

AOA_HD uint64_t neg_mod(uint64_t a, uint64_t p) { return a ? p - a : 0; }

// floor(w * 2^64 / p) for w < p.
AOA_HD uint64_t shoup_quot(uint64_t w, uint64_t p) {
#if defined(__HIP_DEVICE_COMPILE__)
  // Not used on the device: all Shoup quotients are precomputed on the host.
  return 0;
#else
  return (uint64_t)(((unsigned __int128)w << 64) / p);
#endif
}

// x * w mod p, x any 64-bit value, w < p < 2^63, wq = shoup_quot(w, p).
AOA_HD uint64_t mul_shoup(uint64_t x, uint64_t w, uint64_t wq, uint64_t p) {
  uint64_t q = mulhi64(x, wq);
  uint64_t r = x * w - q * p;  // in [0, 2p)
  return r >= p ? r - p : r;
}

// Same, result left in [0, 2p).
AOA_HD uint64_t mul_shoup_lazy(uint64_t x, uint64_t w, uint64_t wq, uint64_t p) {
  uint64_t q = mulhi64(x, wq);
  return x * w - q * p;
}

// Barrett constants for p with nb = bitlen(p) (2 <= nb <= 63).
struct barrett_t {
  uint64_t p;
  uint64_t mu;  // floor(2^(2nb) / p), < 2^(nb+1)
  int nb;
};

// Exact (hi:lo) mod p for a 128-bit value below p * 2^nb (covers a*b, a,b < p).
AOA_HD uint64_t barrett_reduce(uint64_t hi, uint64_t lo, uint64_t p, uint64_t mu, int nb) {
  // q1 = floor(z / 2^(nb-1)) < 2^(nb+1) <= 2^64
  const int s = nb - 1;
  uint64_t q1 = (lo >> s) | (hi << (64 - s));
  // q3 = floor(q1 * mu / 2^(nb+1)); q1*mu < 2^(2nb+2)
  uint64_t ph = mulhi64(q1, mu), pl = q1 * mu;
  const int t = nb + 1;
  uint64_t q3 = (t == 64) ? ph : ((pl >> t) | (ph << (64 - t)));
  // r = z - q3*p in [0, 3p): compute in 128 bits, then correct.
  uint64_t qh = mulhi64(q3, p), ql = q3 * p;
  uint64_t rl = lo - ql;
  uint64_t rh = hi - qh - (lo < ql ? 1u : 0u);
  // at most two corrections
  if (rh != 0 || rl >= p) {
    uint64_t nl = rl - p;
    rh -= (rl < p ? 1u : 0u);
    rl = nl;
  }
  if (rh != 0 || rl >= p) rl -= p;
  return rl;
}

AOA_HD uint64_t mul_barrett(uint64_t a, uint64_t b, uint64_t p, uint64_t mu, int nb) {
  return barrett_reduce(mulhi64(a, b), a * b, p, mu, nb);
}

// (s0 + s1 2^28 + s2 2^56 + s3 2^84) mod q for a total below q 2^nb: the sums
// of wide_accumulate and wide_correct (s3 < 2^44), by barrett_reduce of the
// 128-bit total.
AOA_HD uint64_t wide_reduce(const uint64_t s[4], uint64_t q, uint64_t mu, int nb) {
  const uint64_t t1 = s[1] << 28, t2 = s[2] << 56;
  uint64_t lo = s[0] + t1;
  uint64_t hi = (s[1] >> 36) + (s[2] >> 8) + (s[3] << 20) + (lo < t1 ? 1 : 0);
  lo += t2;
  hi += lo < t2 ? 1 : 0;
  return barrett_reduce(hi, lo, q, mu, nb);
}

// T = s0 + s1 2^28 + s2 2^56 mod q, left up to four multiples of q too large
// (in [0, 5q)), for T < 2^(62+nb) with 33 <= nb <= 56 the bit length of q and
// mu = floor(2^(63+nb) / q) (below 2^64).  The quotient estimate is the high
// word of T' mu, where T' = floor(s0 / 2^(nb-1)) + floor(s1 / 2^(nb-29)) +
// s2 2^(57-nb) is at most 2 below T / 2^(nb-1) (two floors), and the low
// 32x32 partial product is left out (at most 1 more).  So the estimate never
// exceeds floor(T / q) and is at most 4 below it; T minus its multiple of q
// is then below 2^64 and computed modulo 2^64.
AOA_HD uint64_t reduce_sums_lazy(uint64_t s0, uint64_t s1, uint64_t s2, uint64_t q, uint64_t mu, int nb) {
  const uint64_t t = (s0 >> (nb - 1)) + (s1 >> (nb - 29)) + (s2 << (57 - nb));
  const uint32_t t0 = (uint32_t)t, t1 = (uint32_t)(t >> 32);
  const uint32_t m0 = (uint32_t)mu, m1 = (uint32_t)(mu >> 32);
  const uint64_t x = (uint64_t)t0 * m1;
  const uint64_t y = (uint64_t)t1 * m0 + (uint32_t)x;
  const uint64_t qh = (uint64_t)t1 * m1 + (x >> 32) + (y >> 32);
  return s0 + (s1 << 28) + (s2 << 56) - qh * q;
}

// Exact sums of products a c for a prime q wider than 2^56 (ModDown's q0):
// a < 2^56 split at bit 28 (al, ah), c < 2^63 split at bits 28 and 56
// (c0, c1 < 2^28, c2 < 2^7):
// a c = al c0 + (al c1 + ah c0) 2^28 + (al c2 + ah c1) 2^56 + ah c2 2^84,
// six 32x32->64 multiply-adds into s[0..3]; for up to 64 terms no sum
// reaches 2^63.
AOA_HD void wide_accumulate(uint64_t s[4], uint64_t a, uint32_t c0, uint32_t c1, uint32_t c2) {
  const uint32_t al = (uint32_t)a & ((1u << 28) - 1), ah = (uint32_t)(a >> 28);
  s[0] += (uint64_t)al * c0;
  s[1] += (uint64_t)al * c1;
  s[1] += (uint64_t)ah * c0;
  s[2] += (uint64_t)al * c2;
  s[2] += (uint64_t)ah * c1;
  s[3] += (uint64_t)ah * c2;
}

// Adds d < 2^64 to the sums' total, in 28-bit pieces.
AOA_HD void wide_correct(uint64_t s[4], uint64_t d) {
  s[0] += d & ((1u << 28) - 1);
  s[1] += (d >> 28) & ((1u << 28) - 1);
  s[2] += d >> 56;
}

// Full 128-bit product from four 32x32->64 multiply-adds.
AOA_HD void mul128(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
  const uint64_t al = (uint32_t)a, ah = a >> 32, bl = (uint32_t)b, bh = b >> 32;
  const uint64_t ll = al * bl;
  const uint64_t lh = al * bh + (ll >> 32);
  const uint64_t hl = ah * bl + (uint32_t)lh;
  *hi = ah * bh + (lh >> 32) + (hl >> 32);
  *lo = (hl << 32) | (uint32_t)ll;
}

// a * b mod p for a, b < p < 2^62, result in [0, 2p): HAC 14.42 with the
// quotient floor(floor(t / 2^(nb-1)) * mu / 2^(nb+1)) taken as one mulhi by
// mu2 = mu * 2^(63-nb); it is at most 2 below floor(t/p), so t - q*p < 3p
// fits in 64 bits, and one conditional subtraction of 2p leaves [0, 2p).
AOA_HD uint64_t mul_barrett62_lazy(uint64_t a, uint64_t b, uint64_t p, uint64_t mu2, int nb) {
  uint64_t hi, lo;
  mul128(a, b, &hi, &lo);
  const int s = nb - 1;
  const uint64_t q1 = (lo >> s) | (hi << (64 - s));
  const uint64_t q = mulhi64(q1, mu2);
  uint64_t r = lo - q * p;
  const uint64_t p2 = 2 * p;
  return r >= p2 ? r - p2 : r;
}

AOA_HD uint64_t mul_barrett62(uint64_t a, uint64_t b, uint64_t p, uint64_t mu2, int nb) {
  const uint64_t r = mul_barrett62_lazy(a, b, p, mu2, nb);
  return r >= p ? r - p : r;
}

// a * b mod p for a, b < p, any supported prime.
AOA_HD uint64_t mul_var(uint64_t a, uint64_t b, uint64_t p, uint64_t mu, uint64_t mu2, int nb) {
  return (p >> 62) ? mul_barrett(a, b, p, mu, nb) : mul_barrett62(a, b, p, mu2, nb);
}

// The reference's Barrett reduction of a 128-bit value (hi:lo) with
// MU = floor(4^nb / p) (aoa-ntt/src/cuda/htrans/common.cu:257-291, constants at
// :485-506).  Written from its arithmetic: q = floor(floor(a / 2^(2nb-64)) *
// MU / 2^64) (truncated to 64 bits), A = a - q*p, and the result is
// (A mod 2^nb - (A >> nb) * p) mod 2^nb followed by one conditional
// subtraction of p.  Valid for 33 <= nb <= 63.
AOA_HD uint64_t ref_barrett(uint64_t hi, uint64_t lo, uint64_t p, uint64_t mu, int nb) {
  const uint64_t mask = (((uint64_t)1) << nb) - 1;
  const int s = 2 * nb - 64;  // 2..62 for nb in 33..63
  uint64_t xlo = (lo >> s) | (hi << (64 - s));
  uint64_t xhi = hi >> s;
  uint64_t q = mulhi64(xlo, mu) + xhi * mu;
  uint64_t qh = mulhi64(q, p), ql = q * p;
  // A = a - q*p as a 128-bit difference (mod 2^128), split at bit nb.
  uint64_t al = lo - ql;
  uint64_t ah = hi - qh - (lo < ql ? 1u : 0u);
  uint64_t a_lo = al & mask;
  uint64_t a_hi = (al >> nb) | (ah << (64 - nb));
  uint64_t r = (a_lo - a_hi * p) & mask;
  return r >= p ? r - p : r;
}

// ref_barrett(a * w) for a constant w < p with Shoup quotient wq, computed
// from the exact remainder r = a*w mod p.  Writing A = a*w - qhat*p =
// delta*p + r (qhat the reference's Barrett quotient, delta >= 0 its error)
// and eps = 2^nb - p, the reference returns r - eps when eps <= r < delta*eps
// and r otherwise, provided delta*eps < 2^nb.  [lo, thr) must cover
// [eps, delta_max*eps) for every product that can occur; inside it the
// literal formula is evaluated.
AOA_HD uint64_t ref_mulmod_shoup(uint64_t a, uint64_t w, uint64_t wq, uint64_t p, uint64_t mu, int nb,
                                 uint64_t lo, uint64_t thr) {
  uint64_t r = mul_shoup(a, w, wq, p);
  if (r >= lo && r < thr) r = ref_barrett(mulhi64(a, w), a * w, p, mu, nb);
  return r;
}

// Lazy form of ref_mulmod_shoup: returns a value in [0, 2p) congruent to
// ref_barrett(a * w) mod p.  r = Shoup remainder in [0, 2p); its canonical
// value lies in the deviation window only if r < thr or r - p < thr
// (unsigned), a superset of [lo, thr) that also covers [0, lo), where the
// literal formula is just as exact.  thr must be < p.
AOA_HD uint64_t ref_mulmod_shoup_lazy(uint64_t a, uint64_t w, uint64_t wq, uint64_t p, uint64_t mu, int nb,
                                      uint64_t thr) {
  uint64_t r = mul_shoup_lazy(a, w, wq, p);
  if (r < thr || (r - p) < thr) r = ref_barrett(mulhi64(a, w), a * w, p, mu, nb);
  return r;
}

// Screening a product a*c mod p for the deviation window without computing
// its remainder (a < 2^56, c < p < 2^63).  With W = floor(c 2^96 / p) =
// w2 2^64 + w1 2^32 + w0 and r = a*c mod p, the top 32-bit word of
// (a W mod 2^96) is f = floor(r 2^32 / p) or f - 1 (mod 2^32): a W =
// m 2^96 + r 2^96 / p - a theta with 0 <= theta < 1 and a theta < 2^64.
// moddown_screen sums the four partial products that reach that word, drops
// the carry out of the lowest one (a0 w0, at most one more) and adds 2, so it
// returns f, f + 1 or f + 2 (mod 2^32).  Hence r < thr implies
// moddown_screen(...) <= moddown_screen_thr(thr, p).
// bias must be 2^33; it is an argument so that a kernel can hold it in
// registers as the first multiply-add's addend.
AOA_HD uint32_t moddown_screen(uint32_t a0, uint32_t a1, uint32_t w0, uint32_t w1, uint32_t w2,
                               uint64_t bias = 2ull << 32) {
  // a0 w1 + a1 w0 sits at bit 32: its overflow past 2^64 is 2^96, which drops
  uint64_t t = (uint64_t)a0 * w1 + bias;
  t += (uint64_t)a1 * w0;
  return a1 * w1 + a0 * w2 + (uint32_t)(t >> 32);
}


// Bit length of a nonzero 64-bit value.
AOA_HD int bitlen64(uint64_t x) {
  int n = 0;
  while (x) {
    ++n;
    x >>= 1;
  }
  return n;
}

AOA_HD uint64_t pow_mod(uint64_t a, uint64_t e, uint64_t p, uint64_t mu, int nb) {
  uint64_t r = 1 % p, s = a % p;
  while (e) {
    if (e & 1) r = mul_barrett(r, s, p, mu, nb);
    s = mul_barrett(s, s, p, mu, nb);
    e >>= 1;
  }
  return r;
}

}  // namespace aoa
