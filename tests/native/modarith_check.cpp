// Host-side checks of the product's modular arithmetic (aoa_amd/csrc/modarith.hpp
// and the parameter tables of params.cpp), compiled with g++ by the CPU tests.
#include <stdint.h>

#include <random>

#include "modarith.hpp"
#include <cstring>
#include "params.hpp"

using namespace aoa;

extern "C" {

uint64_t chk_ref_barrett(uint64_t hi, uint64_t lo, uint64_t p, uint64_t mu, int nb) {
  return ref_barrett(hi, lo, p, mu, nb);
}

uint64_t chk_mul_barrett(uint64_t a, uint64_t b, uint64_t p, uint64_t mu, int nb) {
  return mul_barrett(a, b, p, mu, nb);
}

// mul_barrett, mul_barrett62 (primes below 2^62), mul128 and mul_shoup
// against exact 128-bit arithmetic, for every prime of a parameter set;
// returns the number of mismatches (or -1 on bad params).
int64_t chk_exact_mults(int N, int k, int kl, int prec, int64_t trials, uint64_t seed) {
  HostParams H;
  if (!build_host_params(N, k, kl, prec, &H).empty()) return -1;
  kl = H.K;  // kl = 0 selects k + 1
  std::mt19937_64 rng(seed);
  int64_t bad = 0;
  for (size_t r = 0; r < H.primes.size(); ++r) {
    const DevPrime& d = H.dprimes[r];
    const uint64_t p = d.p;
    for (int64_t t = 0; t < trials; ++t) {
      uint64_t a = rng() % p, b = rng() % p;
      if (t < 8) {  // edge operands
        a = (t & 1) ? p - 1 : 0;
        b = (t & 2) ? p - 1 : (t & 4 ? 1 : 0);
      }
      const uint64_t e = (uint64_t)(((unsigned __int128)a * b) % p);
      if (mul_barrett(a, b, p, d.mu, (int)d.nb) != e) ++bad;
      if (p < (1ull << 62)) {
        if (mul_barrett62(a, b, p, d.mu2, (int)d.nb) != e) ++bad;
        const uint64_t lz = mul_barrett62_lazy(a, b, p, d.mu2, (int)d.nb);
        if (lz >= 2 * p || lz % p != e) ++bad;
      }
      uint64_t hi, lo;
      mul128(a, b, &hi, &lo);
      const unsigned __int128 full = (unsigned __int128)a * b;
      if (hi != (uint64_t)(full >> 64) || lo != (uint64_t)full) ++bad;
      const uint64_t wq = shoup_quot(b, p);
      const uint64_t x = rng();  // Shoup accepts any 64-bit left operand
      if (mul_shoup(x, b, wq, p) != (uint64_t)(((unsigned __int128)x * b) % p)) ++bad;
    }
  }
  return bad;
}

// ref_mulmod_shoup (the fast ModDown term) against the literal ref_barrett
// on the ModDown operand range: a < max special prime, c = (B/b_j) mod q_i.
// Random operands plus operands built so that the exact remainder sits at
// the edges of the deviation window [eps, thr).  out[0] = mismatches,
// out[1] = products where ref_barrett differs from the exact remainder,
// out[2] = products that took the literal path, out[3] = products that
// moddown_screen passed (Q primes below 2^56), which must include every
// deviating product.
int chk_moddown_terms(int N, int k, int kl, int prec, int64_t trials, uint64_t seed, int64_t* out) {
  HostParams H;
  if (!build_host_params(N, k, kl, prec, &H).empty()) return -1;
  kl = H.K;  // kl = 0 selects k + 1
  std::mt19937_64 rng(seed);
  int64_t bad = 0, quirks = 0, slow = 0, screened = 0;
  for (int i = 0; i < k; ++i) {
    const DevPrime& d = H.dprimes[i];
    const uint64_t p = d.p;
    for (int64_t t = 0; t < trials; ++t) {
      const int j = (int)(rng() % kl);
      const uint64_t bj = H.primes[k + j];
      const uint64_t c = H.bhat_mod_q[((size_t)i * kl + j) * 2];
      const uint64_t cq = H.bhat_mod_q[((size_t)i * kl + j) * 2 + 1];
      uint64_t a = rng() % bj;
      if (t % 4 != 0 && c != 0) {
        // aim the exact remainder at eps-1, eps, thr-1, thr or a random
        // point inside the window: a = target * c^-1 + m p < b_j
        uint64_t target;
        const uint64_t span = (d.quirk_thr > d.quirk_lo && d.quirk_thr != ~(uint64_t)0) ? d.quirk_thr - d.quirk_lo : 1;
        switch (t % 4) {
          case 1: target = d.quirk_lo + (rng() % 3) - 1; break;
          case 2: target = (d.quirk_thr == ~(uint64_t)0 ? d.eps : d.quirk_thr) + (rng() % 3) - 1; break;
          default: target = d.quirk_lo + rng() % span; break;
        }
        target %= p;
        const uint64_t cinv = inv_mod(c, p);
        uint64_t base = (uint64_t)(((unsigned __int128)target * cinv) % p);
        const uint64_t mmax = bj > base ? (bj - 1 - base) / p : 0;
        // half of them with the largest such a (the largest quotient error)
        a = base + (mmax ? ((t & 8) ? mmax : rng() % (mmax + 1)) : 0) * p;
        if (a >= bj) a = base % bj;
      }
      const uint64_t ref = ref_barrett(mulhi64(a, c), a * c, p, d.mu, (int)d.nb);
      const uint64_t fast = ref_mulmod_shoup(a, c, cq, p, d.mu, (int)d.nb, d.quirk_lo, d.quirk_thr);
      const uint64_t thr = d.quirk_thr < p ? d.quirk_thr : p;
      const uint64_t lazy = ref_mulmod_shoup_lazy(a, c, cq, p, d.mu, (int)d.nb, thr);
      const uint64_t exact = (uint64_t)(((unsigned __int128)a * c) % p);
      if (ref != fast) ++bad;
      if (lazy >= 2 * p || lazy % p != ref) ++bad;
      if (ref != exact) ++quirks;
      if (exact >= d.quirk_lo && exact < d.quirk_thr) ++slow;
      {
        // the screen must pass every product that deviates (the device
        // table's words: q0's are not in the host's [i][j] table)
        const MdTerm& m = H.md_terms[(size_t)j * k + i];
        const uint32_t f = moddown_screen((uint32_t)a, (uint32_t)(a >> 32), m.w0, m.w1, m.w2);
        const bool in = f <= H.md_thr[i];
        if (ref != exact && !in) ++bad;
        if (in) ++screened;
      }
    }
  }
  out[0] = bad;
  out[1] = quirks;
  out[2] = slow;
  out[3] = screened;
  return 0;
}

// The quotient error delta = floor(x / q) - qhat of the reference's Barrett
// reduction (qhat as ref_barrett computes it) against the bound behind the
// deviation window, quirk_thr = delta_max * eps, for products x = a c with
// a below the largest special prime and c < q: random x, x at the top of the
// range and x with the largest possible floor(x / 2^(2nb - 64)).  Returns the
// number of products with delta * eps >= quirk_thr, or -1 on bad params.
int64_t chk_barrett_delta(int N, int k, int kl, int prec, int64_t trials, uint64_t seed) {
  HostParams H;
  if (!build_host_params(N, k, kl, prec, &H).empty()) return -1;
  kl = H.K;  // kl = 0 selects k + 1
  std::mt19937_64 rng(seed);
  uint64_t bmax = 0;
  for (int j = 0; j < kl; ++j) bmax = H.primes[k + j] > bmax ? H.primes[k + j] : bmax;
  int64_t bad = 0;
  for (int i = 0; i < k; ++i) {
    const DevPrime& d = H.dprimes[i];
    if (d.quirk_thr == ~(uint64_t)0) continue;
    const uint64_t q = d.p;
    const int nb = (int)d.nb, s = 2 * nb - 64;
    const unsigned __int128 xmax = (unsigned __int128)(bmax - 1) * (q - 1);
    for (int64_t t = 0; t < trials; ++t) {
      unsigned __int128 x;
      if (t % 3 == 0) {
        x = xmax - (rng() % 4096);
      } else if (t % 3 == 1) {
        // the largest X1 with its smallest low part
        x = ((xmax >> s) << s) + (rng() % 16);
        if (x > xmax) x = xmax;
      } else {
        x = (((unsigned __int128)rng() << 64) | rng()) % (xmax + 1);
      }
      const uint64_t hi = (uint64_t)(x >> 64), lo = (uint64_t)x;
      const uint64_t xlo = (lo >> s) | (hi << (64 - s)), xhi = hi >> s;
      const uint64_t qhat = mulhi64(xlo, d.mu) + xhi * d.mu;
      const unsigned __int128 delta = x / q - qhat;
      if (delta * d.eps >= d.quirk_thr) ++bad;
    }
  }
  return bad;
}

// moddown_screen against its bound// moddown_screen against its bound on the Q primes below 2^56 of three
// parameter sets: for random and extreme a < 2^56 and c < p, the returned
// word minus floor(r 2^32 / p) (r = a c mod p) is 0, 1 or 2 modulo 2^32.
// Returns the number of violations.
int64_t chk_moddown_screen_words(int64_t trials, uint64_t seed) {
  std::mt19937_64 rng(seed);
  int64_t bad = 0;
  const int shapes[3][4] = {{65536, 30, 31, 48}, {4096, 4, 5, 52}, {16384, 5, 5, 55}};
  for (const auto& sh : shapes) {
    HostParams H;
    if (!build_host_params(sh[0], sh[1], sh[2], sh[3], &H).empty()) return -1;
    for (int i = 0; i < sh[1]; ++i) {
      const uint64_t p = H.primes[i];
      if (H.dprimes[i].nb > 56) continue;
      for (int64_t t = 0; t < trials; ++t) {
        uint64_t c = rng() % p, a = rng() >> 8;
        if (t < 9) {
          c = (t % 3 == 0) ? p - 1 : (t % 3 == 1 ? 1 : 0);
          a = (t / 3 == 0) ? (1ull << 56) - 1 : (t / 3 == 1 ? 0 : p);
        }
        if (t >= 9 && t < 4096 + 9) a = (uint64_t)(t - 9);  // every a < 2^12
        uint32_t w[3];
        moddown_screen_words(c, p, w);
        const uint64_t r = (uint64_t)(((unsigned __int128)a * c) % p);
        const uint32_t f = (uint32_t)((((unsigned __int128)r) << 32) / p);
        const uint32_t got = moddown_screen((uint32_t)a, (uint32_t)(a >> 32), w[0], w[1], w[2]);
        const uint32_t d = got - f;
        if (d > 2) ++bad;
      }
    }
  }
  return bad;
}

// The three-product accumulation of ModDown: for random and extreme a_j <
// b_j, sum_j a_j c_ij = s0 + (sm - s0 - s2) 2^28 + s2 2^56 with s0, s2, sm
// summed in 64 bits from the 28-bit halves of a_j and the MdConst halves of
// c_ij, compared with the 128-bit sum, for every Q prime below 2^56.
// Returns the number of mismatches (or -1 on bad params).
// ModDown's sums for a Q prime wider than 2^56 (q0) as moddown_wide forms
// them (wide_accumulate over the md_terms pieces of c, the screen and
// wide_correct for the products it passes, wide_reduce) against the
// reference's term-by-term sum of ref_barrett products mod q.  Each trial
// draws K operands a_j < b_j: random, the largest (b_j - 1), or aimed so that
// a_j c_j mod q lies at the edges of or inside the deviation window; trial 0
// takes the largest everywhere (the largest total).  Also checks that the
// pieces make up c.  Returns the number of mismatches, -1 on bad params, -2
// when the screen passed no product (the aimed products missed the window).
// (For the reference's 63-bit q0 the quotient error stays below 2, so none
// of its products deviates; the screened ones add q through wide_correct.)
int64_t chk_moddown_wide(int N, int k, int kl, int prec, int64_t trials, uint64_t seed) {
  HostParams H;
  if (!build_host_params(N, k, kl, prec, &H).empty()) return -1;
  kl = H.K;
  std::mt19937_64 rng(seed);
  int64_t bad = 0, screened = 0;
  for (int i = 0; i < k; ++i) {
    const DevPrime& d = H.dprimes[i];
    if (d.nb <= 56) continue;
    const uint64_t p = d.p;
    for (int j = 0; j < kl; ++j) {
      const MdTerm& m = H.md_terms[(size_t)j * k + i];
      const uint64_t c = H.bhat_mod_q[((size_t)i * kl + j) * 2];
      if (m.cl + ((uint64_t)m.ch << 28) + ((uint64_t)m.cs << 56) != c || m.cl >> 28 || m.ch >> 28) ++bad;
    }
    for (int64_t t = 0; t < trials; ++t) {
      uint64_t s[4] = {0, 0, 0, 0};
      uint64_t want = 0;
      for (int j = 0; j < kl; ++j) {
        const uint64_t bj = H.primes[k + j];
        const uint64_t c = H.bhat_mod_q[((size_t)i * kl + j) * 2];
        const uint64_t cq = H.bhat_mod_q[((size_t)i * kl + j) * 2 + 1];
        const int kind = t == 0 ? 1 : (int)(rng() % 4);
        uint64_t a = kind == 1 ? bj - 1 : rng() % bj;
        if (kind >= 2 && c != 0) {
          const uint64_t span = d.quirk_thr > d.quirk_lo ? d.quirk_thr - d.quirk_lo : 1;
          uint64_t target = kind == 2 ? d.quirk_lo + (rng() % 3) - 1 : d.quirk_lo + rng() % span;
          target %= p;
          const uint64_t base = (uint64_t)(((unsigned __int128)target * inv_mod(c, p)) % p);
          a = base < bj ? base : base % bj;
        }
        const uint64_t ref = ref_barrett(mulhi64(a, c), a * c, p, d.mu, (int)d.nb);
        const uint64_t exact = mul_shoup(a, c, cq, p);
        want = add_mod(want, ref, p);
        const MdTerm& m = H.md_terms[(size_t)j * k + i];
        wide_accumulate(s, a, m.cl, m.ch, m.cs);
        if (moddown_screen((uint32_t)a, (uint32_t)(a >> 32), m.w0, m.w1, m.w2) <= H.md_thr[i]) {
          wide_correct(s, ref + p - exact);
          ++screened;
        } else if (ref != exact)
          ++bad;  // a deviating product the screen let through
      }
      if (wide_reduce(s, p, d.mu, (int)d.nb) != want) ++bad;
    }
  }
  if (bad == 0 && screened == 0) return -2;
  return bad;
}

int64_t chk_moddown_accumulate(int N, int k, int kl, int prec, int64_t trials, uint64_t seed) {
  HostParams H;
  if (!build_host_params(N, k, kl, prec, &H).empty()) return -1;
  kl = H.K;  // kl = 0 selects k + 1
  std::mt19937_64 rng(seed);
  int64_t bad = 0;
  const uint32_t M28 = (1u << 28) - 1;
  for (int i = 0; i < k; ++i) {
    if (H.dprimes[i].nb > 56) continue;
    for (int64_t t = 0; t < trials; ++t) {
      uint64_t s0 = 0, s2 = 0, sm = 0;
      unsigned __int128 want = 0;
      for (int j = 0; j < kl; ++j) {
        const uint64_t bj = H.primes[k + j];
        const uint64_t a = t == 0 ? bj - 1 : (t == 1 ? (bj - 1) & ~(uint64_t)M28 : rng() % bj);
        const MdConst& m = H.md_const[(size_t)i * kl + j];
        const uint64_t c = H.bhat_mod_q[((size_t)i * kl + j) * 2];
        if ((uint64_t)m.cl + ((uint64_t)m.ch << 28) != c || m.cs != m.cl + m.ch) ++bad;
        const uint32_t al = (uint32_t)a & M28, ah = (uint32_t)(a >> 28), as = al + ah;
        s0 += (uint64_t)al * m.cl;
        s2 += (uint64_t)ah * m.ch;
        sm += (uint64_t)as * m.cs;
        want += (unsigned __int128)a * c;
      }
      const unsigned __int128 got =
          (unsigned __int128)s0 + ((unsigned __int128)(sm - s0 - s2) << 28) + ((unsigned __int128)s2 << 56);
      if (got != want) ++bad;
    }
  }
  return bad;
}

// reduce_sums_lazy against 128-bit arithmetic for every Q prime below 2^57:
// the kernel's sums of K random terms; 64 terms at the largest operands
// (a = 2^56 - 1, c = q - 1, the bound T < 2^(62+nb)); values T = m q + d
// (d in -2..2) at the quotient steps, split with s0 and s1 at the top of
// their ranges so that both floors of the estimate drop the most.  The result
// must be congruent to T and below 5q.  Returns the number of failures.
int64_t chk_moddown_reduce(int N, int k, int kl, int prec, int64_t trials, uint64_t seed) {
  HostParams H;
  if (!build_host_params(N, k, kl, prec, &H).empty()) return -1;
  kl = H.K;  // kl = 0 selects k + 1
  std::mt19937_64 rng(seed);
  int64_t bad = 0;
  const uint32_t M28 = (1u << 28) - 1;
  typedef unsigned __int128 u128;
  for (int i = 0; i < k; ++i) {
    const DevPrime& d = H.dprimes[i];
    if (d.nb > 56) continue;
    const uint64_t q = d.p, mu = H.md_red[2 * i];
    if (H.md_red[2 * i + 1] != 5 * q) ++bad;
    if (mu != (uint64_t)((((u128)1) << (63 + d.nb)) / q)) ++bad;
    auto check = [&](uint64_t s0, uint64_t s1, uint64_t s2) {
      const u128 T = (u128)s0 + ((u128)s1 << 28) + ((u128)s2 << 56);
      const uint64_t r = reduce_sums_lazy(s0, s1, s2, q, mu, (int)d.nb);
      if (r >= 5 * q || (u128)r % q != T % q) ++bad;
    };
    for (int64_t t = 0; t < trials; ++t) {
      // the kernel's sums
      uint64_t s0 = 0, s2 = 0, sm = 0;
      const int nt = t == 0 ? 64 : kl;
      for (int j = 0; j < nt; ++j) {
        const uint64_t c = t == 0 ? q - 1 : H.bhat_mod_q[((size_t)i * kl + j) * 2];
        const uint64_t a = t == 0 ? (((uint64_t)1 << 56) - 1) : rng() % H.primes[k + j];
        const uint32_t cl = (uint32_t)(c & M28), ch = (uint32_t)(c >> 28), cs = cl + ch;
        const uint32_t al = (uint32_t)a & M28, ah = (uint32_t)(a >> 28), as = al + ah;
        s0 += (uint64_t)al * cl;
        s2 += (uint64_t)ah * ch;
        sm += (uint64_t)as * cs;
      }
      check(s0, sm - s0 - s2, s2);
      // quotient steps, the low sums at the top of their ranges
      const u128 tmax = ((u128)1) << (61 + d.nb);
      const u128 m = ((u128)rng() << 64 | rng()) % (tmax / q);
      for (int dd = -2; dd <= 2; ++dd) {
        const u128 T = m * q + (u128)(__int128)dd + 2 * (u128)q;
        const uint64_t x0 = ((uint64_t)1 << 61) + (uint64_t)(T & M28) - (((uint64_t)1 << 61) & ~(uint64_t)M28);
        // T = x0 + x1 2^28 + x2 2^56 with x0 near 2^61 and x1 near 2^62
        const u128 rest = (T - x0) >> 28;  // T - x0 is a multiple of 2^28
        const uint64_t x1 = (uint64_t)(rest & ((((u128)1) << 28) - 1)) + (((uint64_t)1 << 62) & ~(uint64_t)M28);
        const u128 rest2 = rest - x1;
        if (T < x0 || rest < x1 || (rest2 & ((((u128)1) << 28) - 1)) != 0) {
          check((uint64_t)(T & M28), (uint64_t)((T >> 28) & M28), (uint64_t)(T >> 56));
          continue;
        }
        check(x0, x1, (uint64_t)(rest2 >> 28));
      }
    }
  }
  return bad;
}

// ModDown's finish with the lazy sums against the textbook formula, for
// every Q prime below 2^57 and level l: ((x - T) B^-1 + add - top) q_l^-1
// mod q_i, computed as (x + 5q - r) (B^-1 q_l^-1) + (q + add - top) q_l^-1 from
// r = reduce_sums_lazy (the rescale path), and (x + 5q - r) B^-1 + add without
// the rescale; then as the kernel's narrow limbs do, with z = q + add - top
// (or z = add) added to the sums as the term z (-B mod q) and one Shoup
// product by B^-1 q_l^-1 (or B^-1).  Returns the number of mismatches.
int64_t chk_moddown_finish(int N, int k, int kl, int prec, int64_t trials, uint64_t seed) {
  HostParams H;
  if (!build_host_params(N, k, kl, prec, &H).empty()) return -1;
  kl = H.K;  // kl = 0 selects k + 1
  std::mt19937_64 rng(seed);
  int64_t bad = 0;
  typedef unsigned __int128 u128;
  for (int l = 1; l < k; ++l) {
    for (int i = 0; i < l; ++i) {
      const DevPrime& d = H.dprimes[i];
      if (d.nb > 56) continue;
      const uint64_t q = d.p, ql = H.primes[l];
      const uint64_t ib = H.inv_b_mod_q[2 * i], ibq = H.inv_b_mod_q[2 * i + 1];
      const uint64_t iql = H.inv_ql_mod_qi[((size_t)l * k + i) * 2], iqlq = H.inv_ql_mod_qi[((size_t)l * k + i) * 2 + 1];
      const uint64_t f = H.md_fin[((size_t)l * k + i) * 2], fq = H.md_fin[((size_t)l * k + i) * 2 + 1];
      for (int64_t t = 0; t < trials; ++t) {
        const uint64_t x = t == 0 ? q - 1 : rng() % q, add = t == 1 ? q - 1 : rng() % q;
        const uint64_t top = t == 2 ? ql - 1 : rng() % ql;
        const uint64_t s0 = rng() >> 3, s1 = rng() >> 2, s2 = (rng() >> 2) & ((((uint64_t)1) << (d.nb + 5)) - 1);
        const u128 T = (u128)s0 + ((u128)s1 << 28) + ((u128)s2 << 56);
        const uint64_t r = reduce_sums_lazy(s0, s1, s2, q, H.md_red[2 * i], (int)d.nb);
        const uint64_t y = x + H.md_red[2 * i + 1] - r;
        const uint64_t tm = (uint64_t)(T % q);
        const uint64_t v = (uint64_t)(((u128)((x + q - tm) % q) * ib) % q);
        const uint64_t want_r = (uint64_t)(((u128)((v + add + q - top % q) % q) * iql) % q);
        const uint64_t got_r = add_mod(mul_shoup(y, f, fq, q), mul_shoup(q + add - top, iql, iqlq, q), q);
        const uint64_t want = (v + add) % q;
        const uint64_t got = add_mod(mul_shoup(y, ib, ibq, q), add, q);
        if (got_r != want_r || got != want) ++bad;
        // the kernel's narrow limbs: z = q + add - top (rescale) or z = add
        // joins the sums as the term z (-B mod q), then one Shoup product
        const uint32_t* e = &H.md_negb[4 * i];
        const uint64_t negb = (uint64_t)e[0] + ((uint64_t)e[1] << 28);
        if (e[2] != e[0] + e[1] || e[0] >> 28 || negb != (q - H.b_mod_q[2 * i]) % q) ++bad;
        auto folded = [&](uint64_t z, uint64_t w, uint64_t wq) {
          const uint32_t M28 = (1u << 28) - 1;
          const uint32_t zl = (uint32_t)z & M28, zh = (uint32_t)(z >> 28);
          const uint64_t f0 = s0 + (uint64_t)zl * e[0], f2 = s2 + (uint64_t)zh * e[1];
          const uint64_t f1 = s1 + (uint64_t)zl * e[1] + (uint64_t)zh * e[0];
          const uint64_t rf = reduce_sums_lazy(f0, f1, f2, q, H.md_red[2 * i], (int)d.nb);
          return mul_shoup(x + H.md_red[2 * i + 1] - rf, w, wq, q);
        };
        if (folded(q + add - top, f, fq) != want_r || folded(add, ib, ibq) != want) ++bad;
      }
    }
  }
  return bad;
}

// ModUp's tables and sums against 128-bit arithmetic, for every level l and
// special prime b_j: the mu_terms entries (the 28-bit halves of
// (Q_l/q_i) mod b_j at the pass position of j, row k: of 2^56 (Q_l/q_0) mod
// b_j) and md_red of the special primes; then the kernel's sums -- q0's value
// split at bit 56, the other terms at bit 28 -- reduced by reduce_sums_lazy
// and brought below b_j with three conditional subtractions, against
// sum_i v_i (Q_l/q_i mod b_j) mod b_j, with v_i = q_i - 1 (the largest values)
// in trial 0 and random residues of q_i otherwise.  Returns the number of
// failures.
int64_t chk_modup_sums(int N, int k, int kl, int prec, int64_t trials, uint64_t seed) {
  HostParams H;
  if (!build_host_params(N, k, kl, prec, &H).empty()) return -1;
  kl = H.K;  // kl = 0 selects k + 1
  std::mt19937_64 rng(seed);
  typedef unsigned __int128 u128;
  const uint32_t M28 = (1u << 28) - 1;
  const int kp = (kl + 15) & ~15;
  int64_t bad = 0;
  auto entry = [&](int l, int row, int j) {
    const int t = j / 4, pos = (t / 4) * 16 + 4 * (j % 4) + t % 4;
    return &H.mu_terms[(((size_t)l * (k + 1) + row) * kp + pos) * 4];
  };
  for (int j = 0; j < kl; ++j) {
    const DevPrime& d = H.dprimes[k + j];
    const uint64_t b = d.p;
    if (H.md_red[2 * (k + j) + 1] != 5 * b) ++bad;
    if (H.md_red[2 * (k + j)] != (uint64_t)((((u128)1) << (63 + d.nb)) / b)) ++bad;
    for (int l = 0; l < k; ++l) {
      for (int i = 0; i <= l; ++i) {
        const uint64_t w = H.qhat_mod_b[((size_t)l * kl + j) * k + i];
        const uint32_t* e = entry(l, i, j);
        if (e[0] != (uint32_t)(w & M28) || e[1] != (uint32_t)(w >> 28) || e[2] != e[0] + e[1]) ++bad;
      }
      const uint64_t w0 = H.qhat_mod_b[((size_t)l * kl + j) * k];
      const uint64_t wh = (uint64_t)(((u128)w0 << 56) % b);
      const uint32_t* eh = entry(l, k, j);
      if (eh[0] != (uint32_t)(wh & M28) || eh[1] != (uint32_t)(wh >> 28) || eh[2] != eh[0] + eh[1]) ++bad;
      for (int64_t t = 0; t < trials; ++t) {
        uint64_t s0 = 0, s2 = 0, sm = 0;
        u128 want = 0;
        for (int i = 0; i <= l; ++i) {
          const uint64_t qi = H.primes[i];
          const uint64_t v = t == 0 ? qi - 1 : rng() % qi;
          want = (want + (u128)(v % b) * H.qhat_mod_b[((size_t)l * kl + j) * k + i]) % b;
          const uint32_t* e = entry(l, i, j);
          // the kernel reduces v mod b_j first only where q_i > b_j (i >= 1)
          uint64_t vv = (i > 0 && qi > b) ? v % b : v;
          if (i == 0) {
            const uint32_t hi = (uint32_t)(v >> 56);
            vv = v & ((((uint64_t)1) << 56) - 1);
            s0 += (uint64_t)hi * eh[0];
            sm += (uint64_t)hi * eh[2];
          }
          const uint32_t vl = (uint32_t)vv & M28, vh = (uint32_t)(vv >> 28);
          s0 += (uint64_t)vl * e[0];
          s2 += (uint64_t)vh * e[1];
          sm += (uint64_t)(vl + vh) * e[2];
        }
        uint64_t r = reduce_sums_lazy(s0, sm - s0 - s2, s2, b, H.md_red[2 * (k + j)], (int)d.nb);
        for (uint64_t m = 4 * b; m >= b; m >>= 1) r = r >= m ? r - m : r;
        if ((u128)r != want) ++bad;
      }
    }
  }
  return bad;
}

// The product's host tables (same numbering as aoa_params_table), built
// without a GPU.  Returns the number of words (out may be null).
// The row-ordered twiddle tables against the index each row-pass thread
// uses: for every two-pass split N = N1 x S, direction, prime, row, stage
// 4 <= r < log2 S of the row pass, thread sh < S/16 and m < 2^r / (S/16), the
// kernel's entry 2^(g0+r) + row 2^r + m (S/16) + sh of the reordered table
// must be the natural table's entry for thread sh's m-th butterfly, which
// holds 16 consecutive values: 2^(g0+r) + row 2^r + sh 2^r/(S/16) + m.  Every
// other entry is unchanged.  Returns the number of entries compared, or -1
// on the first mismatch, -2 when a split that must have a table has none.
int64_t chk_row_twiddles(int N, int k, int kl, int prec) {
  HostParams H;
  if (!build_host_params(N, k, kl, prec, &H).empty()) return -3;
  const int R = H.k + H.K;
  int64_t n = 0;
  for (int S = 64; S <= 256; S *= 2) {
    const int N1 = N / S;
    const bool valid = N1 * S == N && (N1 == 64 || N1 == 128 || N1 == 256);
    int g0 = 0, logs = 0;
    while ((1 << g0) < N1) ++g0;
    while ((1 << logs) < S) ++logs;
    const int ts = S / 16;
    for (int dir = 0; dir < 2; ++dir) {
      const std::vector<uint64_t>& T = dir ? H.tw_inv : H.tw_fwd;
      std::vector<uint64_t> TR;
      row_twiddle_order(T, N, R, S, &TR);
      if (!valid) {
        if (!TR.empty()) return -2;
        continue;
      }
      if (TR.size() != T.size()) return -2;
      std::vector<char> seen((size_t)N, 0);
      for (int pr = 0; pr < R; ++pr) {
        const uint64_t* t = &T[(size_t)pr * N * 2];
        const uint64_t* tr = &TR[(size_t)pr * N * 2];
        std::fill(seen.begin(), seen.end(), 0);
        for (int r = 4; r < logs; ++r)
          for (int row = 0; row < N1; ++row)
            for (int sh = 0; sh < ts; ++sh)
              for (int m = 0; m < (1 << r) / ts; ++m) {
                const size_t base = ((size_t)1 << (g0 + r)) + ((size_t)row << r);
                const size_t kern = base + (size_t)m * ts + sh, nat = base + (size_t)sh * ((1 << r) / ts) + m;
                if (tr[2 * kern] != t[2 * nat] || tr[2 * kern + 1] != t[2 * nat + 1]) return -1;
                seen[kern] = 1;
                ++n;
              }
        for (int i = 0; i < N; ++i)
          if (!seen[i] && (tr[2 * i] != t[2 * i] || tr[2 * i + 1] != t[2 * i + 1])) return -1;
      }
    }
  }
  return n;
}

int64_t chk_table(int N, int k, int kl, int prec, int which, uint64_t* out) {
  HostParams H;
  if (!build_host_params(N, k, kl, prec, &H).empty()) return -1;
  kl = H.K;  // kl = 0 selects k + 1
  std::vector<uint64_t> v;
  auto take = [&](const std::vector<uint64_t>& src, int stride) {
    for (size_t i = 0; i < src.size(); i += stride) v.push_back(src[i]);
  };
  switch (which) {
    case 0: take(H.tw_fwd, 2); break;
    case 1: take(H.tw_inv, 2); break;
    case 2: take(H.inv_qhat, 2); break;
    case 3: take(H.qhat_mod_b, 1); break;
    case 4: take(H.inv_bhat, 2); break;
    case 5: take(H.bhat_mod_q, 2); break;
    case 6: take(H.inv_b_mod_q, 2); break;
    case 7: take(H.b_mod_q, 2); break;
    case 8: take(H.inv_ql_mod_qi, 2); break;
    case 9:
      for (auto& d : H.dprimes) v.push_back(d.mu);
      break;
    case 10: v = H.primes; break;
    case 15: take(H.bscale, 2); break;
    case 11:  // encoder roots exp(2 pi i t / 2N) as (re, im) bit patterns
      v.resize(H.enc_w.size());
      std::memcpy(v.data(), H.enc_w.data(), v.size() * 8);
      break;
    case 12:
      for (uint32_t g : H.rot_group) v.push_back(g);
      break;
    case 13: v = H.gauss_cdt; break;
    case 14:
      for (const MdConst& m : H.md_const)
        for (uint32_t x : {m.cl, m.ch, m.cs, m.w0, m.w1, m.w2, m.thr, m.pad}) v.push_back(x);
      break;
    default: return -1;
  }
  if (out)
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
  return (int64_t)v.size();
}

}  // extern "C"
