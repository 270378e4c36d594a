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
// out[2] = products that took the literal path.
int chk_moddown_terms(int N, int k, int kl, int prec, int64_t trials, uint64_t seed, int64_t* out) {
  HostParams H;
  if (!build_host_params(N, k, kl, prec, &H).empty()) return -1;
  std::mt19937_64 rng(seed);
  int64_t bad = 0, quirks = 0, slow = 0;
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
        a = base + (mmax ? (rng() % (mmax + 1)) : 0) * p;
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
    }
  }
  out[0] = bad;
  out[1] = quirks;
  out[2] = slow;
  return 0;
}

// The product's host tables (same numbering as aoa_params_table), built
// without a GPU.  Returns the number of words (out may be null).
int64_t chk_table(int N, int k, int kl, int prec, int which, uint64_t* out) {
  HostParams H;
  if (!build_host_params(N, k, kl, prec, &H).empty()) return -1;
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
    case 11:  // encoder roots exp(2 pi i t / 2N) as (re, im) bit patterns
      v.resize(H.enc_w.size());
      std::memcpy(v.data(), H.enc_w.data(), v.size() * 8);
      break;
    case 12:
      for (uint32_t g : H.rot_group) v.push_back(g);
      break;
    case 13: v = H.gauss_cdt; break;
    default: return -1;
  }
  if (out)
    for (size_t i = 0; i < v.size(); ++i) out[i] = v[i];
  return (int64_t)v.size();
}

}  // extern "C"
