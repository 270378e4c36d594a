// Host-side construction of a parameter set (primes, twiddles, RNS constants).
#include "params.hpp"

#include <cmath>
#include <cstring>

namespace aoa {

#include "prime_table.inc"

uint32_t bit_reverse(uint32_t x, int bits) {
  uint32_t r = 0;
  for (int i = 0; i < bits; ++i) {
    r = (r << 1) | (x & 1u);
    x >>= 1;
  }
  return r;
}

uint64_t inv_mod(uint64_t a, uint64_t p) {
  // extended Euclid on signed 128-bit integers
  __int128 t = 0, nt = 1, r = p, nr = a % p;
  while (nr != 0) {
    __int128 q = r / nr, tmp;
    tmp = t - q * nt; t = nt; nt = tmp;
    tmp = r - q * nr; r = nr; nr = tmp;
  }
  if (t < 0) t += p;
  return (uint64_t)t;
}

static uint64_t mulm(uint64_t a, uint64_t b, uint64_t p) {
  return (uint64_t)(((unsigned __int128)a * b) % p);
}

static uint64_t powm(uint64_t a, uint64_t e, uint64_t p) {
  uint64_t r = 1 % p, s = a % p;
  while (e) {
    if (e & 1) r = mulm(r, s, p);
    s = mulm(s, s, p);
    e >>= 1;
  }
  return r;
}

// Prime layout of the reference (aoa-ntt/src/cuda/manager.cu:426-464): the
// 63-bit prime first, then k-1 primes of `prec` bits; the special base B is
// 55-bit for prec 48/52 and the continuation of the same bucket for 45/55.
static std::string select_primes(int k, int kl, int prec, std::vector<aoa_prime_entry>* q,
                                 std::vector<aoa_prime_entry>* b) {
  const aoa_prime_entry* qb = nullptr;
  int qn = 0;
  const aoa_prime_entry* bb = nullptr;
  int bn = 0, boff = 0;
  switch (prec) {
    case 45: qb = AOA_BUCKET_45; qn = 100; bb = AOA_BUCKET_45; bn = 100; boff = k - 1; break;
    case 48: qb = AOA_BUCKET_48; qn = 100; bb = AOA_BUCKET_55; bn = 500; boff = 0; break;
    case 52: qb = AOA_BUCKET_52; qn = 30; bb = AOA_BUCKET_55; bn = 500; boff = 0; break;
    case 55: qb = AOA_BUCKET_55; qn = 500; bb = AOA_BUCKET_55; bn = 500; boff = k - 1; break;
    default: return "We don't support this precision.";
  }
  if (k < 1 || kl < 1) return "k and kl must be positive";
  // ModUp / ModDown keep one 2 KiB LDS column per limb per workgroup.
  if (k > 64 || kl > 64) return "at most 64 Q limbs and 64 special limbs are supported";
  if (k - 1 > qn) return "not enough primes of this precision for k";
  if (boff + kl > bn) return "not enough special primes for kl";
  q->clear();
  b->clear();
  q->push_back(AOA_BUCKET_63[0]);
  for (int i = 0; i < k - 1; ++i) q->push_back(qb[i]);
  for (int j = 0; j < kl; ++j) b->push_back(bb[boff + j]);
  // ModUp splits its operands (< b_j) at bit 28 and ModDown keeps lazy sums
  // of K terms below 2^57 each: both need special primes below 2^56.
  uint64_t bmax = 0;
  for (auto& e : *b) {
    if (e.p >> 56) return "special primes must be below 2^56";
    bmax = e.p > bmax ? e.p : bmax;
  }
  // ModDown's lazy sums of limb i hold K products a c (a < b_j, c < q_i), one
  // folded term below 2^57 q_i and at most K screen corrections below 2 q_i:
  // reduce_sums_lazy needs their total below 2^(62+nb), which
  // K b_max + 2^57 + 2K <= 2^62 ensures.
  typedef unsigned __int128 u128;
  if ((u128)kl * bmax + ((u128)1 << 57) + 2 * (u128)kl > ((u128)1 << 62))
    return "special primes too large for ModDown's lazy sums";
  // ModUp's lazy sums for b_j hold at most k - 1 products below b_j^2 and
  // q0's two terms below (2^56 + 2^7) b_j: (k - 1) b_max + 2^56 + 2^7 <= 2^62
  // keeps them below 2^(62+nb) for reduce_sums_lazy.
  if ((u128)(k > 1 ? k - 1 : 0) * bmax + ((u128)1 << 56) + 128 > ((u128)1 << 62))
    return "special primes too large for ModUp's lazy sums";
  // ModDown's exact sums for a Q prime wider than 2^56 (q0) total below
  // q (K b_max + 2K) (K products and K screen corrections below 2q); one
  // Barrett reduction needs that below q 2^nb.
  for (const aoa_prime_entry& e : *q) {
    int nb = 0;
    while (nb < 64 && (e.p >> nb)) ++nb;
    if (nb > 56 && (u128)kl * bmax + 2 * (u128)kl > ((u128)1 << nb))
      return "special primes too large for ModDown's sums of a wide Q prime";
  }
  // ModDown's rescale forms q_i + add_i - top with top < q_l for i < l, which
  // needs q_l < q_i: the Q primes decrease (q0 first, then the bucket in
  // decreasing order).
  for (size_t i = 1; i < q->size(); ++i)
    if ((*q)[i].p >= (*q)[i - 1].p) return "Q primes must decrease";
  return "";
}

std::string select_rns_primes(int k, int kl, int prec, std::vector<uint64_t>* primes) {
  std::vector<aoa_prime_entry> q, b;
  std::string err = select_primes(k, kl, prec, &q, &b);
  if (!err.empty()) return err;
  primes->clear();
  for (auto& e : q) primes->push_back(e.p);
  for (auto& e : b) primes->push_back(e.p);
  return "";
}

// The words (w0, w1, w2) of floor(c 2^96 / p) for c < p < 2^63, by long
// division one 32-bit word at a time.
void moddown_screen_words(uint64_t c, uint64_t p, uint32_t w[3]) {
  unsigned __int128 r = c;
  for (int x = 2; x >= 0; --x) {
    r <<= 32;
    w[x] = (uint32_t)(r / p);
    r %= p;
  }
}

// Screen threshold for remainders below thr: floor(thr 2^32 / p) + 2, or
// 2^32 - 1 (every product screened in) when that does not fit.
uint32_t moddown_screen_thr(uint64_t thr, uint64_t p) {
  const unsigned __int128 t = (((unsigned __int128)thr << 32) / p) + 2;
  return t > 0xFFFFFFFFu ? 0xFFFFFFFFu : (uint32_t)t;
}

void row_twiddle_order(const std::vector<uint64_t>& tw, int N, int R, int S, std::vector<uint64_t>* out) {
  out->clear();
  const int N1 = S > 0 ? N / S : 0;
  auto ok = [](int v) { return v == 64 || v == 128 || v == 256; };
  if (!ok(S) || !ok(N1) || N1 * S != N || tw.size() != (size_t)R * N * 2) return;
  int g0 = 0, logs = 0;
  while ((1 << g0) < N1) ++g0;
  while ((1 << logs) < S) ++logs;
  const int ts = S / 16;
  out->assign(tw.begin(), tw.end());
  for (int pr = 0; pr < R; ++pr) {
    const uint64_t* src = &tw[(size_t)pr * N * 2];
    uint64_t* dst = &(*out)[(size_t)pr * N * 2];
    for (int r = 4; r < logs; ++r) {
      const int blk = 1 << r, c = blk / ts;
      for (int b = 0; b < N1; ++b) {
        const size_t base = ((size_t)1 << (g0 + r)) + (size_t)b * blk;
        for (int sh = 0; sh < ts; ++sh)
          for (int m = 0; m < c; ++m) {
            const size_t from = base + (size_t)sh * c + m, to = base + (size_t)m * ts + sh;
            dst[2 * to] = src[2 * from];
            dst[2 * to + 1] = src[2 * from + 1];
          }
      }
    }
  }
}

std::string build_host_params(int N, int k, int kl, int prec, HostParams* out) {
  if (N < 2048 || N > 65536 || (N & (N - 1)) != 0) return "NTT is not working for this N";
  if (kl <= 0) kl = k + 1;
  std::vector<aoa_prime_entry> qe, be;
  std::string err = select_primes(k, kl, prec, &qe, &be);
  if (!err.empty()) return err;

  HostParams& P = *out;
  P = HostParams();
  P.N = N;
  P.logN = 0;
  while ((1 << P.logN) < N) ++P.logN;
  P.k = k;
  P.K = kl;
  P.prec = prec;
  for (auto& e : qe) { P.primes.push_back(e.p); P.roots.push_back(e.g); }
  for (auto& e : be) { P.primes.push_back(e.p); P.roots.push_back(e.g); }
  const int R = k + kl;

  for (int r = 0; r < R; ++r) {
    uint64_t p = P.primes[r];
    if ((p - 1) % (2 * (uint64_t)N) != 0) return "prime does not support this ring dimension";
    DevPrime d;
    std::memset(&d, 0, sizeof(d));
    d.p = p;
    d.nb = (uint32_t)bitlen64(p);
    d.mu = (uint64_t)((((unsigned __int128)1) << (2 * d.nb)) / p);
    d.mu2 = d.mu << (63 - d.nb);
    d.ninv = inv_mod((uint64_t)N % p, p);
    d.ninv_q = shoup_quot(d.ninv, p);
    d.eps = (((uint64_t)1) << d.nb) - p;
    d.r64 = (uint64_t)((((unsigned __int128)1) << 64) % p);
    d.r64_q = shoup_quot(d.r64, p);
    P.dprimes.push_back(d);
  }
  for (int j = 0; j < kl; ++j)
    for (int i = 1; i < k; ++i)
      if (P.primes[i] > P.primes[k + j]) P.dprimes[k + j].modup_reduce = 1;

  // Twiddles: psi = g^((p-1)/2N); forward table psi^brev(i), inverse psi^-brev(i).
  P.tw_fwd.assign((size_t)R * N * 2, 0);
  P.tw_inv.assign((size_t)R * N * 2, 0);
  std::vector<uint64_t> pw(N), ipw(N);
  for (int r = 0; r < R; ++r) {
    uint64_t p = P.primes[r];
    uint64_t psi = powm(P.roots[r], (p - 1) / (2 * (uint64_t)N), p);
    uint64_t ipsi = inv_mod(psi, p);
    pw[0] = 1;
    ipw[0] = 1;
    for (int i = 1; i < N; ++i) {
      pw[i] = mulm(pw[i - 1], psi, p);
      ipw[i] = mulm(ipw[i - 1], ipsi, p);
    }
    uint64_t* tf = &P.tw_fwd[(size_t)r * N * 2];
    uint64_t* ti = &P.tw_inv[(size_t)r * N * 2];
    for (int i = 0; i < N; ++i) {
      uint32_t e = bit_reverse((uint32_t)i, P.logN);
      tf[2 * i] = pw[e];
      tf[2 * i + 1] = shoup_quot(pw[e], p);
      ti[2 * i] = ipw[e];
      ti[2 * i + 1] = shoup_quot(ipw[e], p);
    }
    DevPrime& d = P.dprimes[r];
    d.w1ninv = mulm(ti[2], d.ninv, p);
    d.w1ninv_q = shoup_quot(d.w1ninv, p);
  }

  // RNS constants (aoa-ntt/src/cuda/manager.cu:566-759), as residues of
  // products of primes; no big integers are needed.
  const std::vector<uint64_t> Qp(P.primes.begin(), P.primes.begin() + k);
  const std::vector<uint64_t> Bp(P.primes.begin() + k, P.primes.end());
  P.inv_qhat.assign((size_t)k * k * 2, 0);
  P.qhat_mod_b.assign((size_t)k * kl * k, 0);
  P.inv_ql_mod_qi.assign((size_t)k * k * 2, 0);
  for (int l = 0; l < k; ++l) {
    for (int i = 0; i <= l; ++i) {
      // (Q_l / q_i) mod q_i, Q_l = q_0 ... q_l
      uint64_t qi = Qp[i];
      uint64_t prod = 1 % qi;
      for (int m = 0; m <= l; ++m)
        if (m != i) prod = mulm(prod, Qp[m] % qi, qi);
      uint64_t inv = inv_mod(prod, qi);
      P.inv_qhat[((size_t)l * k + i) * 2] = inv;
      P.inv_qhat[((size_t)l * k + i) * 2 + 1] = shoup_quot(inv, qi);
      for (int j = 0; j < kl; ++j) {
        uint64_t bj = Bp[j];
        uint64_t pr = 1 % bj;
        for (int m = 0; m <= l; ++m)
          if (m != i) pr = mulm(pr, Qp[m] % bj, bj);
        P.qhat_mod_b[((size_t)l * kl + j) * k + i] = pr;
      }
    }
    for (int i = 0; i < k; ++i) {
      if (i == l) continue;
      uint64_t v = inv_mod(Qp[l] % Qp[i], Qp[i]);
      P.inv_ql_mod_qi[((size_t)l * k + i) * 2] = v;
      P.inv_ql_mod_qi[((size_t)l * k + i) * 2 + 1] = shoup_quot(v, Qp[i]);
    }
  }
  // ModUp's constants (Q_l/q_i) mod b_j split at bit 28, in the order the
  // kernel reads them: one block of kp = K rounded up to 16 entries per (l, i),
  // special limb j = g + 4t (g = j % 4) at position (t / 4) 16 + 4 g + t % 4,
  // so that the limbs of one wave's pass are consecutive entries.  Row i = k
  // holds (2^56 (Q_l/q_0)) mod b_j: the kernel splits q0's value v at bit 56
  // and multiplies its top bits by it.
  const int kp = (kl + 15) & ~15;
  P.mu_terms.assign((size_t)k * (k + 1) * kp * 4, 0);
  auto put_mu = [&](int l, int row, int j, uint64_t w) {
    const int t = j / 4, pos = (t / 4) * 16 + 4 * (j % 4) + t % 4;
    uint32_t* e = &P.mu_terms[(((size_t)l * (k + 1) + row) * kp + pos) * 4];
    e[0] = (uint32_t)(w & ((1u << 28) - 1));
    e[1] = (uint32_t)(w >> 28);
    e[2] = e[0] + e[1];
  };
  for (int l = 0; l < k; ++l)
    for (int j = 0; j < kl; ++j) {
      for (int i = 0; i <= l; ++i) put_mu(l, i, j, P.qhat_mod_b[((size_t)l * kl + j) * k + i]);
      const uint64_t w0 = P.qhat_mod_b[((size_t)l * kl + j) * k];
      put_mu(l, k, j, (uint64_t)(((unsigned __int128)w0 << 56) % P.primes[k + j]));
    }
  P.inv_bhat.assign((size_t)kl * 2, 0);
  for (int j = 0; j < kl; ++j) {
    uint64_t bj = Bp[j];
    uint64_t pr = 1 % bj;
    for (int m = 0; m < kl; ++m)
      if (m != j) pr = mulm(pr, Bp[m] % bj, bj);
    uint64_t inv = inv_mod(pr, bj);
    P.inv_bhat[2 * j] = inv;
    P.inv_bhat[2 * j + 1] = shoup_quot(inv, bj);
  }
  P.bscale.assign((size_t)kl * 4, 0);
  for (int j = 0; j < kl; ++j) {
    const uint64_t bj = Bp[j], inv = P.inv_bhat[2 * j];
    const DevPrime& d = P.dprimes[k + j];
    const uint64_t a = mulm(d.ninv, inv, bj), b = mulm(d.w1ninv, inv, bj);
    P.bscale[4 * j] = a;
    P.bscale[4 * j + 1] = shoup_quot(a, bj);
    P.bscale[4 * j + 2] = b;
    P.bscale[4 * j + 3] = shoup_quot(b, bj);
  }
  P.bhat_mod_q.assign((size_t)k * kl * 2, 0);
  P.md_const.assign((size_t)k * kl, MdConst{});
  P.md_terms.assign((size_t)kl * k, MdTerm{});
  P.md_thr.assign((size_t)k, 0);
  P.md_red.assign((size_t)(k + kl) * 2, 0);
  P.md_fin.assign((size_t)k * k * 2, 0);
  P.md_negb.assign((size_t)k * 4, 0);
  P.inv_b_mod_q.assign((size_t)k * 2, 0);
  P.b_mod_q.assign((size_t)k * 2, 0);
  uint64_t bmax = 0;
  for (uint64_t b : Bp) bmax = b > bmax ? b : bmax;
  for (int i = 0; i < k; ++i) {
    uint64_t qi = Qp[i];
    uint64_t bfull = 1 % qi;
    for (int m = 0; m < kl; ++m) bfull = mulm(bfull, Bp[m] % qi, qi);
    for (int j = 0; j < kl; ++j) {
      uint64_t pr = 1 % qi;
      for (int m = 0; m < kl; ++m)
        if (m != j) pr = mulm(pr, Bp[m] % qi, qi);
      P.bhat_mod_q[((size_t)i * kl + j) * 2] = pr;
      P.bhat_mod_q[((size_t)i * kl + j) * 2 + 1] = shoup_quot(pr, qi);
    }
    // Bound on the reference's Barrett quotient error delta = floor(x/q) - qhat
    // for x = a c < bmax q_i, where qhat = floor(X1 MU / 2^64) with
    // X1 = floor(x / 2^s), s = 2nb - 64 and MU = (2^(2nb) - rho) / q_i,
    // rho = 2^(2nb) mod q_i.  Then x/q - X1 MU / 2^64 = (x mod 2^s)/q +
    // X1 rho / (q 2^64) < 1 + X1 rho / (q 2^64) (as 2^s < q), so
    // delta < 2 + X1 rho / (q 2^64) and delta <= floor(X1max rho / (q 2^64)) + 2.
    // The reference result deviates from the exact remainder only for
    // eps <= r < delta * eps, which needs delta * eps < 2^nb; otherwise every
    // product takes the literal formula.
    DevPrime& d = P.dprimes[i];
    unsigned __int128 dmax;
    if (d.nb >= 33) {
      const unsigned __int128 x1max = ((unsigned __int128)(bmax - 1) * (qi - 1)) >> (2 * d.nb - 64);
      const unsigned __int128 rho = (((unsigned __int128)1) << (2 * d.nb)) % qi;
      dmax = x1max * rho / (((unsigned __int128)qi) << 64) + 2;
    } else {
      dmax = ~(unsigned __int128)0 >> 1;
    }
    unsigned __int128 thr = dmax > ((unsigned __int128)1 << 64) ? ~(unsigned __int128)0 >> 1 : dmax * d.eps;
    if (thr < ((unsigned __int128)1 << (d.nb - 1))) {
      d.quirk_lo = d.eps;
      d.quirk_thr = (uint64_t)thr;
    } else {
      d.quirk_lo = 0;
      d.quirk_thr = ~(uint64_t)0;
    }
    // Karatsuba split and screen words of every c, for Q primes below 2^56.
    // One screen threshold per prime: a bound per (i, j) would pass fewer
    // products but costs the kernel's inner loop one more scalar operand per
    // product.  A wider prime (the 63-bit q0) gets c in three pieces, at bits
    // 28 and 56, and the screen words in the device table only.
    const uint64_t wthr = d.quirk_thr < qi ? d.quirk_thr : qi;
    const uint32_t sthr = moddown_screen_thr(wthr, qi);
    if (d.nb > 56) {
      for (int j = 0; j < kl; ++j) {
        const uint64_t c = P.bhat_mod_q[((size_t)i * kl + j) * 2];
        uint32_t w[3];
        moddown_screen_words(c, qi, w);
        P.md_terms[(size_t)j * k + i] = MdTerm{(uint32_t)(c & ((1u << 28) - 1)),
                                               (uint32_t)((c >> 28) & ((1u << 28) - 1)), (uint32_t)(c >> 56), w[0],
                                               w[1], w[2]};
      }
      P.md_thr[i] = sthr;
    } else {
      for (int j = 0; j < kl; ++j) {
        const uint64_t c = P.bhat_mod_q[((size_t)i * kl + j) * 2];
        MdConst& m = P.md_const[(size_t)i * kl + j];
        uint32_t w[3];
        moddown_screen_words(c, qi, w);
        m.cl = (uint32_t)(c & ((1u << 28) - 1));
        m.ch = (uint32_t)(c >> 28);
        m.cs = m.cl + m.ch;
        m.w0 = w[0];
        m.w1 = w[1];
        m.w2 = w[2];
        m.thr = sthr;
        m.pad = 2;  // moddown_screen's bias 2^33, as its high word
        P.md_terms[(size_t)j * k + i] = MdTerm{m.cl, m.ch, m.cs, m.w0, m.w1, m.w2};
      }
      P.md_thr[i] = sthr;
    }
    uint64_t inv = inv_mod(bfull, qi);
    P.inv_b_mod_q[2 * i] = inv;
    P.inv_b_mod_q[2 * i + 1] = shoup_quot(inv, qi);
    if (d.nb >= 33 && d.nb <= 56) {
      P.md_red[2 * i] = (uint64_t)((((unsigned __int128)1) << (63 + d.nb)) / qi);
      P.md_red[2 * i + 1] = 5 * qi;
    } else {
      P.md_red[2 * i] = 0;
      P.md_red[2 * i + 1] = qi;
    }
    for (int l = 0; l < k; ++l) {
      const uint64_t f = mulm(inv, P.inv_ql_mod_qi[((size_t)l * k + i) * 2], qi);
      P.md_fin[((size_t)l * k + i) * 2] = f;
      P.md_fin[((size_t)l * k + i) * 2 + 1] = shoup_quot(f, qi);
    }
    if (d.nb <= 56) {
      const uint64_t nb_ = (qi - bfull) % qi;
      P.md_negb[4 * i] = (uint32_t)(nb_ & ((1u << 28) - 1));
      P.md_negb[4 * i + 1] = (uint32_t)(nb_ >> 28);
      P.md_negb[4 * i + 2] = P.md_negb[4 * i] + P.md_negb[4 * i + 1];
    }
    P.b_mod_q[2 * i] = bfull;
    P.b_mod_q[2 * i + 1] = shoup_quot(bfull, qi);
  }
  // ModUp reduces its lazy sums per special prime (below 2^56, so nb <= 56)
  for (int j = 0; j < kl; ++j) {
    const DevPrime& d = P.dprimes[k + j];
    P.md_red[2 * (k + j)] = (uint64_t)((((unsigned __int128)1) << (63 + d.nb)) / d.p);
    P.md_red[2 * (k + j) + 1] = 5 * d.p;
  }
  // Encoder tables: the roots exp(2 pi i t / 2N) and the rotation group
  // 5^j mod 2N (aoa-ntt/src/ckks/encoder.cu:36-48).  The reference's ksiPows
  // come from cos(angle) and sin(angle) in host code that an optimising g++
  // turns into one glibc sincos call, whose results differ from separate sin
  // and cos in the last bit for a few angles; sincos is called here through a
  // volatile pointer so that no compiler substitutes anything else.
  {
    void (*volatile sincos_fn)(double, double*, double*) = ::sincos;
    const size_t M = 2 * (size_t)N;
    P.enc_w.resize(2 * M);
    for (size_t t = 0; t < M; ++t) {
      const double ang = 2.0 * M_PI * (double)t / (double)M;
      sincos_fn(ang, &P.enc_w[2 * t + 1], &P.enc_w[2 * t]);
    }
    // Rounded Gaussian e = round(sigma z): P(e <= m) = Phi((m + 1/2) / sigma),
    // scaled to 2^64 and rounded; the upper half from the complement, so
    // that every entry keeps its relative precision; the last entry
    // saturates at 2^64 - 1.
    P.gauss_cdt.resize(2 * kGaussHalf);
    for (int j = 0; j < 2 * kGaussHalf; ++j) {
      const double x = ((double)(j - kGaussHalf) + 0.5) / kGaussSigma;
      if (x < 0) {
        P.gauss_cdt[j] = (uint64_t)std::nearbyint(std::ldexp(0.5 * std::erfc(-x / std::sqrt(2.0)), 64));
      } else {
        const uint64_t tail = (uint64_t)std::nearbyint(std::ldexp(0.5 * std::erfc(x / std::sqrt(2.0)), 64));
        P.gauss_cdt[j] = tail == 0 ? ~(uint64_t)0 : (uint64_t)0 - tail;
      }
    }
    P.rot_group.resize(N / 2);
    uint64_t g = 1;
    for (int j = 0; j < N / 2; ++j) {
      P.rot_group[j] = (uint32_t)g;
      g = g * 5 % M;
    }
  }
  return "";
}

}  // namespace aoa
