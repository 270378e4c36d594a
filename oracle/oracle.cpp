// ORACLE -- test infrastructure only.
//
// A host-only C++ restatement of the reference's CKKS evaluation path
// (pdroalves/aoa, aoa-ntt/).  It is the checker for the GPU library and the
// CPU baseline of the benchmark; product code never links or calls it.
//
// Each function follows the reference file:line given beside it, with the same
// tables and the same arithmetic: the custom Barrett `mod`, the two-pass
// hierarchical NTT with its gjk / nthroot / gN tables, the RNS constants
// computed with (here: minimal) big integers, and the poly state machine of
// poly_t.cpp that decides which transforms run.
//
// Pinning (tests/test_oracle.py): mod vs exact 128-bit remainder; NTT vs the
// O(N^2) closed form sum_n x[n] psi^(n(2 brev(pos)+1)); INTT o NTT = id;
// negacyclic products vs schoolbook; RNS constants vs Python big integers;
// ModUp vs the CRT identity; ModDown vs exact big-integer ModDown plus the
// Barrett deviation; rescale vs (a - a_l)/q_l.  The encoder FFT is pinned by
// FP64 round trips and by the canonical-embedding definition; encode/decode
// by the embedding (the residues, read as a polynomial over the integers,
// evaluate to scale * slot values at exp(i pi 5^j / N)), by agreement of the
// signed value across limbs and by decode o encode = id.  The sampler's
// restatement (Philox4x32-10 and the mapping of its output to each
// distribution) is pinned by the generator's published known-answer vectors,
// the Gaussian table against an independent normal CDF and the frequencies and
// moments of every distribution; the end-to-end C1 sequence by decryption
// within the reference's bound and by the defining relations of its keys.
#include <omp.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace orc {

struct aoa_prime_entry {
  uint64_t p;
  uint32_t g;
};
#include "../aoa_amd/csrc/prime_table.inc"

// ------------------------------------------------ 128-bit helpers (common.cu:140-170)
struct u128 {
  uint64_t lo = 0, hi = 0;
};
static inline uint64_t mul64hi(uint64_t a, uint64_t b) {
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
}
static inline u128 mullazy(uint64_t a, uint64_t b) {
  u128 c;
  c.hi = mul64hi(a, b);
  c.lo = a * b;
  return c;
}
static inline u128 rshift(u128 x, int b) {  // common.cu:129-136
  u128 y = x;
  y.lo = (y.lo >> b) + (y.hi << (64 - b));
  y.hi = (y.hi >> b);
  return y;
}

struct Oracle {
  int N = 0, k = 0, kl = 0, prec = 0, Na = 0, Nb = 0;
  std::vector<uint64_t> P;      // COPRIMES_BUCKET: Q primes then B primes
  std::vector<uint32_t> G;      // PROOTS
  std::vector<int> NB;          // RNSCoprimes_NumBits
  std::vector<uint64_t> MU;     // h_BARRETT_MU
  std::vector<uint64_t> nthroot, invnthroot;  // [R][N]
  std::map<int, std::vector<uint64_t>> gjk, invgjk;  // [R][log S][S]
  std::vector<uint64_t> gN, ginvN;                   // [R][N]
  std::vector<uint64_t> InvQqi, Qqibi, Bbiqi, InvBbi, InvModBqi, Bqi, Invqlqi;
  // Not in the reference: when set, ModDown reduces its cross-base products
  // exactly instead of with the reference's Barrett (the product's exact mode).
  bool exact_moddown = false;
  int R() const { return k + kl; }
};

// ------------------------------------------------ mod-arith (common.cu:174-304)
// mod(uint128_t, rid): the reference's Barrett reduction, common.cu:257-291.
static uint64_t ref_mod(const Oracle& o, u128 a, int rid) {
  const uint64_t p = o.P[rid];
  const int numbits = o.NB[rid];
  const uint64_t z = ((uint64_t)1 << numbits) - 1;
  const uint64_t m0 = a.lo & z;
  const uint64_t m1 = (a.hi << (64 - numbits)) + (a.lo >> numbits);
  const uint64_t MU = o.MU[rid];
  const u128 x = rshift(a, 2 * numbits - 64);
  const uint64_t q = (mul64hi(x.lo, MU) + (x.hi * MU));
  const u128 qp = mullazy(q, p);
  const uint64_t q0 = qp.lo & z;
  const uint64_t q1 = (qp.hi << (64 - numbits)) + (qp.lo >> numbits);
  const uint64_t r0 = (m0 - q0);
  const uint64_t r1 = (m1 - q1) - (m0 < q0);
  const uint64_t r = (r0 - r1 * p) & z;
  return r >= p ? r - p : r;
}
static inline uint64_t mulmod(const Oracle& o, uint64_t a, uint64_t b, int rid) {  // common.cu:174-193
  return ref_mod(o, mullazy(a, b), rid);
}
// addmod / submod: SEL-based, common.h:83-90
static inline uint64_t addmod(const Oracle& o, uint64_t a, uint64_t b, int rid) {
  const uint64_t p = o.P[rid];
  return (a + b) >= p ? a + b - p : a + b;
}
static inline uint64_t submod(const Oracle& o, uint64_t a, uint64_t b, int rid) {
  const uint64_t p = o.P[rid];
  return a < b ? a + (p - b) : a - b;
}
static uint64_t fast_pow(const Oracle& o, uint64_t a, uint64_t b, int rid) {  // common.cu:294-304
  uint64_t r = 1, s = a;
  while (b > 0) {
    if (b % 2 == 1) r = mulmod(o, r, s, rid);
    s = mulmod(o, s, s, rid);
    b /= 2;
  }
  return r;
}

// ------------------------------------------------ minimal big integers for precompute
struct Big {
  std::vector<uint32_t> w;  // little endian
};
static Big big_one() {
  Big b;
  b.w.push_back(1);
  return b;
}
static void big_mul(Big& a, uint64_t m) {
  unsigned __int128 carry = 0;
  for (auto& x : a.w) {
    unsigned __int128 t = (unsigned __int128)x * m + carry;
    x = (uint32_t)t;
    carry = t >> 32;
  }
  while (carry) {
    a.w.push_back((uint32_t)carry);
    carry >>= 32;
  }
}
// a /= d, returns remainder
static uint64_t big_divmod(Big& a, uint64_t d) {
  unsigned __int128 rem = 0;
  for (int i = (int)a.w.size() - 1; i >= 0; --i) {
    unsigned __int128 cur = (rem << 32) | a.w[i];
    a.w[i] = (uint32_t)(cur / d);
    rem = cur % d;
  }
  while (a.w.size() > 1 && a.w.back() == 0) a.w.pop_back();
  return (uint64_t)rem;
}
static uint64_t big_mod(const Big& a, uint64_t d) {
  Big c = a;
  return big_divmod(c, d);
}
static uint64_t inv_mod(uint64_t a, uint64_t p) {  // NTL::InvMod
  __int128 t = 0, nt = 1, r = p, nr = a % p;
  while (nr != 0) {
    __int128 q = r / nr, tmp;
    tmp = t - q * nt; t = nt; nt = tmp;
    tmp = r - q * nr; r = nr; nr = tmp;
  }
  if (t < 0) t += p;
  return (uint64_t)t;
}

// ------------------------------------------------ primes (manager.cu:426-464, :483-564)
static std::string fill_primes(Oracle& o) {
  std::vector<aoa_prime_entry> b;
  const aoa_prime_entry* qb;
  int qn;
  switch (o.prec) {
    case 45: qb = AOA_BUCKET_45; qn = 100; break;
    case 48: qb = AOA_BUCKET_48; qn = 100; break;
    case 52: qb = AOA_BUCKET_52; qn = 30; break;
    case 55: qb = AOA_BUCKET_55; qn = 500; break;
    default: return "We don't support this precision.";
  }
  o.P.clear();
  o.G.clear();
  o.P.push_back(AOA_BUCKET_63[0].p);
  o.G.push_back(AOA_BUCKET_63[0].g);
  if (o.k - 1 > qn) return "bucket too small";
  for (int i = 0; i < o.k - 1; ++i) {
    o.P.push_back(qb[i].p);
    o.G.push_back(qb[i].g);
  }
  for (int j = 0; j < o.kl; ++j) {
    const aoa_prime_entry* e;
    if (o.prec == 45) e = &AOA_BUCKET_45[o.k - 1 + j];
    else if (o.prec == 55) e = &AOA_BUCKET_55[o.k - 1 + j];
    else e = &AOA_BUCKET_55[j];
    o.P.push_back(e->p);
    o.G.push_back(e->g);
  }
  return "";
}

// ------------------------------------------------ Barrett constants (common.cu:485-521)
static void init_barrett(Oracle& o) {
  o.NB.assign(o.R(), 0);
  o.MU.assign(o.R(), 0);
  for (int r = 0; r < o.R(); ++r) {
    uint64_t p = o.P[r];
    int nb = 0;
    while ((nb < 64) && ((p >> nb) != 0)) ++nb;
    o.NB[r] = nb;
    o.MU[r] = (uint64_t)((((unsigned __int128)1) << (2 * nb)) / p);  // floor(4^nb / p)
  }
}

// ------------------------------------------------ NTT tables (common.cu:314-357, ntt.cu:685-1031)
static int ilog2(int n) {
  int r = 0;
  while ((1 << (r + 1)) <= n) ++r;
  return r;
}
static void compute_gjk(Oracle& o, int S) {
  const int lg = ilog2(S);
  std::vector<uint64_t>& g1 = o.gjk[S];
  std::vector<uint64_t>& g2 = o.invgjk[S];
  g1.assign((size_t)o.R() * lg * S, 0);
  g2.assign((size_t)o.R() * lg * S, 0);
  for (int rid = 0; rid < o.R(); ++rid) {
    const uint64_t p = o.P[rid];
    const uint64_t n = (p - 1) / (2 * S);
    const uint64_t g = fast_pow(o, o.G[rid], n, rid);
    for (int stride = 0; stride < lg; stride++)
      for (int j = 0; j < S; j++) {
        size_t idx = (size_t)rid * lg * S + stride * S + j;
        int e = j * ((S * 2) >> (lg - stride));
        g1[idx] = fast_pow(o, g, e, rid);
      }
    for (int stride = lg - 1; stride >= 0; stride--)
      for (int j = 0; j < S; j++) {
        size_t idx = (size_t)rid * lg * S + stride * S + j;
        int e = (S * 2 - j) * ((S * 2) >> (lg - stride));
        g2[idx] = fast_pow(o, g, e, rid);
      }
  }
}
static std::vector<int> brev_map(int n) {  // ntt.cu:961-990
  std::vector<int> m(n, 0);
  for (int i = 1; i < n; i++) m[i] = i;
  int j = 0;
  for (int i = 1; i < n; i++) {
    int b = n >> 1;
    while (j >= b) {
      j -= b;
      b >>= 1;
    }
    j += b;
    if (j > i) {
      m[i] = j;
      m[j] = i;
    }
  }
  return m;
}
static void factor(int N, int* Na, int* Nb) {  // ntt.h:24-44
  switch (N) {
    case 128: *Na = 16; *Nb = 8; break;
    case 256: *Na = 16; *Nb = 16; break;
    case 512: *Na = 32; *Nb = 16; break;
    case 1024: *Na = 32; *Nb = 32; break;
    case 2048: *Na = 64; *Nb = 32; break;
    case 4096: *Na = 64; *Nb = 64; break;
    case 8192: *Na = 128; *Nb = 64; break;
    case 16384: *Na = 128; *Nb = 128; break;
    case 32768: *Na = 256; *Nb = 128; break;
    case 65536: *Na = 256; *Nb = 256; break;
    default: *Na = *Nb = 0;
  }
}
static void init_ntt(Oracle& o) {
  const int N = o.N;
  factor(N, &o.Na, &o.Nb);
  compute_gjk(o, o.Na);
  if (o.Nb != o.Na) compute_gjk(o, o.Nb);
  o.nthroot.assign((size_t)o.R() * N, 0);
  o.invnthroot.assign((size_t)o.R() * N, 0);
  o.gN.assign((size_t)o.R() * N, 0);
  o.ginvN.assign((size_t)o.R() * N, 0);
  std::vector<int> bNb = brev_map(o.Nb);
#pragma omp parallel for
  for (int rid = 0; rid < o.R(); ++rid) {
    const uint64_t p = o.P[rid];
    const uint64_t psi = fast_pow(o, o.G[rid], (p - 1) / (2 * (uint64_t)N), rid);
    const uint64_t invpsi = inv_mod(psi, p);
    uint64_t a = 1, b = 1;
    for (int cid = 0; cid < N; cid++) {  // compute_nthroot_ntt, hierarchical variant (no N^-1)
      o.nthroot[(size_t)rid * N + cid] = a;
      o.invnthroot[(size_t)rid * N + cid] = b;
      a = mulmod(o, a, psi, rid);
      b = mulmod(o, b, invpsi, rid);
    }
    const uint64_t g = fast_pow(o, o.G[rid], (p - 1) / (uint64_t)N, rid);
    const uint64_t ginv = inv_mod(g, p);
    const uint64_t nainv = inv_mod((uint64_t)(o.Na * o.Nb) % p, p);
    for (int cid = 0; cid < N; cid++) {
      o.gN[(size_t)rid * N + cid] = fast_pow(o, g, (uint64_t)(cid / o.Nb) * bNb[cid % o.Nb], rid);
      o.ginvN[(size_t)rid * N + cid] =
          mulmod(o, fast_pow(o, ginv, (uint64_t)bNb[cid / o.Na] * (cid % o.Na), rid), nainv, rid);
    }
  }
}

// ------------------------------------------------ RNS precompute (manager.cu:566-759)
static void precompute(Oracle& o) {
  const int k = o.k, kl = o.kl;
  std::vector<uint64_t> Q(o.P.begin(), o.P.begin() + k), B(o.P.begin() + k, o.P.end());
  Big Qprod = big_one(), Bprod = big_one();
  for (auto q : Q) big_mul(Qprod, q);
  for (auto b : B) big_mul(Bprod, b);
  o.InvModBqi.assign(k, 0);
  o.Bqi.assign(k, 0);
  for (int i = 0; i < k; ++i) {
    o.Bqi[i] = big_mod(Bprod, Q[i]);
    o.InvModBqi[i] = inv_mod(o.Bqi[i], Q[i]);
  }
  o.InvBbi.assign(kl, 0);
  for (int j = 0; j < kl; ++j) {
    Big bb = Bprod;
    big_divmod(bb, B[j]);  // B / b_j
    o.InvBbi[j] = inv_mod(big_mod(bb, B[j]), B[j]);
  }
  o.Invqlqi.assign((size_t)k * k, 0);
  for (int l = 0; l < k; ++l)
    for (int i = 0; i < k; ++i)
      if (l != i) o.Invqlqi[(size_t)l * k + i] = inv_mod(Q[l] % Q[i], Q[i]);
  o.Bbiqi.assign((size_t)k * kl, 0);  // pushed for qi: for bi  -> [i*kl + j]
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < kl; ++j) {
      Big bb = Bprod;
      big_divmod(bb, B[j]);
      o.Bbiqi[(size_t)i * kl + j] = big_mod(bb, Q[i]);
    }
  o.InvQqi.assign((size_t)k * k, 0);
  o.Qqibi.assign((size_t)k * k * kl, 0);
  for (int l = 0; l < k; ++l) {
    Big Ql = big_one();
    for (int m = 0; m <= l; ++m) big_mul(Ql, Q[m]);
    for (int i = 0; i <= l; ++i) {  // only i <= l is ever read (manager.cu:196-219)
      Big e = Ql;
      big_divmod(e, Q[i]);  // multiply_subset_except(Q, 0, l+1, i)
      o.InvQqi[(size_t)l * k + i] = inv_mod(big_mod(e, Q[i]), Q[i]);
      for (int j = 0; j < kl; ++j) o.Qqibi[((size_t)l * k + i) * kl + j] = big_mod(e, B[j]);
    }
  }
}

// ------------------------------------------------ NTT (ntt.cu:128-475)
static void doNTT(const Oracle& o, uint64_t* s, int S, int rid) {  // ntt.cu:196-214, :128-158
  const int lg = ilog2(S);
  const uint64_t* gjk = o.gjk.at(S).data() + (size_t)rid * lg * S;
  for (int stride = 0; stride < lg; stride++) {
    const int m = S / (2 << stride);
    for (int cid = 0; cid < S / 2; ++cid) {
      const int j = cid * 2 * m / S;
      const int i = j + (cid % (S / (2 * m))) * 2 * m;
      const uint64_t xi = s[i], xim = s[i + m];
      const uint64_t a = gjk[stride * S + j];
      s[i] = addmod(o, xi, xim, rid);
      s[i + m] = mulmod(o, a, submod(o, xi, xim, rid), rid);
    }
  }
}
static void doINTT(const Oracle& o, uint64_t* s, int S, int rid) {  // ntt.cu:217-246, :160-193
  const int lg = ilog2(S);
  const uint64_t* ig = o.invgjk.at(S).data() + (size_t)rid * lg * S;
  for (int stride = lg - 1; stride >= 0; stride--) {
    const int m = S / (2 << stride);
    for (int cid = 0; cid < S / 2; ++cid) {
      const int j = cid * 2 * m / S;
      const int i = j + (cid % (S / (2 * m))) * 2 * m;
      const uint64_t xi = s[i], xim = s[i + m];
      const uint64_t a = ig[stride * S + j];
      const uint64_t axim = mulmod(o, a, xim, rid);
      s[i] = addmod(o, xi, axim, rid);
      s[i + m] = submod(o, xi, axim, rid);
    }
  }
}
// Forward NTT of one residue: hntt_tr<Nb> then hntt_tr<Na> (ntt.cu:36-52, :321-363).
static void ntt_residue(const Oracle& o, uint64_t* x, int rid) {
  const int N = o.N, Na = o.Na, Nb = o.Nb;
  std::vector<uint64_t> tmp(N), s(std::max(Na, Nb));
  const uint64_t* C = o.nthroot.data() + (size_t)rid * N;
  for (int b = 0; b < Na; ++b) {
    for (int t = 0; t < Nb; ++t) s[t] = mulmod(o, x[t * Na + b], C[t * Na + b], rid);
    doNTT(o, s.data(), Nb, rid);
    for (int t = 0; t < Nb; ++t) tmp[b * Nb + t] = s[t];
  }
  const uint64_t* C2 = o.gN.data() + (size_t)rid * N;
  for (int b = 0; b < Nb; ++b) {
    for (int t = 0; t < Na; ++t) s[t] = mulmod(o, tmp[t * Nb + b], C2[t * Nb + b], rid);
    doNTT(o, s.data(), Na, rid);
    for (int t = 0; t < Na; ++t) x[b * Na + t] = s[t];
  }
}
// Inverse NTT of one residue: hintt<Na> then hintt_tr<Nb> (ntt.cu:53-66, :372-460).
static void intt_residue(const Oracle& o, uint64_t* x, int rid) {
  const int N = o.N, Na = o.Na, Nb = o.Nb;
  std::vector<uint64_t> tmp(N), s(std::max(Na, Nb));
  const uint64_t* C = o.ginvN.data() + (size_t)rid * N;
  for (int r = 0; r < Nb; ++r) {
    for (int t = 0; t < Na; ++t) s[t] = x[r * Na + t];
    doINTT(o, s.data(), Na, rid);
    for (int t = 0; t < Na; ++t) tmp[r * Na + t] = mulmod(o, s[t], C[r * Na + t], rid);
  }
  const uint64_t* C2 = o.invnthroot.data() + (size_t)rid * N;
  for (int b = 0; b < Na; ++b) {
    for (int t = 0; t < Nb; ++t) s[t] = tmp[t * Na + b];
    doINTT(o, s.data(), Nb, rid);
    for (int t = 0; t < Nb; ++t) x[t * Na + b] = mulmod(o, s[t], C2[t * Na + b], rid);
  }
}
static void transform(const Oracle& o, uint64_t* d, int res0, int nres, int dir) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int r = res0; r < res0 + nres; ++r) {
    if (dir == 0)
      ntt_residue(o, d + (size_t)r * o.N, r);
    else
      intt_residue(o, d + (size_t)r * o.N, r);
  }
}

// ------------------------------------------------ poly layer (poly_t.cpp)
enum { RNSSTATE = 1, TRANSSTATE = 2 };
enum { QBase = 0, QBBase = 2 };
struct Poly {
  std::vector<uint64_t> d;
  int state = RNSSTATE;
  int base = QBase;
};
static int nres_of(const Oracle& o, int base) { return base == QBase ? o.k : o.k + o.kl; }
static Poly poly_new(const Oracle& o, int base) {
  Poly p;
  p.base = base;
  p.state = RNSSTATE;
  p.d.assign((size_t)nres_of(o, base) * o.N, 0);
  return p;
}
static void execute(const Oracle& o, Poly& p, int dir) {  // ntt.cu:463-475
  if ((p.state == TRANSSTATE && dir == 0) || (p.state == RNSSTATE && dir == 1)) return;
  transform(o, p.d.data(), 0, nres_of(o, p.base), dir);
  p.state = dir == 0 ? TRANSSTATE : RNSSTATE;
}
static void vec_op(const Oracle& o, int op, uint64_t* c, const uint64_t* a, const uint64_t* b, int nres) {
  // ntt_vector_op, ntt.cu:479-509
  const size_t n = (size_t)nres * o.N;
#pragma omp parallel for
  for (size_t t = 0; t < n; ++t) {
    const int rid = (int)(t / o.N);
    switch (op) {
      case 0: c[t] = addmod(o, a[t], b[t], rid); break;
      case 1: c[t] = submod(o, a[t], b[t], rid); break;
      case 2: c[t] = mulmod(o, a[t], b[t], rid); break;
      case 3: c[t] = (o.P[rid] - a[t]) * (a[t] != 0); break;
    }
  }
}
static void vec_op_int(const Oracle& o, int op, uint64_t* c, const uint64_t* a, uint64_t b, int nres) {
  // ntt_vector_op_int, ntt.cu:531-558: the scalar is used as given (not
  // reduced), with the same SEL-based add/sub and Barrett mulmod.
  const size_t n = (size_t)nres * o.N;
#pragma omp parallel for
  for (size_t t = 0; t < n; ++t) {
    const int rid = (int)(t / o.N);
    switch (op) {
      case 0: c[t] = addmod(o, a[t], b, rid); break;
      case 1: c[t] = submod(o, a[t], b, rid); break;
      case 2: c[t] = mulmod(o, a[t], b, rid); break;
    }
  }
}
static void conjugation_kernel(const Oracle& o, uint64_t* b, const uint64_t* a, int nres) {
  // execute_conjugation, encoder.cu:391-401 (NTT domain): b[cid] = a[N-1-cid]
  for (int r = 0; r < nres; ++r)
    for (int cid = 0; cid < o.N; ++cid) b[(size_t)r * o.N + cid] = a[(size_t)r * o.N + (o.N - 1 - cid)];
}
static void poly_add_add(const Oracle& o, Poly& c1, Poly& a1, Poly& b1, Poly& c2, Poly& a2, Poly& b2) {
  // poly_t.cpp:237-272
  const int base = std::min(std::min(a1.base, b1.base), std::min(a2.base, b2.base));
  if (a1.state != b1.state) {
    execute(o, a1, 1);
    execute(o, b1, 1);
  }
  if (a2.state != b2.state) {
    execute(o, a2, 1);
    execute(o, b2, 1);
  }
  const int nr = nres_of(o, base);
  std::vector<uint64_t> r1((size_t)nr * o.N), r2((size_t)nr * o.N);
  vec_op(o, 0, r1.data(), a1.d.data(), b1.d.data(), nr);
  vec_op(o, 0, r2.data(), a2.d.data(), b2.d.data(), nr);
  std::memcpy(c1.d.data(), r1.data(), r1.size() * 8);
  std::memcpy(c2.d.data(), r2.data(), r2.size() * 8);
  c1.state = a1.state;
  c2.state = a2.state;
}
static void poly_binop(const Oracle& o, int op, Poly& c, Poly& a, Poly& b) {
  // poly_add / poly_sub (poly_t.cpp:137-155, :173-191): INVERSE both on a
  // state mismatch; poly_mul (:201-220): FORWARD both.
  const int base = std::min(a.base, b.base);
  if (op == 2) {
    execute(o, a, 0);
    execute(o, b, 0);
  } else if (a.state != b.state) {
    execute(o, a, 1);
    execute(o, b, 1);
  }
  const int nr = nres_of(o, base);
  std::vector<uint64_t> r((size_t)nr * o.N);
  vec_op(o, op, r.data(), a.d.data(), b.d.data(), nr);
  std::memcpy(c.d.data(), r.data(), r.size() * 8);
  c.state = a.state;
}
static void modup_kernel(const Oracle& o, uint64_t* qb, int level) {  // manager.cu:180-223
  const int N = o.N, k = o.k, kl = o.kl;
#pragma omp parallel for
  for (int cid = 0; cid < N; ++cid)
    for (int rid_B = 0; rid_B < kl; ++rid_B) {
      uint64_t val = 0;
      for (int rid_Q = 0; rid_Q <= level; rid_Q++) {
        uint64_t aux = mulmod(o, qb[cid + (size_t)rid_Q * N], o.InvQqi[(size_t)level * k + rid_Q], rid_Q);
        aux %= o.P[rid_B + k];
        const uint64_t c = o.Qqibi[((size_t)level * k + rid_Q) * kl + rid_B];
        val = addmod(o, val, mulmod(o, aux, c, rid_B + k), rid_B + k);
      }
      qb[cid + (size_t)(rid_B + k) * N] = val;
    }
}
static void moddown_kernel(const Oracle& o, uint64_t* out, const uint64_t* qb, int level) {  // manager.cu:312-341
  const int N = o.N, k = o.k, kl = o.kl;
#pragma omp parallel for
  for (int cid = 0; cid < N; ++cid)
    for (int rid_Q = 0; rid_Q <= level; ++rid_Q) {
      uint64_t val1 = 0;
      for (int rid_B = 0; rid_B < kl; rid_B++) {
        const uint64_t aux1 = mulmod(o, qb[cid + (size_t)(rid_B + k) * N], o.InvBbi[rid_B], rid_B + k);
        const uint64_t c = o.Bbiqi[(size_t)rid_Q * kl + rid_B];
        const uint64_t tmp = o.exact_moddown ? (uint64_t)(((unsigned __int128)aux1 * c) % o.P[rid_Q])
                                             : mulmod(o, aux1, c, rid_Q);
        val1 = addmod(o, val1, tmp, rid_Q);
      }
      const uint64_t tmp = submod(o, qb[cid + (size_t)rid_Q * N], val1, rid_Q);
      out[cid + (size_t)rid_Q * N] = mulmod(o, tmp, o.InvModBqi[rid_Q], rid_Q);
    }
}
static void rescale_kernel(const Oracle& o, uint64_t* a, uint64_t* b, int level) {  // manager.cu:90-118
  const int N = o.N, k = o.k;
#pragma omp parallel for
  for (int t = 0; t < N * level; ++t) {
    const int cid = t % N, rid = t / N;
    const uint64_t s = o.Invqlqi[(size_t)level * k + rid];
    uint64_t ca = a[cid + (size_t)rid * N], ra = a[cid + (size_t)level * N];
    uint64_t cb = b[cid + (size_t)rid * N], rb = b[cid + (size_t)level * N];
    ca = submod(o, ca, ra, rid);
    cb = submod(o, cb, rb, rid);
    a[cid + (size_t)rid * N] = mulmod(o, ca, s, rid);
    b[cid + (size_t)rid * N] = mulmod(o, cb, s, rid);
  }
}
static void poly_modup(const Oracle& o, Poly& a, Poly& b, int level) {  // poly_t.cpp:345-361
  std::memcpy(a.d.data(), b.d.data(), (size_t)o.k * o.N * 8);  // poly_copy copies |b| = Q bytes
  a.state = b.state;
  execute(o, a, 1);
  modup_kernel(o, a.d.data(), level);
}
static void poly_moddown(const Oracle& o, Poly& a, Poly& b, int level) {  // poly_t.cpp:363-378
  execute(o, b, 1);
  moddown_kernel(o, a.d.data(), b.d.data(), level);
  a.state = RNSSTATE;
}

// Galois automorphism kernel (encoder.cu:300-322), scatter form.
static void rotation_kernel(const Oracle& o, uint64_t* b, const uint64_t* a, long pw, int nres) {
  const int N = o.N;
  for (int rid = 0; rid < nres; ++rid)
    for (int cid = 0; cid < N; ++cid) {
      long new_cid = (long)cid * pw % (N << 1);
      const int flag = (new_cid >= N);
      if (flag) new_cid -= N;
      const uint64_t v = a[cid + (size_t)rid * N];
      b[new_cid + (size_t)rid * N] = flag ? submod(o, o.P[rid], v, rid) : v;
    }
}
static long powmod_small(long base, long e, long m) {
  long r = 1 % m, s = base % m;
  while (e) {
    if (e & 1) r = r * s % m;
    s = s * s % m;
    e >>= 1;
  }
  return r;
}

// ------------------------------------------------ encoder FFT (encoder.cu:10-124)
// ksiPows[j] = (cos(angle), sin(angle)) (encoder.cu:44-48): an optimising g++
// (the reference's Release build, -O3) emits one glibc sincos call for the
// pair, which differs from separate cos and sin in the last bit for a few
// angles; sincos is called explicitly, through a volatile pointer.
static void libm_sincos(double a, double* s, double* c) {
  void (*volatile fn)(double, double*, double*) = ::sincos;
  fn(a, s, c);
}
static void array_bit_reverse(std::complex<double>* vals, long size) {  // encoder.cu:12-23
  for (long i = 1, j = 0; i < size; ++i) {
    long bit = size >> 1;
    for (; j >= bit; bit >>= 1) j -= bit;
    j += bit;
    if (i < j) std::swap(vals[i], vals[j]);
  }
}
// fftSpecial (encoder.cu:33-73): Nh = ring dimension N, M = 4N.
static void fft_special(std::complex<double>* vals, long size, long ringN) {
  const long Nh = ringN, M = 4 * ringN;
  std::vector<long> rotGroup(Nh);
  long g = 1;
  for (long i = 0; i < Nh; ++i) {
    rotGroup[i] = g;
    g = (g * 5) % M;
  }
  std::vector<std::complex<double>> ksi(M + 1);
  for (long j = 0; j < M; ++j) {
    const double angle = 2.0 * M_PI * j / M;
    double sn, cs;
    libm_sincos(angle, &sn, &cs);
    ksi[j] = std::complex<double>(cs, sn);
  }
  ksi[M] = ksi[0];
  array_bit_reverse(vals, size);
  for (long len = 2; len <= size; len <<= 1)
    for (long i = 0; i < size; i += len) {
      const long lenh = len >> 1, lenq = len << 2;
      for (long j = 0; j < lenh; ++j) {
        const long idx = (rotGroup[j] % lenq) * M / lenq;
        std::complex<double> u = vals[i + j], v = vals[i + j + lenh] * ksi[idx];
        vals[i + j] = u + v;
        vals[i + j + lenh] = u - v;
      }
    }
}
// fftSpecialInv (encoder.cu:75-124): N = ring dimension, M = 2N, scaled by 1/size.
static void fft_special_inv(std::complex<double>* vals, long size, long ringN) {
  const long Nh = ringN / 2, M = 2 * ringN;
  std::vector<long> rotGroup(Nh);
  long g = 1;
  for (long i = 0; i < Nh; ++i) {
    rotGroup[i] = g;
    g = (g * 5) % M;
  }
  std::vector<std::complex<double>> ksi(M + 1);
  for (long j = 0; j < M; ++j) {
    const double angle = 2.0 * M_PI * j / M;
    double sn, cs;
    libm_sincos(angle, &sn, &cs);
    ksi[j] = std::complex<double>(cs, sn);
  }
  ksi[M] = ksi[0];
  for (long len = size; len >= 1; len >>= 1)
    for (long i = 0; i < size; i += len) {
      const long lenh = len >> 1, lenq = len << 2;
      for (long j = 0; j < lenh; ++j) {
        const long idx = (lenq - (rotGroup[j] % lenq)) * M / lenq;
        std::complex<double> u = vals[i + j] + vals[i + j + lenh];
        std::complex<double> v = (vals[i + j] - vals[i + j + lenh]) * ksi[idx];
        vals[i + j] = u;
        vals[i + j + lenh] = v;
      }
    }
  array_bit_reverse(vals, size);
  for (long i = 0; i < size; ++i) vals[i] /= (double)size;
}

// ------------------------------------------------ encoder (encoder.cu:126-298)
// ADJUSTNEG (encoder.cu:128): a negative value whose magnitude is a multiple
// of p gives p, not 0.
static inline uint64_t adjust_neg(int64_t x, uint64_t p) {
  return x >= 0 ? (uint64_t)x % p : p - ((uint64_t)(-x) % p);
}
// ADJUSTSCALE (encoder.cu:129).
static inline double adjust_scale(uint64_t x, uint64_t p, uint64_t factor) {
  return x <= (p >> 1) ? ((double)x) / factor : (((double)x) - ((double)p)) / (double)factor;
}
// ckks_encode (encoder.cu:166-221) with ckks_convert_C2PC (:141-164): the
// caller's values (filled, then zeros) go through fftSpecialInv; residues of
// nres limbs, zero outside the slot positions.
static void encode(const Oracle& o, uint64_t* out, std::complex<double>* vals, int slots, uint64_t scale, int nres) {
  fft_special_inv(vals, slots, o.N);
  std::memset(out, 0, (size_t)nres * o.N * sizeof(uint64_t));
  const int gap = (o.N >> 1) / slots;
  for (int lid = 0; lid < nres; ++lid)
    for (int sid = 0; sid < slots; ++sid) {
      const int64_t mir = vals[sid].real() * scale;
      const int64_t mii = vals[sid].imag() * scale;
      const uint64_t p = o.P[lid];
      out[(size_t)sid * gap + (size_t)lid * o.N] = adjust_neg(mir, p);
      out[(size_t)sid * gap + (o.N >> 1) + (size_t)lid * o.N] = adjust_neg(mii, p);
    }
}
// ckks_decode (encoder.cu:259-298) with ckks_convert_PC2C (:232-257): limb 0
// (coefficient domain), then fftSpecial.
static void decode(const Oracle& o, std::complex<double>* out, const uint64_t* a, int slots, uint64_t scale) {
  const uint64_t p = o.P[0];
  const int gap = (o.N >> 1) / slots;
  for (int sid = 0; sid < slots; ++sid)
    out[sid] = std::complex<double>(adjust_scale(a[(size_t)sid * gap], p, scale),
                                    adjust_scale(a[(size_t)sid * gap + (o.N >> 1)], p, scale));
  fft_special(out, slots, o.N);
}

// ------------------------------------------------ sampler
// The reference draws from cuRAND (Sobol64 for uniform, XORWOW normals for
// the Gaussian) and host rand() (sampler.cu:21-225), which cannot be
// reproduced.  The product's sampler keeps the reference's distributions with
// a counter-based generator; this is its restatement: Philox4x32-10 as
// published by Salmon et al. (SC 2011, Random123), the counter / key layout,
// and the mapping of 128 random bits to each distribution.
struct Ctr4 {
  uint32_t v[4];
};
static Ctr4 philox4x32_10(Ctr4 c, uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.v[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.v[2];
    Ctr4 n;
    n.v[0] = (uint32_t)(p1 >> 32) ^ c.v[1] ^ k0;
    n.v[1] = (uint32_t)p1;
    n.v[2] = (uint32_t)(p0 >> 32) ^ c.v[3] ^ k1;
    n.v[3] = (uint32_t)p0;
    c = n;
  }
  return c;
}
enum { S_UNIFORM = 0, S_GAUSSIAN = 1, S_ZO = 2, S_HW = 3, S_NARROW = 4 };
static Ctr4 draw(uint64_t seed, uint64_t nonce, uint32_t tag, uint32_t idx) {
  Ctr4 c{{idx, tag, (uint32_t)nonce, (uint32_t)(nonce >> 32)}};
  return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}
// P(round(3.2 z) <= j - 32) * 2^64, rounded; upper half via the complement;
// the last entry saturates.
static std::vector<uint64_t> gauss_cdt() {
  std::vector<uint64_t> t(64);
  for (int j = 0; j < 64; ++j) {
    const double x = ((double)(j - 32) + 0.5) / 3.2;
    if (x < 0) {
      t[j] = (uint64_t)std::nearbyint(std::ldexp(0.5 * std::erfc(-x / std::sqrt(2.0)), 64));
    } else {
      const uint64_t tail = (uint64_t)std::nearbyint(std::ldexp(0.5 * std::erfc(x / std::sqrt(2.0)), 64));
      t[j] = tail == 0 ? ~(uint64_t)0 : (uint64_t)0 - tail;
    }
  }
  return t;
}
static uint64_t small_res(int64_t v, uint64_t q) { return v >= 0 ? (uint64_t)v : q + v; }
static void sample(const Oracle& o, uint64_t* out, int res0, int nres, int kind, uint64_t seed, uint64_t nonce) {
  const int N = o.N;
  if (kind == S_UNIFORM) {
    for (int l = res0; l < res0 + nres; ++l)
      for (int i = 0; i < N; ++i) {
        const Ctr4 c = draw(seed, nonce, ((uint32_t)S_UNIFORM << 24) | (uint32_t)l, (uint32_t)i);
        const unsigned __int128 X = ((unsigned __int128)(((uint64_t)c.v[3] << 32) | c.v[2]) << 64) |
                                    (((uint64_t)c.v[1] << 32) | c.v[0]);
        // floor(X q / 2^128) from the 192-bit product
        const uint64_t q = o.P[l];
        const unsigned __int128 lo = (unsigned __int128)(uint64_t)X * q;
        const unsigned __int128 hi = (unsigned __int128)(uint64_t)(X >> 64) * q;
        out[(size_t)l * N + i] = (uint64_t)((hi + (lo >> 64)) >> 64);
      }
    return;
  }
  if (kind == S_HW) {
    for (int l = res0; l < res0 + nres; ++l) std::memset(out + (size_t)l * N, 0, (size_t)N * 8);
    std::vector<char> used(N, 0);
    const int weight = std::min(64, N);
    int n = 0;
    for (uint32_t a = 0; a < 4096 && n < weight; ++a) {
      const Ctr4 c = draw(seed, nonce, ((uint32_t)S_HW << 24) | 0xFFFFFFu, a);
      const uint32_t p = c.v[0] & (uint32_t)(N - 1);
      if (used[p]) continue;
      used[p] = 1;
      const int64_t v = (c.v[1] & 1u) ? -1 : 1;
      for (int l = res0; l < res0 + nres; ++l) out[(size_t)l * N + p] = small_res(v, o.P[l]);
      ++n;
    }
    return;
  }
  static const std::vector<uint64_t> cdt = gauss_cdt();
  for (int i = 0; i < N; ++i) {
    const Ctr4 c = draw(seed, nonce, ((uint32_t)kind << 24) | 0xFFFFFFu, (uint32_t)i);
    const uint64_t u = ((uint64_t)c.v[1] << 32) | c.v[0];
    int64_t v;
    if (kind == S_GAUSSIAN) {
      int cnt = 0;
      for (int j = 0; j < 64; ++j) cnt += u >= cdt[j];
      v = cnt - 32;
    } else if (kind == S_ZO) {
      // exactly N/2 nonzeros (sample_Z0, aoa-ntt/src/cuda/sampler.cu:119-150):
      // six Feistel rounds permute the index, H of ceil(logN/2) bits and Lo
      // of floor(logN/2); i is nonzero when its image has H's top bit clear
      int logN = 0;
      while ((1 << logN) < N) ++logN;
      const int hb = (logN + 1) / 2, lb = logN / 2;
      uint32_t h = (uint32_t)i >> lb, lo = (uint32_t)i & ((1u << lb) - 1u);
      for (uint32_t r = 0; r < 6; ++r) {
        const uint32_t tag = ((uint32_t)S_ZO << 24) | (0xFFFF00u + r);
        if (r & 1)
          h ^= draw(seed, nonce, tag, lo).v[0] & ((1u << hb) - 1u);
        else
          lo ^= draw(seed, nonce, tag, h).v[0] & ((1u << lb) - 1u);
      }
      v = (h >> (hb - 1)) ? 0 : ((c.v[1] & 1u) ? -1 : 1);
    } else {
      v = (int64_t)(uint64_t)(((unsigned __int128)u * 3) >> 64) - 1;
    }
    for (int l = res0; l < res0 + nres; ++l) out[(size_t)l * N + i] = small_res(v, o.P[l]);
  }
}

// ------------------------------------------------ C1 pipeline
// One plaintext product end to end, on the sequence of the reference's
// ckks_keygen / ckks_encrypt / ckks_mul / ckks_decrypt (ckkskeys.cpp:61-131,
// ckks.cpp:87-105, :273-358, :107-132) with the library's sampler: keys, then
// encode -> encrypt -> HomMul+relin+rescale -> decrypt -> decode.  Public
// key on Q limbs: b = e - a s; relinearisation key on QB limbs:
// b = rho(s^2) + e - a s; all in the NTT domain.  Nonces: s 0, pk a 1, pk e
// 2, evk a 3, evk e 4; an encryption at nonce n draws u (ZO) at n, e0 at n+1,
// e1 at n+2.
struct C1Keys {
  std::vector<uint64_t> s, pkb, pka, evkb, evka;
};
static void c1_keygen(const Oracle& o, C1Keys& K, uint64_t seed) {
  const int k = o.k, R = o.k + o.kl;
  const size_t N = o.N, qn = (size_t)k * N, qbn = (size_t)R * N;
  K.s.assign(qbn, 0);
  sample(o, K.s.data(), 0, R, S_HW, seed, 0);
  transform(o, K.s.data(), 0, R, 0);
  K.pka.assign(qn, 0);
  sample(o, K.pka.data(), 0, k, S_UNIFORM, seed, 1);
  std::vector<uint64_t> e(qbn), t(qbn);
  sample(o, e.data(), 0, k, S_GAUSSIAN, seed, 2);
  transform(o, e.data(), 0, k, 0);
  vec_op(o, 2, t.data(), K.pka.data(), K.s.data(), k);
  K.pkb.assign(qn, 0);
  vec_op(o, 1, K.pkb.data(), e.data(), t.data(), k);
  K.evka.assign(qbn, 0);
  sample(o, K.evka.data(), 0, R, S_UNIFORM, seed, 3);
  sample(o, e.data(), 0, R, S_GAUSSIAN, seed, 4);
  transform(o, e.data(), 0, R, 0);
  std::vector<uint64_t> s2(qbn, 0), rho(qbn, 0);
  vec_op(o, 2, s2.data(), K.s.data(), K.s.data(), R);
  for (int i = 0; i < k; ++i)  // rho: Q limbs times B mod q_i, special limbs 0 (manager.cu:149-176)
    for (size_t c = 0; c < N; ++c) rho[i * N + c] = mulmod(o, s2[i * N + c], o.Bqi[i], i);
  vec_op(o, 2, t.data(), K.evka.data(), K.s.data(), R);
  K.evkb.assign(qbn, 0);
  vec_op(o, 0, K.evkb.data(), rho.data(), e.data(), R);
  vec_op(o, 1, K.evkb.data(), K.evkb.data(), t.data(), R);
}
static void c1_encrypt(const Oracle& o, const C1Keys& K, uint64_t* c0, uint64_t* c1, const std::complex<double>* z,
                       int slots, uint64_t scale, uint64_t seed, uint64_t nonce) {
  const int k = o.k;
  const size_t qn = (size_t)k * o.N;
  std::vector<uint64_t> m(qn), u(qn), e0(qn), e1(qn), t(qn);
  std::vector<std::complex<double>> v(z, z + slots);
  encode(o, m.data(), v.data(), slots, scale, k);
  transform(o, m.data(), 0, k, 0);
  sample(o, u.data(), 0, k, S_ZO, seed, nonce);
  sample(o, e0.data(), 0, k, S_GAUSSIAN, seed, nonce + 1);
  sample(o, e1.data(), 0, k, S_GAUSSIAN, seed, nonce + 2);
  transform(o, u.data(), 0, k, 0);
  transform(o, e0.data(), 0, k, 0);
  transform(o, e1.data(), 0, k, 0);
  vec_op(o, 2, t.data(), K.pkb.data(), u.data(), k);
  vec_op(o, 0, c0, t.data(), e0.data(), k);
  vec_op(o, 0, c0, c0, m.data(), k);
  vec_op(o, 2, t.data(), K.pka.data(), u.data(), k);
  vec_op(o, 0, c1, t.data(), e1.data(), k);
}

}  // namespace orc

using namespace orc;

// ================================================================ C ABI
extern "C" {

void* orc_create(int N, int k, int kl, int prec) {
  Oracle* o = new Oracle();
  o->N = N;
  o->k = k;
  o->kl = kl > 0 ? kl : k + 1;
  o->prec = prec;
  if (!fill_primes(*o).empty()) {
    delete o;
    return nullptr;
  }
  init_barrett(*o);
  int Na, Nb;
  factor(N, &Na, &Nb);
  if (Na == 0) {
    delete o;
    return nullptr;
  }
  init_ntt(*o);
  precompute(*o);
  return o;
}
void orc_destroy(void* h) { delete (Oracle*)h; }
int orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}
int orc_primes(void* h, uint64_t* out) {
  Oracle* o = (Oracle*)h;
  for (int r = 0; r < o->R(); ++r) out[r] = o->P[r];
  return o->R();
}
void orc_set_moddown_exact(void* h, int exact) { ((Oracle*)h)->exact_moddown = exact != 0; }
// Encoder FFT on interleaved (re, im) doubles; inverse = encoding direction.
void orc_fft_special(double* v, long size, long ringN, int inverse) {
  auto* c = reinterpret_cast<std::complex<double>*>(v);
  if (inverse)
    fft_special_inv(c, size, ringN);
  else
    fft_special(c, size, ringN);
}
// Encoder on interleaved (re, im) doubles: vals (slots values) is
// transformed in place as in the reference.
void orc_encode(void* h, uint64_t* out, double* vals, int slots, uint64_t scale, int nres) {
  encode(*(Oracle*)h, out, reinterpret_cast<std::complex<double>*>(vals), slots, scale, nres);
}
void orc_decode(void* h, double* vals, const uint64_t* in, int slots, uint64_t scale) {
  decode(*(Oracle*)h, reinterpret_cast<std::complex<double>*>(vals), in, slots, scale);
}
void orc_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
  Ctr4 c{{ctr[0], ctr[1], ctr[2], ctr[3]}};
  c = philox4x32_10(c, key[0], key[1]);
  for (int i = 0; i < 4; ++i) out[i] = c.v[i];
}
void orc_gauss_cdt(uint64_t* out) {
  const std::vector<uint64_t> t = gauss_cdt();
  for (int i = 0; i < 64; ++i) out[i] = t[i];
}
void orc_sample(void* h, uint64_t* out, int res0, int nres, int kind, uint64_t seed, uint64_t nonce) {
  sample(*(Oracle*)h, out, res0, nres, kind, seed, nonce);
}
void* orc_c1_keygen(void* h, uint64_t seed) {
  C1Keys* K = new C1Keys;
  c1_keygen(*(Oracle*)h, *K, seed);
  return K;
}
void orc_c1_free(void* keys) { delete (C1Keys*)keys; }
// Keys in the NTT domain: which 0 s (QB), 1 pk b, 2 pk a (Q), 3 evk b, 4 evk a (QB).
int64_t orc_c1_key(void* keys, int which, uint64_t* out) {
  const C1Keys& K = *(C1Keys*)keys;
  const std::vector<uint64_t>* v[5] = {&K.s, &K.pkb, &K.pka, &K.evkb, &K.evka};
  if (which < 0 || which > 4) return -1;
  if (out) std::memcpy(out, v[which]->data(), v[which]->size() * 8);
  return (int64_t)v[which]->size();
}
void orc_hommul(void* h, uint64_t* c30, uint64_t* c31, const uint64_t* c10, const uint64_t* c11,
                const uint64_t* c20, const uint64_t* c21, int in_state, const uint64_t* evkb,
                const uint64_t* evka, int level, int rescale);
// z1, z2: slots complex values each (re, im interleaved); out: slots values;
// c3 (optional, 2 x k limbs): the HomMul output in the coefficient domain.
void orc_c1_run(void* h, void* keys, const double* z1, const double* z2, int slots, uint64_t scale, uint64_t seed,
                uint64_t nonce, double* out, uint64_t* c3) {
  const Oracle& o = *(Oracle*)h;
  const C1Keys& K = *(C1Keys*)keys;
  const int k = o.k, L = k - 1;
  const size_t qn = (size_t)k * o.N;
  std::vector<uint64_t> a0(qn), a1(qn), b0(qn), b1(qn), r0(qn), r1(qn), t(qn);
  c1_encrypt(o, K, a0.data(), a1.data(), reinterpret_cast<const std::complex<double>*>(z1), slots, scale, seed, nonce);
  c1_encrypt(o, K, b0.data(), b1.data(), reinterpret_cast<const std::complex<double>*>(z2), slots, scale, seed,
             nonce + 3);
  orc_hommul(h, r0.data(), r1.data(), a0.data(), a1.data(), b0.data(), b1.data(), TRANSSTATE, K.evkb.data(),
             K.evka.data(), k - 1, 1);
  if (c3) {
    std::memcpy(c3, r0.data(), qn * 8);
    std::memcpy(c3 + qn, r1.data(), qn * 8);
  }
  // decrypt on the L = k-1 remaining limbs: m = c0 + c1 s
  transform(o, r0.data(), 0, L, 0);
  transform(o, r1.data(), 0, L, 0);
  vec_op(o, 2, t.data(), r1.data(), K.s.data(), L);
  vec_op(o, 0, t.data(), t.data(), r0.data(), L);
  transform(o, t.data(), 0, L, 1);
  const uint64_t s2 = (uint64_t)((unsigned __int128)scale * scale / o.P[k - 1]);
  decode(o, reinterpret_cast<std::complex<double>*>(out), t.data(), slots, s2);
}
int orc_roots(void* h, uint64_t* out) {
  Oracle* o = (Oracle*)h;
  for (int r = 0; r < o->R(); ++r) out[r] = o->G[r];
  return o->R();
}
int orc_factor(void* h, int* Na, int* Nb) {
  Oracle* o = (Oracle*)h;
  *Na = o->Na;
  *Nb = o->Nb;
  return 0;
}
// Constant tables, same numbering as aoa_params_table for the RNS tables:
// 2 InvQqi [k][k], 3 Qqibi as (l, j, i), 4 InvBbi, 5 Bbiqi [i][j], 6 InvModBqi,
// 7 Bqi, 8 Invqlqi [k][k], 9 MU.
int64_t orc_table(void* h, int which, uint64_t* out) {
  Oracle* o = (Oracle*)h;
  const int k = o->k, kl = o->kl;
  std::vector<uint64_t> v;
  switch (which) {
    case 2: v = o->InvQqi; break;
    case 3:
      v.assign((size_t)k * kl * k, 0);
      for (int l = 0; l < k; ++l)
        for (int i = 0; i <= l; ++i)
          for (int j = 0; j < kl; ++j) v[((size_t)l * kl + j) * k + i] = o->Qqibi[((size_t)l * k + i) * kl + j];
      break;
    case 4: v = o->InvBbi; break;
    case 5: v = o->Bbiqi; break;
    case 6: v = o->InvModBqi; break;
    case 7: v = o->Bqi; break;
    case 8: v = o->Invqlqi; break;
    case 9: v = o->MU; break;
    default: return -1;
  }
  if (out) std::memcpy(out, v.data(), v.size() * 8);
  return (int64_t)v.size();
}
uint64_t orc_mod128(void* h, int rid, uint64_t hi, uint64_t lo) {
  u128 a;
  a.hi = hi;
  a.lo = lo;
  return ref_mod(*(Oracle*)h, a, rid);
}
uint64_t orc_mulmod(void* h, int rid, uint64_t a, uint64_t b) { return mulmod(*(Oracle*)h, a, b, rid); }
void orc_ntt(void* h, uint64_t* d, int res0, int nres, int dir) { transform(*(Oracle*)h, d, res0, nres, dir); }
void orc_eltwise(void* h, int op, uint64_t* c, const uint64_t* a, const uint64_t* b, int nres) {
  vec_op(*(Oracle*)h, op, c, a, b, nres);
}
void orc_eltwise_scalar(void* h, int op, uint64_t* c, const uint64_t* a, uint64_t b, int nres) {
  vec_op_int(*(Oracle*)h, op, c, a, b, nres);
}
void orc_conjugate(void* h, uint64_t* out, const uint64_t* in, int nres) {
  conjugation_kernel(*(Oracle*)h, out, in, nres);
}
void orc_rescale(void* h, uint64_t* a, uint64_t* b, int level) { rescale_kernel(*(Oracle*)h, a, b, level); }
void orc_modup(void* h, uint64_t* qb, int level) { modup_kernel(*(Oracle*)h, qb, level); }
void orc_moddown(void* h, uint64_t* out, const uint64_t* qb, int level) {
  moddown_kernel(*(Oracle*)h, out, qb, level);
}
void orc_automorphism(void* h, uint64_t* out, const uint64_t* in, int nres, int rot_slots) {
  Oracle* o = (Oracle*)h;
  long pw = powmod_small(5, rot_slots, 2L * o->N);
  rotation_kernel(*o, out, in, pw, nres);
}

// ckks_mul (ckks.cpp:334-358) = ckks_mul_without_rescale (:273-332) + ckks_rescale (:220-230).
// c1x/c2x: Q polys (k limbs) in state `in_state` (1 coefficient, 2 NTT); evk in
// the NTT domain (k+kl limbs).  Outputs the coefficient-domain c3 (k limbs;
// limbs 0..level-1 meaningful after the rescale).
void orc_hommul(void* h, uint64_t* c30, uint64_t* c31, const uint64_t* c10, const uint64_t* c11,
                const uint64_t* c20, const uint64_t* c21, int in_state, const uint64_t* evkb,
                const uint64_t* evka, int level, int rescale) {
  const Oracle& o = *(Oracle*)h;
  const size_t qn = (size_t)o.k * o.N, qbn = (size_t)(o.k + o.kl) * o.N;
  Poly a0 = poly_new(o, QBase), a1 = poly_new(o, QBase), b0 = poly_new(o, QBase), b1 = poly_new(o, QBase);
  std::memcpy(a0.d.data(), c10, qn * 8);
  std::memcpy(a1.d.data(), c11, qn * 8);
  std::memcpy(b0.d.data(), c20, qn * 8);
  std::memcpy(b1.d.data(), c21, qn * 8);
  a0.state = a1.state = b0.state = b1.state = in_state;
  Poly eb = poly_new(o, QBBase), ea = poly_new(o, QBBase);
  std::memcpy(eb.d.data(), evkb, qbn * 8);
  std::memcpy(ea.d.data(), evka, qbn * 8);
  eb.state = ea.state = TRANSSTATE;
  // CKKSContext temporaries start in RNSSTATE (poly_init)
  Poly axbx1 = poly_new(o, QBase), axbx2 = poly_new(o, QBase), axax = poly_new(o, QBase),
       bxbx = poly_new(o, QBase);
  Poly d2QB = poly_new(o, QBBase), dtQB0 = poly_new(o, QBBase), dtQB1 = poly_new(o, QBBase);
  Poly dtQ0 = poly_new(o, QBase), dtQ1 = poly_new(o, QBase);
  Poly r0 = poly_new(o, QBase), r1 = poly_new(o, QBase);
  poly_add_add(o, axbx1, a1, a0, axbx2, b1, b0);
  poly_binop(o, 2, axbx1, axbx1, axbx2);  // d1
  poly_binop(o, 2, bxbx, a0, b0);         // d0
  poly_binop(o, 2, axax, a1, b1);         // d2
  poly_modup(o, d2QB, axax, level);
  poly_binop(o, 2, dtQB0, d2QB, eb);
  poly_binop(o, 2, dtQB1, d2QB, ea);
  poly_moddown(o, dtQ0, dtQB0, level);
  poly_moddown(o, dtQ1, dtQB1, level);
  poly_add_add(o, r0, bxbx, dtQ0, r1, axbx1, dtQ1);
  poly_binop(o, 1, r1, r1, bxbx);
  poly_binop(o, 1, r1, r1, axax);
  if (rescale) {
    execute(o, r0, 1);  // poly_rescale INVERSEs both (poly_t.cpp:395-411)
    execute(o, r1, 1);
    rescale_kernel(o, r0.d.data(), r1.d.data(), level);
  }
  std::memcpy(c30, r0.d.data(), qn * 8);
  std::memcpy(c31, r1.d.data(), qn * 8);
}

// ckks_rotate_left (ckks.cpp:982-1009) for a coefficient-domain input; rtk in
// the NTT domain.  Output in the coefficient domain.
void orc_rotate_left(void* h, uint64_t* o0, uint64_t* o1, const uint64_t* c0, const uint64_t* c1,
                     const uint64_t* rtkb, const uint64_t* rtka, int rot_slots, int level) {
  const Oracle& o = *(Oracle*)h;
  const size_t qn = (size_t)o.k * o.N, qbn = (size_t)(o.k + o.kl) * o.N;
  const long pw = powmod_small(5, rot_slots, 2L * o.N);
  Poly out0 = poly_new(o, QBase), axax = poly_new(o, QBase);
  rotation_kernel(o, out0.d.data(), c0, pw, o.k);
  rotation_kernel(o, axax.d.data(), c1, pw, o.k);
  Poly eb = poly_new(o, QBBase), ea = poly_new(o, QBBase);
  std::memcpy(eb.d.data(), rtkb, qbn * 8);
  std::memcpy(ea.d.data(), rtka, qbn * 8);
  eb.state = ea.state = TRANSSTATE;
  Poly d2QB = poly_new(o, QBBase), dtQB0 = poly_new(o, QBBase), dtQB1 = poly_new(o, QBBase);
  Poly dtQ0 = poly_new(o, QBase), out1 = poly_new(o, QBase);
  poly_modup(o, d2QB, axax, level);
  poly_binop(o, 2, dtQB0, d2QB, eb);
  poly_binop(o, 2, dtQB1, d2QB, ea);
  poly_moddown(o, dtQ0, dtQB0, level);
  poly_moddown(o, out1, dtQB1, level);
  poly_binop(o, 0, out0, out0, dtQ0);
  std::memcpy(o0, out0.d.data(), qn * 8);
  std::memcpy(o1, out1.d.data(), qn * 8);
}


// dot_prod (manager.cu:53-67): v = sum_i a[i][t] b[i][t] with mullazy and
// add128 (common.cu:157-213, no reduction inside the sum), then mod(v, rid),
// over every residue of a QB polynomial.  a, b: k host pointers each.
void orc_dot(void* h, uint64_t* c, const uint64_t* const* a, const uint64_t* const* b, int k) {
  const Oracle& o = *(Oracle*)h;
  const size_t n = (size_t)o.R() * o.N;
  for (size_t t = 0; t < n; ++t) {
    const int rid = (int)(t / o.N);
    u128 v;
    v.hi = v.lo = 0;
    for (int i = 0; i < k; ++i) {
      const u128 m = mullazy(a[i][t], b[i][t]);
      u128 z;
      z.lo = v.lo + m.lo;
      z.hi = v.hi + m.hi + (z.lo < v.lo);
      v = z;
    }
    c[t] = ref_mod(o, v, rid);
  }
}

// dr2 (manager.cu:370-395) over every residue of a QB polynomial: ct21 =
// ct01 ct11, ct22 = (ct01 ct12)(ct02 ct11), ct23 = ct02 ct12, each a mulmod.
void orc_dr2(void* h, uint64_t* c21, uint64_t* c22, uint64_t* c23, const uint64_t* a01, const uint64_t* a02,
             const uint64_t* a11, const uint64_t* a12) {
  const Oracle& o = *(Oracle*)h;
  const size_t n = (size_t)o.R() * o.N;
  for (size_t t = 0; t < n; ++t) {
    const int rid = (int)(t / o.N);
    c21[t] = mulmod(o, a01[t], a11[t], rid);
    c22[t] = mulmod(o, mulmod(o, a01[t], a12[t], rid), mulmod(o, a02[t], a11[t], rid), rid);
    c23[t] = mulmod(o, a02[t], a12[t], rid);
  }
}

// ckks_rotate_right (ckks.cpp:953-980): rotate_slots_right is rotate_slots_left
// by N/2 - rotSlots (encoder.cu:356-363), then the same key switch with the
// right rotation key.
void orc_rotate_right(void* h, uint64_t* o0, uint64_t* o1, const uint64_t* c0, const uint64_t* c1,
                      const uint64_t* rtkb, const uint64_t* rtka, int rot_slots, int level) {
  const Oracle& o = *(Oracle*)h;
  orc_rotate_left(h, o0, o1, c0, c1, rtkb, rtka, (o.N >> 1) - rot_slots, level);
}

}  // extern "C"
