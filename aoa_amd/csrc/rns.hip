// RNS kernels for gfx950: element-wise ops, the HomMul tensor product, fast
// basis conversion (ModUp / ModDown), rescale and the Galois automorphism.
//
// Semantics follow the reference kernels (aoa-ntt/src/cuda/manager.cu:90-367,
// aoa-ntt/src/cuda/htrans/ntt.cu:479-667, aoa-ntt/src/ckks/encoder.cu:300-322);
// every product is exact except the one ModDown cross-base product, which uses
// the reference's Barrett formula so that residues stay bit-identical.
#include <algorithm>

#include "device_common.hpp"
#include "kernels.hpp"

namespace aoa {

__device__ __forceinline__ uint64_t reduce128(uint64_t hi, uint64_t lo, const DevPrime& d) {
  uint64_t a = mul_shoup(hi, d.r64, d.r64_q, d.p);
  uint64_t b = barrett_reduce(0, lo, d.p, d.mu, (int)d.nb);
  return add_mod(a, b, d.p);
}

// (hi, lo) = s0 + s1 2^28 + s2 2^56 as a 128-bit value.
__device__ __forceinline__ void join_28(uint64_t s0, uint64_t s1, uint64_t s2, uint64_t* hi, uint64_t* lo) {
  uint64_t l = s0 + (s1 << 28);
  uint64_t h = (s1 >> 36) + (l < s0 ? 1u : 0u);
  const uint64_t t = l + (s2 << 56);
  h += (s2 >> 8) + (t < l ? 1u : 0u);
  *hi = h;
  *lo = t;
}

// ------------------------------------------------------------ element-wise --

// Two elements per thread; a 256-thread block covers 512 elements of one limb.
__global__ __launch_bounds__(256) void k_eltwise(DevParams P, int op, uint64_t* out, const uint64_t* a,
                                                 const uint64_t* b, const uint64_t* c, int qlimbs, int prime0) {
  const size_t e0 = (size_t)blockIdx.x * 512;
  const int li = (int)(e0 >> P.logN);
  const int pr = li < qlimbs ? prime0 + li : P.k + (li - qlimbs);
  const DevPrime d = P.primes[pr];
  const size_t e = e0 + (size_t)threadIdx.x * 2;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const size_t i = e + t;
    uint64_t r;
    switch (op) {
      case ELT_ADD: r = add_mod(a[i], b[i], d.p); break;
      case ELT_SUB: r = sub_mod(a[i], b[i], d.p); break;
      case ELT_MUL: r = mul_var(a[i], b[i], d.p, d.mu, d.mu2, (int)d.nb); break;
      case ELT_NEG: r = neg_mod(a[i], d.p); break;
      case ELT_MULADD: r = add_mod(mul_var(a[i], b[i], d.p, d.mu, d.mu2, (int)d.nb), c[i], d.p); break;
      default: r = 0; break;
    }
    out[i] = r;
  }
}

void launch_eltwise(const DevParams& P, int op, uint64_t* out, const uint64_t* a, const uint64_t* b,
                    const uint64_t* c, int nlimbs, int qlimbs, int prime0, hipStream_t st) {
  const size_t total = (size_t)nlimbs * P.N;
  k_eltwise<<<dim3((unsigned)(total / 512)), 256, 0, st>>>(P, op, out, a, b, c, qlimbs, prime0);
}

// out = a op b for a 64-bit scalar b used as given (not reduced), with the
// reference's own formulas (aoa-ntt/src/cuda/htrans/ntt.cu:531-558,
// aoa-ntt/include/newckks/cuda/htrans/common.h:83-90, common.cu:174-193):
// add and subtract by one conditional correction in wrapping 64-bit
// arithmetic, multiply by its Barrett reduction of the 128-bit product.
// These are what ckks_add / ckks_mul by a constant feed to the residues, so
// the literal forms keep them bit-identical even when b exceeds the prime.
__global__ __launch_bounds__(256) void k_eltwise_scalar(DevParams P, int op, uint64_t* out, const uint64_t* a,
                                                        uint64_t b, int prime0) {
  const size_t e0 = (size_t)blockIdx.x * 512;
  const int li = (int)(e0 >> P.logN);
  const DevPrime d = P.primes[prime0 + li];
  const size_t e = e0 + (size_t)threadIdx.x * 2;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const uint64_t x = a[e + t];
    uint64_t r;
    if (op == ELT_ADD) {
      const uint64_t s = x + b;
      r = s >= d.p ? s - d.p : s;
    } else if (op == ELT_SUB) {
      r = x < b ? x + (d.p - b) : x - b;
    } else {
      r = ref_barrett(mulhi64(x, b), x * b, d.p, d.mu, (int)d.nb);
    }
    out[e + t] = r;
  }
}

void launch_eltwise_scalar(const DevParams& P, int op, uint64_t* out, const uint64_t* a, uint64_t b, int nlimbs,
                           int prime0, hipStream_t st) {
  const size_t total = (size_t)nlimbs * P.N;
  k_eltwise_scalar<<<dim3((unsigned)(total / 512)), 256, 0, st>>>(P, op, out, a, b, prime0);
}

// Public-key encryption's combination (aoa-ntt/src/ckks/ckks.cpp:87-105),
// NTT domain, limbs 0..nlimbs-1: c0 = pk_b u + e0 + m, c1 = pk_a u + e1.
__global__ __launch_bounds__(256) void k_encrypt_combine(DevParams P, uint64_t* c0, uint64_t* c1,
                                                         const uint64_t* pkb, const uint64_t* pka,
                                                         const uint64_t* u, const uint64_t* e0,
                                                         const uint64_t* e1, const uint64_t* m) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  const DevPrime d = P.primes[(int)(e >> P.logN)];
  const uint64_t uu = u[e];
  const uint64_t b = add_mod(mul_var(pkb[e], uu, d.p, d.mu, d.mu2, (int)d.nb), e0[e], d.p);
  c0[e] = add_mod(b, m[e], d.p);
  c1[e] = add_mod(mul_var(pka[e], uu, d.p, d.mu, d.mu2, (int)d.nb), e1[e], d.p);
}

void launch_encrypt_combine(const DevParams& P, uint64_t* c0, uint64_t* c1, const uint64_t* pkb,
                            const uint64_t* pka, const uint64_t* u, const uint64_t* e0, const uint64_t* e1,
                            const uint64_t* m, int nlimbs, hipStream_t st) {
  const size_t total = (size_t)nlimbs * P.N;
  k_encrypt_combine<<<dim3((unsigned)(total / 256)), 256, 0, st>>>(P, c0, c1, pkb, pka, u, e0, e1, m);
}

// Conjugation in the NTT domain: out[cid] = in[N-1-cid] per limb
// (aoa-ntt/src/ckks/encoder.cu:391-401).  With the bit-reversed evaluation
// order, reversing the index maps the evaluation at psi^e to psi^-e, i.e. the
// automorphism X -> X^-1.
__global__ __launch_bounds__(256) void k_conjugate(DevParams P, uint64_t* out, const uint64_t* in) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t base = (e >> P.logN) << P.logN;
  const size_t cid = e - base;
  out[e] = in[base + ((size_t)P.N - 1 - cid)];
}

void launch_conjugate(const DevParams& P, uint64_t* out, const uint64_t* in, int nlimbs, hipStream_t st) {
  const size_t total = (size_t)nlimbs * P.N;
  k_conjugate<<<dim3((unsigned)(total / 256)), 256, 0, st>>>(P, out, in);
}

// d0 = a0 b0, d2 = a1 b1, d1 = (a0 + a1)(b0 + b1) - d0 - d2 (= a0 b1 + a1 b0),
// two adjacent elements per thread of one limb (pointers at element e of the
// limb, e even), moved as 16-byte vectors.
struct alignas(16) u64x2 {
  uint64_t v[2];
};

__device__ __forceinline__ void tensor_body(const DevPrime& d, uint64_t* d0, uint64_t* d1, uint64_t* d2,
                                            const uint64_t* a0, const uint64_t* a1, const uint64_t* b0,
                                            const uint64_t* b1, size_t e) {
  const u64x2 X0 = *reinterpret_cast<const u64x2*>(a0 + e), X1 = *reinterpret_cast<const u64x2*>(a1 + e);
  const u64x2 Y0 = *reinterpret_cast<const u64x2*>(b0 + e), Y1 = *reinterpret_cast<const u64x2*>(b1 + e);
  u64x2 R0, R1, R2;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const uint64_t x0 = X0.v[t], x1 = X1.v[t], y0 = Y0.v[t], y1 = Y1.v[t];
    const uint64_t p0 = mul_var(x0, y0, d.p, d.mu, d.mu2, (int)d.nb);
    const uint64_t p2 = mul_var(x1, y1, d.p, d.mu, d.mu2, (int)d.nb);
    const uint64_t s = mul_var(add_mod(x0, x1, d.p), add_mod(y0, y1, d.p), d.p, d.mu, d.mu2, (int)d.nb);
    R0.v[t] = p0;
    R2.v[t] = p2;
    R1.v[t] = sub_mod(sub_mod(s, p0, d.p), p2, d.p);
  }
  *reinterpret_cast<u64x2*>(d0 + e) = R0;
  *reinterpret_cast<u64x2*>(d1 + e) = R1;
  *reinterpret_cast<u64x2*>(d2 + e) = R2;
}

__global__ __launch_bounds__(256) void k_tensor(DevParams P, uint64_t* d0, uint64_t* d1, uint64_t* d2,
                                                const uint64_t* a0, const uint64_t* a1,
                                                const uint64_t* b0, const uint64_t* b1) {
  const size_t e0 = (size_t)blockIdx.x * 512;
  const int li = (int)(e0 >> P.logN);
  tensor_body(P.primes[li], d0, d1, d2, a0, a1, b0, b1, e0 + (size_t)threadIdx.x * 2);
}

void launch_tensor(const DevParams& P, uint64_t* d0, uint64_t* d1, uint64_t* d2, const uint64_t* a0,
                   const uint64_t* a1, const uint64_t* b0, const uint64_t* b1, int L, hipStream_t st) {
  const size_t total = (size_t)L * P.N;
  k_tensor<<<dim3((unsigned)(total / 512)), 256, 0, st>>>(P, d0, d1, d2, a0, a1, b0, b1);
}

__global__ __launch_bounds__(256) void k_tensor_d(DevParams P, const TensorDesc* dd) {
  const int per = P.N / 512;
  const TensorDesc t = dd[blockIdx.x / per];
  const size_t e = (size_t)(blockIdx.x % per) * 512 + (size_t)threadIdx.x * 2;
  tensor_body(P.primes[t.prime], glob(t.d0), glob(t.d1), glob(t.d2), glob(t.a0), glob(t.a1), glob(t.b0), glob(t.b1), e);
}

void launch_tensor_desc(const DevParams& P, const TensorDesc* d, int ndesc, hipStream_t st) {
  k_tensor_d<<<dim3((unsigned)(ndesc * (P.N / 512))), 256, 0, st>>>(P, d);
}

// Basis conversion workgroups: 64 consecutive coefficients x 4 limb groups
// (wave g of the workgroup handles limbs g, g+4, g+8, ...), so a launch has
// N/64 workgroups per polynomial and each wave reads 512 contiguous bytes.
constexpr int BC_COEF = 64;
constexpr int BC_GROUPS = 4;

// ----------------------------------------------------------------- ModUp ----
// For j < K: out_j = sum_{i<=l} ([x_i (Q_l/q_i)^-1]_{q_i} mod b_j) (Q_l/q_i mod b_j)
// mod b_j (aoa-ntt/src/cuda/manager.cu:180-223).  The per-i values are shared
// through LDS; the sum over i is accumulated exactly in 128 bits and reduced
// once, which equals the reference's term-by-term modular sum.

// Sums over i <= level of v_i (Q_l/q_i mod b_j) for NL special limbs jl[],
// v_i = [x_i (Q_l/q_i)^-1]_{q_i} from LDS, each reduced mod b_j first where
// q_i > b_j (always for q0).  With v, w < 2^56 split at bit 28 and
// vs = vl + vh, ws = wl + wh: v w = vl wl + (vs ws - vl wl - vh wh) 2^28 +
// vh wh 2^56, three 32x32->64 multiply-adds per term into 64-bit sums that
// cannot overflow for up to 64 terms, one lazy reduction per limb
// (reduce_sums_lazy; the host checks its bound).  The halves of w come split
// from the mu_terms table, the NL limbs' entries side by side, so that a term
// takes one scalar load and no scalar ALU work.
template <int NL, bool LAZY>
__device__ __forceinline__ void modup_pass(const DevParams& P, const uint64_t* tl, int c, const int* jl, int level,
                                           uint64_t* outB, size_t cid) {
  constexpr uint32_t M28 = (1u << 28) - 1;
  const int L = level + 1, k = P.k;
  // The pass's limbs jl[n] = jl[0] + 4n are consecutive entries of mu_terms:
  // (level, i, pos0 + n) sits at e0 + i kp 4 + 4n.
  const int kp = (P.K + 15) & ~15;
  const int t0 = jl[0] / BC_GROUPS;
  const int pos0 = (t0 / 4) * 16 + 4 * (jl[0] % BC_GROUPS) + t0 % 4;
  const cptr<uint32_t> e0 = as_const(P.mu_terms) + ((size_t)level * (k + 1) * kp + pos0) * 4;
  DevPrime d[NL];
  uint32_t red = 0;
  uint64_t s0[NL], s2[NL], sm[NL];
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    d[n] = ld_prime(P.primes, k + jl[n]);
    red |= d[n].modup_reduce;
    s0[n] = s2[n] = sm[n] = 0;
  }
  auto acc = [&](int n, uint64_t v, cptr<uint32_t> e) {
    const uint32_t vl = (uint32_t)v & M28, vh = (uint32_t)(v >> 28);
    s0[n] += (uint64_t)vl * e[0];
    s2[n] += (uint64_t)vh * e[1];
    sm[n] += (uint64_t)(vl + vh) * e[2];
  };
  {
    // q0's value v < 2^63, wider than every special prime: v = lo + hi 2^56,
    // lo against (Q_l/q_0) mod b_j and hi (7 bits) against 2^56 (Q_l/q_0)
    // mod b_j (row k), hi taken as the low half of the split with no high half
    const uint64_t v = tl[c];
    const uint32_t hi = (uint32_t)(v >> 56);
    const cptr<uint32_t> eh = e0 + (size_t)k * kp * 4;
#pragma unroll
    for (int n = 0; n < NL; ++n) {
      acc(n, v & ((1ull << 56) - 1), e0 + 4 * n);
      s0[n] += (uint64_t)hi * eh[4 * n];
      sm[n] += (uint64_t)hi * eh[4 * n + 2];
    }
  }
  if (red) {
    for (int i = 1; i < L; ++i) {
      const uint64_t v = tl[i * BC_COEF + c];
      const uint64_t qi = ld_prime_p(P.primes, i);
      const cptr<uint32_t> e = e0 + (size_t)i * kp * 4;
#pragma unroll
      for (int n = 0; n < NL; ++n)
        acc(n, qi > d[n].p ? barrett_reduce(0, v, d[n].p, d[n].mu, (int)d[n].nb) : v, e + 4 * n);
    }
  } else {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
#pragma unroll 2
    for (int i = 1; i < L; ++i) {
      const uint64_t v = tl[i * BC_COEF + c];
      const uint32_t vl = (uint32_t)v & M28, vh = (uint32_t)(v >> 28), vs = vl + vh;
      // whole 16-byte entries, so that the NL entries load as one block
      const cptr<u32x4> e = (cptr<u32x4>)(e0 + (size_t)i * kp * 4);
#pragma unroll
      for (int n = 0; n < NL; ++n) {
        const u32x4 w = e[n];
        s0[n] += (uint64_t)vl * w.x;
        s2[n] += (uint64_t)vh * w.y;
        sm[n] += (uint64_t)vs * w.z;
      }
    }
  }
  // one lazy reduction per limb (below 5 b_j), then down to [0, b_j), or with
  // LAZY to [0, 4 b_j), which the forward NTT's lazy butterflies take
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    const uint64_t b = d[n].p;
    uint64_t r = reduce_sums_lazy(s0[n], sm[n] - s0[n] - s2[n], s2[n], b, as_const(P.md_red)[2 * (k + jl[n])],
                                  (int)d[n].nb);
    for (uint64_t m = 4 * b; m >= (LAZY ? 4 * b : b); m >>= 1) r = r >= m ? r - m : r;
    outB[(size_t)jl[n] * P.N + cid] = r;
  }
}

template <bool LAZY>
__device__ __forceinline__ void modup_body(const DevParams& P, const uint64_t* x, uint64_t* outB, int level,
                                           int blk, uint64_t* tl) {
  const int L = level + 1;
  const size_t N = (size_t)P.N;
  const int c = threadIdx.x % BC_COEF;
  // wave-uniform limb group: keeps limb indices and constant loads scalar
  const int g = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / BC_COEF));
  const size_t cid = (size_t)blk * BC_COEF + c;
  const tw2* iq = reinterpret_cast<const tw2*>(P.inv_qhat) + (size_t)level * P.k;
  for (int i = g; i < L; i += BC_GROUPS) {
    const uint64_t p = P.primes[i].p;
    const tw2 w = iq[i];
    tl[i * BC_COEF + c] = mul_shoup(x[(size_t)i * N + cid], w.w, w.q, p);
  }
  __syncthreads();
  // This wave's special limbs g, g+4, ... in passes of four, two or one.
  int j = g;
  while (j < P.K) {
    const int left = (P.K - j + BC_GROUPS - 1) / BC_GROUPS;
    const int jl[4] = {j, j + BC_GROUPS, j + 2 * BC_GROUPS, j + 3 * BC_GROUPS};
    if (left >= 4) {
      modup_pass<4, LAZY>(P, tl, c, jl, level, outB, cid);
      j += 4 * BC_GROUPS;
    } else if (left >= 2) {
      modup_pass<2, LAZY>(P, tl, c, jl, level, outB, cid);
      j += 2 * BC_GROUPS;
    } else {
      modup_pass<1, LAZY>(P, tl, c, jl, level, outB, cid);
      j += BC_GROUPS;
    }
  }
}

// Seven waves per SIMD (72 registers, one spilled): 3 % faster than six.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_modup(DevParams P, const uint64_t* x, uint64_t* outB, int level) {
  extern __shared__ uint64_t tl[];  // [L][64]
  modup_body<false>(P, x, outB, level, blockIdx.x, tl);
}

void launch_modup(const DevParams& P, const uint64_t* x, uint64_t* outB, int level, hipStream_t st) {
  const size_t lds = (size_t)(level + 1) * BC_COEF * sizeof(uint64_t);
  k_modup<<<dim3(P.N / BC_COEF), 256, lds, st>>>(P, x, outB, level);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_modup_d(DevParams P, const ModUpDesc* dd) {
  extern __shared__ uint64_t tl[];
  const int per = P.N / BC_COEF;
  const ModUpDesc u = dd[blockIdx.x / per];
  // the batched form's outputs go only to the forward NTT: left in [0, 4 b_j)
  modup_body<true>(P, glob(u.x), glob(u.outB), u.level, blockIdx.x % per, tl);
}

void launch_modup_desc(const DevParams& P, const ModUpDesc* d, int ndesc, int max_level, hipStream_t st) {
  const size_t lds = (size_t)(max_level + 1) * BC_COEF * sizeof(uint64_t);
  k_modup_d<<<dim3((unsigned)(ndesc * (P.N / BC_COEF))), 256, lds, st>>>(P, d);
}

// --------------------------------------------------------------- ModDown ----
// out_i = (x_i - sum_j ref([x_bj (B/b_j)^-1]_bj * (B/b_j mod q_i))) B^-1 mod q_i
// (aoa-ntt/src/cuda/manager.cu:312-341), optionally + add_i (rotations, and
// HomMul with the single-workgroup transform), optionally rescaled:
// out_i <- (out_i - out_l) q_l^-1 mod q_i (manager.cu:90-118).
//
// ref(a*c) is the reference's Barrett reduction of the product (ref_barrett).
// With A = a*c - qhat*p = delta*p + r (r the exact remainder, delta the
// Barrett quotient error, 0 <= delta*eps < 2^nb, eps = 2^nb - p), its result
// is r - eps when eps <= r < delta*eps and r otherwise; delta_max (from the
// largest product and 2^(2nb) mod p, params.cpp) bounds delta.  Every limb
// takes exact sums of its products (moddown_conv for primes below 2^57,
// moddown_wide for the 63-bit q0), corrected term by term where the screen
// flags a product; moddown_exact skips the screen and its corrections.

// Sums over j in [j0, j1) of the ModDown products for the NL limbs
// i0, i0 + 4, ... (primes below 2^57), reduced mod q_i up to four multiples
// of q_i too many (in [0, 5 q_i)), into res[].  Each
// product a c (a < 2^56 from LDS, c = (B/b_j) mod q_i) is accumulated exactly
// from the 28-bit halves, a c = al cl + (as cs - al cl - ah ch) 2^28 + ah ch 2^56
// (as = al + ah, cs = cl + ch): three 32x32->64 multiply-adds into 64-bit
// sums that cannot overflow for up to 64 terms, one 128-bit reduction per
// limb.  Exact sums equal the reference's term-by-term modular sum whenever
// no term deviates.  SCREEN runs moddown_screen on every product; when a
// product of some lane passes it (one term in 10 000 to 300 000 at 48-bit
// q_i, growing with 2^48 - q_i) the wave takes a branch in which the lanes
// whose product passed add ref(a c) - (a c mod q_i) + q_i (in (0, 2 q_i)) to
// their low sum (and to sm, so the middle sum keeps its value): every other
// term lies outside the deviation window, where the two agree.  The pass/fail
// masks come straight from the compares and are not carried across
// iterations.  fold(i, &z) may add one more term z (-B mod q_i), z < 2^57
// (the caller's addend and the rescale's dropped limb), so that the caller
// finishes the limb with one Shoup product.  The host checks that K products,
// the folded term and K corrections stay below 2^(62+nb) (build_host_params).
template <int NL, bool SCREEN, class Fold>
__device__ __forceinline__ void moddown_conv(const DevParams& P, const uint64_t* aux, int c, int i0,
                                             int j0, int j1, uint64_t* res, Fold fold) {
  constexpr uint32_t M28 = (1u << 28) - 1;
  constexpr int TW = (int)(sizeof(MdTerm) / 4);
  const int k = P.k, K = P.K;
  const cptr<uint32_t> terms = as_const(P.md_terms);
  const tw2* bq = reinterpret_cast<const tw2*>(P.bhat_mod_q);
  uint64_t s0[NL], s2[NL], sm[NL];
  uint32_t thr[NL];
  // moddown_screen's bias, opaque to the compiler so that it stays the addend
  // of the first multiply-add instead of becoming an addition of its own
  uint64_t bias = 2ull << 32;
  asm volatile("" : "+v"(bias));
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    s0[n] = s2[n] = sm[n] = 0;
    thr[n] = as_const(P.md_thr)[i0 + n * BC_GROUPS];
  }
  for (int j = j0; j < j1; ++j) {
    const uint64_t a = aux[j * BC_COEF + c];
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    const uint32_t al = a0 & M28, ah = (uint32_t)(a >> 28), as = al + ah;
    const cptr<uint32_t> t = terms + ((size_t)j * k + i0) * TW;
    bool h[NL];
    uint64_t any = 0;
#pragma unroll
    for (int n = 0; n < NL; ++n) {
      const cptr<uint32_t> e = t + n * BC_GROUPS * TW;
      s0[n] += (uint64_t)al * e[0];
      s2[n] += (uint64_t)ah * e[1];
      sm[n] += (uint64_t)as * e[2];
      if (SCREEN) {
        h[n] = moddown_screen(a0, a1, e[3], e[4], e[5], bias) <= thr[n];
        any |= __builtin_amdgcn_ballot_w64(h[n]);
      }
    }
    if (SCREEN && any) {
      // j made opaque here so that the compiler does not keep running
      // pointers into bq across the loop for this rarely taken branch
      int jo = j;
      asm volatile("" : "+s"(jo));
#pragma unroll
      for (int n = 0; n < NL; ++n) {
        if (h[n]) {
          const int i = i0 + n * BC_GROUPS;
          const DevPrime dq = ld_prime(P.primes, i);
          const tw2 w = ld_tw2(bq + (size_t)i * K, jo);
          const uint64_t x = mul_shoup(a, w.w, w.q, dq.p);
          const uint64_t f = ref_barrett(mulhi64(a, w.w), a * w.w, dq.p, dq.mu, (int)dq.nb);
          const uint64_t d = f + dq.p - x;
          s0[n] += d;
          sm[n] += d;  // leaves the middle sum sm - s0 - s2 as it was
        }
      }
    }
  }
#pragma unroll
  for (int n = 0; n < NL; ++n) {
    const int i = i0 + n * BC_GROUPS;
    uint64_t z;
    if (fold(i, &z)) {
      // one more term z (-B mod q_i), z < 2^57: halves below 2^29
      const cptr<uint32_t> e = as_const(P.md_negb) + 4 * i;
      const uint32_t zl = (uint32_t)z & M28, zh = (uint32_t)(z >> 28), zs = zl + zh;
      s0[n] += (uint64_t)zl * e[0];
      s2[n] += (uint64_t)zh * e[1];
      sm[n] += (uint64_t)zs * e[2];
    }
    const cptr<uint64_t> red = as_const(P.md_red) + 2 * i;
    res[n] = reduce_sums_lazy(s0[n], sm[n] - s0[n] - s2[n], s2[n], ld_prime_p(P.primes, i), red[0],
                              ld_prime_nb(P.primes, i));
  }
}

template <int NL, class Fold>
__device__ __forceinline__ void moddown_convs(const DevParams& P, const uint64_t* aux, int c, int i0,
                                              int j0, int j1, uint64_t* res, Fold fold) {
  if (P.moddown_exact)
    moddown_conv<NL, false>(P, aux, c, i0, j0, j1, res, fold);
  else
    moddown_conv<NL, true>(P, aux, c, i0, j0, j1, res, fold);
}

// Sum over j in [j0, j1) of the ModDown products of limb i of a prime wider
// than 2^56 (the 63-bit q0), as the reference reduces them, canonical.  With
// c = (B/b_j) mod q_i split at bits 28 and 56 (c0, c1 < 2^28, c2 < 2^7; the
// md_terms entry holds c0, c1, c2 and the screen words) and a < 2^56 at bit
// 28: six 32x32->64 multiply-adds per term into four 64-bit sums
// (wide_accumulate).  SCREEN as in moddown_conv: the lanes whose product
// passes add ref(a c) - (a c mod q_i) + q_i (below 2^64) in 28-bit pieces
// (wide_correct).  The total, below q_i (K b_max + 2K) <= q_i 2^nb (the host
// checks it), is reduced once with Barrett (wide_reduce).
template <bool SCREEN>
__device__ __forceinline__ uint64_t moddown_wide(const DevParams& P, const uint64_t* aux, int c, int i, int j0,
                                                 int j1) {
  constexpr int TW = (int)(sizeof(MdTerm) / 4);
  const int k = P.k, K = P.K;
  const cptr<uint32_t> terms = as_const(P.md_terms) + (size_t)i * TW;
  const uint32_t thr = as_const(P.md_thr)[i];
  uint64_t bias = 2ull << 32;
  asm volatile("" : "+v"(bias));
  uint64_t s[4] = {0, 0, 0, 0};
  for (int j = j0; j < j1; ++j) {
    const uint64_t a = aux[j * BC_COEF + c];
    const cptr<uint32_t> e = terms + (size_t)j * k * TW;
    wide_accumulate(s, a, e[0], e[1], e[2]);
    if (SCREEN) {
      const bool h = moddown_screen((uint32_t)a, (uint32_t)(a >> 32), e[3], e[4], e[5], bias) <= thr;
      if (__builtin_amdgcn_ballot_w64(h) && h) {
        const DevPrime dq = ld_prime(P.primes, i);
        const tw2 w = ld_tw2(reinterpret_cast<const tw2*>(P.bhat_mod_q) + (size_t)i * K, j);
        const uint64_t x = mul_shoup(a, w.w, w.q, dq.p);
        const uint64_t f = ref_barrett(mulhi64(a, w.w), a * w.w, dq.p, dq.mu, (int)dq.nb);
        wide_correct(s, f + dq.p - x);
      }
    }
  }
  const DevPrime dq = ld_prime(P.primes, i);
  return wide_reduce(s, dq.p, dq.mu, (int)dq.nb);
}

// No folded term.
struct NoFold {
  __device__ bool operator()(int, uint64_t*) const { return false; }
};

// bscaled: inB's limbs already carry (B/b_j)^-1 and lie in [0, 2 b_j).
__device__ __forceinline__ void moddown_body(const DevParams& P, const uint64_t* inQ, const uint64_t* inB,
                                             const uint64_t* add, uint64_t* out, int level, int rescale,
                                             int bscaled, int blk, uint64_t* aux) {
  const size_t N = (size_t)P.N;
  const int c = threadIdx.x % BC_COEF;
  // wave-uniform limb group: keeps limb indices and constant loads scalar
  const int g = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / BC_COEF));
  const size_t cid = (size_t)blk * BC_COEF + c;
  const int K = P.K, k = P.k;
  const tw2* ib = reinterpret_cast<const tw2*>(P.inv_bhat);
  if (bscaled) {
    for (int j = g; j < K; j += BC_GROUPS) {
      const uint64_t bp = P.primes[k + j].p;
      const uint64_t x = inB[(size_t)j * N + cid];
      aux[j * BC_COEF + c] = x >= bp ? x - bp : x;
    }
  } else {
    for (int j = g; j < K; j += BC_GROUPS) {
      const uint64_t bp = P.primes[k + j].p;
      const tw2 w = ib[j];
      aux[j * BC_COEF + c] = mul_shoup(inB[(size_t)j * N + cid], w.w, w.q, bp);
    }
  }
  __syncthreads();
  const tw2* ibq = reinterpret_cast<const tw2*>(P.inv_b_mod_q);
  // Sum over j in [j0, j1) of the reference's products for limb i of a prime
  // wider than 56 bits (q0), canonical.
  auto conv = [&](int i, int j0, int j1) -> uint64_t {
    return P.moddown_exact ? moddown_wide<false>(P, aux, c, i, j0, j1) : moddown_wide<true>(P, aux, c, i, j0, j1);
  };
  // (x_i - acc) B^-1 (+ add_i) mod q_i from the sum acc of limb i, which is
  // below md_red[i][1] (5 q_i, or q0 for the canonical q0 sums).
  auto finish = [&](int i, uint64_t acc) -> uint64_t {
    const uint64_t p = ld_prime_p(P.primes, i);
    const tw2 bi = ld_tw2(ibq, i);
    const uint64_t y = inQ[(size_t)i * N + cid] + as_const(P.md_red)[2 * i + 1] - acc;
    uint64_t v = mul_shoup(y, bi.w, bi.q, p);
    if (add) v = add_mod(v, add[(size_t)i * N + cid], p);
    return v;
  };
  // x_i + md_red[i][1] - acc: (x_i - acc) mod q_i as a value below 2^64
  auto lift = [&](int i, uint64_t acc) -> uint64_t {
    return inQ[(size_t)i * N + cid] + as_const(P.md_red)[2 * i + 1] - acc;
  };
  // This wave's limbs g, g+4, ... below iend, in passes of four, three, two or one.
  // A limb of a prime wider than 56 bits (q0) is summed exactly by conv
  // (canonical) and written by wide(i, acc); the others take the lazy sums of
  // moddown_convs, with the term z (-B) when fold(i, &z) says so, and are
  // written by narrow(i, acc).
  // q0's sum (a 63-bit prime's terms cost about 1.4 times a narrow limb's)
  // is split over the four waves, as the rescale's top limb is below, so that
  // the wave that writes limb 0 carries no more than the others: the partial
  // sums meet in LDS and acc0 is their total, valid after the next barrier.
  uint64_t* part = aux + (size_t)K * BC_COEF;            // [4][64]: the top limb's partial sums
  uint64_t* part0 = part + (size_t)BC_GROUPS * BC_COEF;  // [4][64]: q0's
  const int jq = (K + BC_GROUPS - 1) / BC_GROUPS;
  const int j0 = g * jq, j1 = min(K, j0 + jq);
  const bool wide0 = ld_prime_nb(P.primes, 0) > 56;
  if (wide0) part0[g * BC_COEF + c] = j1 > j0 ? conv(0, j0, j1) : 0;
  auto acc0 = [&]() {
    const uint64_t q0 = ld_prime_p(P.primes, 0);
    uint64_t a = 0;
#pragma unroll
    for (int t = 0; t < BC_GROUPS; ++t) a = add_mod(a, part0[t * BC_COEF + c], q0);
    return a;
  };
  auto limbs = [&](int iend, auto wide, auto fold, auto narrow) {
    int i = g;
    while (i < iend) {
      const int left = (iend - i + BC_GROUPS - 1) / BC_GROUPS;
      if (ld_prime_nb(P.primes, i) > 56) {
        wide(i, i == 0 ? acc0() : conv(i, 0, K));
        i += BC_GROUPS;
        continue;
      }
      int li[4] = {i, i + BC_GROUPS, i + 2 * BC_GROUPS, i + 3 * BC_GROUPS};
      // the first n limbs of li all have primes below 2^57
      auto is_narrow = [&](int n) {
        bool ok = true;
        for (int t = 1; t < n; ++t) ok &= ld_prime_nb(P.primes, li[t]) <= 56;
        return ok;
      };
      uint64_t r[4];
      if (left >= 4 && is_narrow(4)) {
        moddown_convs<4>(P, aux, c, i, 0, K, r, fold);
        for (int n = 0; n < 4; ++n) narrow(li[n], r[n]);
        i += 4 * BC_GROUPS;
      } else if (left == 3 && is_narrow(3)) {
        moddown_convs<3>(P, aux, c, i, 0, K, r, fold);
        for (int n = 0; n < 3; ++n) narrow(li[n], r[n]);
        i += 3 * BC_GROUPS;
      } else if (left >= 2 && is_narrow(2)) {
        moddown_convs<2>(P, aux, c, i, 0, K, r, fold);
        for (int n = 0; n < 2; ++n) narrow(li[n], r[n]);
        i += 2 * BC_GROUPS;
      } else {
        moddown_convs<1>(P, aux, c, i, 0, K, r, fold);
        narrow(i, r[0]);
        i += BC_GROUPS;
      }
    }
  };
  if (!rescale) {
    if (wide0) __syncthreads();
    // add_i joins the sums as z = add_i: (x_i - acc + add_i B) B^-1 = (x_i - acc) B^-1 + add_i
    limbs(
        level + 1, [&](int i, uint64_t acc) { out[(size_t)i * N + cid] = finish(i, acc); },
        [&](int i, uint64_t* z) {
          if (!add) return false;
          *z = add[(size_t)i * N + cid];
          return true;
        },
        [&](int i, uint64_t acc) {
          const tw2 bi = ld_tw2(ibq, i);
          out[(size_t)i * N + cid] = mul_shoup(lift(i, acc), bi.w, bi.q, ld_prime_p(P.primes, i));
        });
    return;
  }
  // Rescale needs limb `level` everywhere: its sum over j is split across the
  // four groups and combined through LDS.
  {
    uint64_t r = 0;
    if (ld_prime_nb(P.primes, level) > 56) {
      r = conv(level, j0, j1 > j0 ? j1 : j0);
    } else if (j1 > j0) {
      moddown_convs<1>(P, aux, c, level, j0, j1, &r, NoFold());
    }
    part[g * BC_COEF + c] = r;
  }
  __syncthreads();
  const uint64_t ml = as_const(P.md_red)[2 * level + 1];
  uint64_t acc = 0;
  // four partial sums below ml each: their sum stays below 2^64 and below ml
  // once reduced by the multiples of pl that ml allows
#pragma unroll
  for (int t = 0; t < BC_GROUPS; ++t) acc = add_mod(acc, part[t * BC_COEF + c], ml);
  const uint64_t top = finish(level, acc);
  // out_i = ((x_i - acc_i) B^-1 + add_i - top) q_l^-1
  //       = (x_i - acc_i + z_i B) (B^-1 q_l^-1),  z_i = q_i + add_i - top  (mod q_i)
  // with top < q_l < q_i (the Q primes decrease): 0 < z_i < 2 q_i.  The
  // narrow limbs add the term z_i (-B) to their sums and take one Shoup
  // product; q0 (summed by moddown_wide) multiplies both parts.
  const tw2* iql = reinterpret_cast<const tw2*>(P.inv_ql_mod_qi) + (size_t)level * k;
  const tw2* fin = reinterpret_cast<const tw2*>(P.md_fin) + (size_t)level * k;
  auto zval = [&](int i) { return ld_prime_p(P.primes, i) + (add ? add[(size_t)i * N + cid] : 0) - top; };
  limbs(
      level,
      [&](int i, uint64_t acc) {
        const uint64_t p = ld_prime_p(P.primes, i);
        const tw2 f = ld_tw2(fin, i), sq = ld_tw2(iql, i);
        out[(size_t)i * N + cid] = add_mod(mul_shoup(lift(i, acc), f.w, f.q, p), mul_shoup(zval(i), sq.w, sq.q, p), p);
      },
      [&](int i, uint64_t* z) {
        *z = zval(i);
        return true;
      },
      [&](int i, uint64_t acc) {
        const tw2 f = ld_tw2(fin, i);
        out[(size_t)i * N + cid] = mul_shoup(lift(i, acc), f.w, f.q, ld_prime_p(P.primes, i));
      });
}

// Eight waves per SIMD, as k_moddown_d: 2.5 % more plain HomMuls per second.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_moddown(DevParams P, ModDownArgs A) {
  extern __shared__ uint64_t aux[];  // [K][64], then [4][64] partial sums of the top limb and of q0
  const int y = blockIdx.y;
  moddown_body(P, A.inQ[y], A.inB[y], A.add[y], A.out[y], A.level, A.rescale, 0, blockIdx.x, aux);
}

void launch_moddown(const DevParams& P, const ModDownArgs& a, hipStream_t st) {
  const size_t lds = (size_t)(P.K + 2 * BC_GROUPS) * BC_COEF * sizeof(uint64_t);
  k_moddown<<<dim3(P.N / BC_COEF, a.npolys), 256, lds, st>>>(P, a);
}

// Eight waves per SIMD (58 registers): the batched grid is many times the
// machine, and more waves hide the product loop's table loads; 4 % faster
// than seven waves (70 registers) and than five (81, the default).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void k_moddown_d(DevParams P, const ModDownDesc* dd,
                                                                                           int bscaled) {
  extern __shared__ uint64_t aux[];
  const int per = P.N / BC_COEF;
  const ModDownDesc m = dd[blockIdx.x / per];
  // generic (flat) accesses here: the global form ran 1 % slower (more
  // scalar registers held for the bases, which this kernel already spills)
  moddown_body(P, m.inQ, m.inB, m.add, m.out, m.level, m.rescale, bscaled, blockIdx.x % per, aux);
}

void launch_moddown_desc(const DevParams& P, const ModDownDesc* d, int ndesc, bool bscaled, hipStream_t st) {
  const size_t lds = (size_t)(P.K + 2 * BC_GROUPS) * BC_COEF * sizeof(uint64_t);
  k_moddown_d<<<dim3((unsigned)(ndesc * (P.N / BC_COEF))), 256, lds, st>>>(P, d, bscaled ? 1 : 0);
}

// -------------------------------------------------------------- checksum ----
// acc += sum_i x[i] (2 (i + off) + 1) mod 2^64 over n words: a position-
// weighted sum that changes when any word changes or words move.  One atomic
// add per wave.
__global__ __launch_bounds__(256) void k_checksum(const uint64_t* x, uint64_t n, uint64_t off, uint64_t* acc) {
  uint64_t s = 0;
  const uint64_t stride = (uint64_t)gridDim.x * 256;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) s += x[i] * (2 * (i + off) + 1);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd((unsigned long long*)acc, (unsigned long long)s);
}

void launch_checksum(const uint64_t* x, size_t n, uint64_t off, uint64_t* acc, hipStream_t st) {
  if (!n) return;
  const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
  k_checksum<<<dim3((unsigned)blocks), 256, 0, st>>>(x, (uint64_t)n, off, acc);
}

// --------------------------------------------------------------- rescale ----
__global__ __launch_bounds__(256) void k_rescale(DevParams P, uint64_t* c0, uint64_t* c1, int level) {
  const size_t N = (size_t)P.N;
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  const int i = (int)(((size_t)blockIdx.x * 256) >> P.logN);
  const size_t cid = e & (N - 1);
  uint64_t* c = blockIdx.y ? c1 : c0;
  const uint64_t p = P.primes[i].p;
  const tw2 s = reinterpret_cast<const tw2*>(P.inv_ql_mod_qi)[(size_t)level * P.k + i];
  const uint64_t top = c[(size_t)level * N + cid];
  c[e] = mul_shoup(sub_mod(c[e], top, p), s.w, s.q, p);
}

void launch_rescale(const DevParams& P, uint64_t* c0, uint64_t* c1, int level, hipStream_t st) {
  if (level <= 0) return;
  const size_t total = (size_t)level * P.N;
  k_rescale<<<dim3((unsigned)(total / 256), 2), 256, 0, st>>>(P, c0, c1, level);
}

// ------------------------------------------------------------------- rho ----
// Q limbs times (B mod q_i), special limbs zero (aoa-ntt/src/cuda/manager.cu:149-160,
// where d_RNSBqi is zero past the Q limbs).
__global__ __launch_bounds__(256) void k_rho(DevParams P, uint64_t* out, const uint64_t* in) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  const int li = (int)(((size_t)blockIdx.x * 256) >> P.logN);
  if (li >= P.k) {
    out[e] = 0;
    return;
  }
  const tw2 b = reinterpret_cast<const tw2*>(P.b_mod_q)[li];
  out[e] = mul_shoup(in[e], b.w, b.q, P.primes[li].p);
}

void launch_rho(const DevParams& P, uint64_t* out, const uint64_t* in, hipStream_t st) {
  const size_t total = (size_t)(P.k + P.K) * P.N;
  k_rho<<<dim3((unsigned)(total / 256)), 256, 0, st>>>(P, out, in);
}

// ------------------------------------------------------- dot product, dr2 ----
// c = sum_i a[i] b[i] over every residue of a QB polynomial: the products are
// summed in 128 bits and reduced once with the reference's Barrett reduction
// (dot_prod, aoa-ntt/src/cuda/manager.cu:53-67, with add128 and mod of
// common.cu:195-213, :257-291), so a sum outside Barrett's exact range keeps
// the reference's result.  a and b are device arrays of k device pointers.
__global__ __launch_bounds__(256) void k_dot(DevParams P, uint64_t* c, const uint64_t* const* a,
                                             const uint64_t* const* b, int k) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  const DevPrime d = P.primes[(int)(((size_t)blockIdx.x * 256) >> P.logN)];
  uint64_t hi = 0, lo = 0;
  for (int i = 0; i < k; ++i) {
    const uint64_t x = glob(a[i])[e], y = glob(b[i])[e];
    const uint64_t l = x * y;
    lo += l;
    hi += mulhi64(x, y) + (lo < l ? 1u : 0u);
  }
  c[e] = ref_barrett(hi, lo, d.p, d.mu, (int)d.nb);
}

void launch_dot(const DevParams& P, uint64_t* c, const uint64_t* const* a, const uint64_t* const* b, int k,
                hipStream_t st) {
  const size_t total = (size_t)(P.k + P.K) * P.N;
  k_dot<<<dim3((unsigned)(total / 256)), 256, 0, st>>>(P, c, a, b, k);
}

// The reference's dr2 kernel (aoa-ntt/src/cuda/manager.cu:370-395) over every
// residue of a QB polynomial, with its mulmod (Barrett of the full product):
// c21 = a01 a11, c22 = (a01 a12)(a02 a11) -- a product where the tensor's
// middle term would be a sum; kept as the reference computes it --,
// c23 = a02 a12.
__global__ __launch_bounds__(256) void k_dr2(DevParams P, uint64_t* c21, uint64_t* c22, uint64_t* c23,
                                             const uint64_t* a01, const uint64_t* a02, const uint64_t* a11,
                                             const uint64_t* a12) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  const DevPrime d = P.primes[(int)(((size_t)blockIdx.x * 256) >> P.logN)];
  auto mm = [&](uint64_t x, uint64_t y) { return ref_barrett(mulhi64(x, y), x * y, d.p, d.mu, (int)d.nb); };
  const uint64_t x01 = a01[e], x02 = a02[e], x11 = a11[e], x12 = a12[e];
  c21[e] = mm(x01, x11);
  c22[e] = mm(mm(x01, x12), mm(x02, x11));
  c23[e] = mm(x02, x12);
}

void launch_dr2(const DevParams& P, uint64_t* c21, uint64_t* c22, uint64_t* c23, const uint64_t* a01,
                const uint64_t* a02, const uint64_t* a11, const uint64_t* a12, hipStream_t st) {
  const size_t total = (size_t)(P.k + P.K) * P.N;
  k_dr2<<<dim3((unsigned)(total / 256)), 256, 0, st>>>(P, c21, c22, c23, a01, a02, a11, a12);
}

// ---------------------------------------------------------- automorphism ----
// Forward map (encoder.cu:300-322): X^c -> X^(c g mod 2N), negated on wrap.
// Gather form: out[j] = in[t] if t = j g^-1 mod 2N < N, else -in[t - N].
// The gather reads one limb at a stride of g^-1: every 8-byte read falls in
// another cache line, so the source limb (N 8 bytes, 512 KiB at N=2^16) has
// to stay in L2 while its workgroups run.  Workgroups take their block of the
// output through xcd_chunk_id, so the blocks of one limb run on one XCD (one
// L2) one after another instead of on all eight.
__global__ __launch_bounds__(256) void k_automorphism(DevParams P, uint64_t* out, const uint64_t* in,
                                                      uint64_t ginv) {
  const size_t N = (size_t)P.N;
  const size_t blk = (size_t)xcd_chunk_id(blockIdx.x, gridDim.x);
  const size_t e = blk * 256 + threadIdx.x;
  const int li = (int)((blk * 256) >> P.logN);
  const uint64_t j = e & (N - 1);
  const uint64_t p = P.primes[li].p;
  const uint64_t t = (j * ginv) & (2 * N - 1);
  const uint64_t* src = in + (size_t)li * N;
  uint64_t v = t < N ? src[t] : neg_mod(src[t - N], p);
  out[e] = v;
}

uint64_t galois_inverse(uint64_t galois, int N) {
  // g^-1 mod 2N: the group (Z/2NZ)* has exponent N/2, so g^-1 = g^(N/2 - 1).
  const uint64_t M = 2 * (uint64_t)N;
  uint64_t r = 1, s = galois % M, e = (uint64_t)N / 2 - 1;
  while (e) {
    if (e & 1) r = (r * s) % M;
    s = (s * s) % M;
    e >>= 1;
  }
  return r;
}

void launch_automorphism(const DevParams& P, uint64_t* out, const uint64_t* in, int nlimbs,
                         uint64_t galois, hipStream_t st) {
  const size_t total = (size_t)nlimbs * P.N;
  k_automorphism<<<dim3((unsigned)(total / 256)), 256, 0, st>>>(P, out, in, galois_inverse(galois, P.N));
}

// Descriptor form, one entry per limb: out = sigma(in) (+ addsrc when set).
// Blocks mapped to XCDs as in k_automorphism.
__global__ __launch_bounds__(256) void k_automorphism_d(DevParams P, const AutoDesc* dd) {
  const int per = P.N / 256;
  const int blk = xcd_chunk_id(blockIdx.x, gridDim.x);
  const AutoDesc a = dd[blk / per];
  const uint64_t N = (uint64_t)P.N;
  const uint64_t j = (uint64_t)(blk % per) * 256 + threadIdx.x;
  const uint64_t p = P.primes[a.prime].p;
  const uint64_t t = (j * a.ginv) & (2 * N - 1);
  const uint64_t* in = glob(a.in);
  const uint64_t* addsrc = glob(a.addsrc);
  uint64_t v = t < N ? in[t] : neg_mod(in[t - N], p);
  if (addsrc) v = add_mod(v, addsrc[j], p);
  glob(a.out)[j] = v;
}

void launch_automorphism_desc(const DevParams& P, const AutoDesc* d, int ndesc, hipStream_t st) {
  k_automorphism_d<<<dim3((unsigned)(ndesc * (P.N / 256))), 256, 0, st>>>(P, d);
}

}  // namespace aoa
