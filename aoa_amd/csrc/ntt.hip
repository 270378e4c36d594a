// Negacyclic NTT over RNS limbs for gfx950.
//
// Output order and values equal the reference's hierarchical NTT
// (aoa-ntt/src/cuda/htrans/ntt.cu:321-460): out[pos] = sum_n x[n] *
// psi^(n * (2*brev(pos) + 1)) mod p, i.e. the standard Cooley-Tukey
// negacyclic NTT with merged twist and bit-reversed output; the inverse is the
// exact inverse (Gentleman-Sande, N^-1 folded into the last stage).
//
// The N-point transform is split as N = N1 x N2.  The column pass runs the
// first log2(N1) stages on S = N1 elements with stride N2; the row pass runs
// the remaining log2(N2) stages on contiguous rows of S = N2.  Both passes are
// in place (or read a separate source in the first pass), so no ping-pong
// buffer is needed.  A 256-thread workgroup transforms a 4096-element tile:
// TS = S/16 threads per sub-transform, 16 values per thread in registers, four
// stages per register round and one LDS exchange between rounds.  Twiddles
// psi^brev(i) come from a per-prime table with Shoup quotients.
//
// Every pass exists in two launch forms: limbs described by kernel arguments
// (plain calls) and limbs listed in a device descriptor table (batches).
#include "device_common.hpp"
#include "kernels.hpp"

namespace aoa {

template <int S>
struct Geo {
  static constexpr int LOGS = (S == 64) ? 6 : (S == 128 ? 7 : 8);
  static constexpr int TS = S / 16;   // threads per sub-transform
  static constexpr int C = 256 / TS;  // sub-transforms per tile
  static constexpr int COL_LDS = S * (C + 1);
  static constexpr int ROW_LDS = C * (S + S / 16);
};

// Butterflies.  LZ (primes below 2^61): Harvey's lazy forms, values kept in
// [0, 4p) through the forward transform and in [0, 2p) through the inverse,
// one correction per butterfly instead of three.  Otherwise (the 63-bit
// prime) every value stays canonical.
template <bool LZ>
__device__ __forceinline__ void ct_bfly(uint64_t& a, uint64_t& b, const tw2 t, uint64_t p) {
  if (LZ) {
    const uint64_t p2 = 2 * p;
    uint64_t u = a;
    u = u >= p2 ? u - p2 : u;
    const uint64_t v = mul_shoup_lazy(b, t.w, t.q, p);
    a = u + v;
    b = u + p2 - v;
  } else {
    uint64_t v = mul_shoup(b, t.w, t.q, p);
    uint64_t u = a;
    uint64_t s = u + v;
    a = s >= p ? s - p : s;
    b = u >= v ? u - v : u + (p - v);
  }
}

template <bool LZ>
__device__ __forceinline__ void gs_bfly(uint64_t& a, uint64_t& b, const tw2 t, uint64_t p) {
  uint64_t u = a, v = b;
  if (LZ) {
    const uint64_t p2 = 2 * p;
    const uint64_t s = u + v;
    a = s >= p2 ? s - p2 : s;
    b = mul_shoup_lazy(u + p2 - v, t.w, t.q, p);
  } else {
    uint64_t s = u + v;
    a = s >= p ? s - p : s;
    b = mul_shoup(u + (p - v), t.w, t.q, p);
  }
}

// [0, 4p) -> [0, p)
__device__ __forceinline__ uint64_t canon4(uint64_t x, uint64_t p) {
  const uint64_t p2 = 2 * p;
  x = x >= p2 ? x - p2 : x;
  return x >= p ? x - p : x;
}

__device__ __forceinline__ int swz(int s) { return s + (s >> 4); }

// ------------------------------------------------------------ pass bodies --
// src/dst point at limb 0 of the limb; `pr` is its prime index; `tile` the
// tile within the limb.

// Column pass: global stages 0..log2(S)-1 on columns of an S x N2 matrix.
template <int S, bool LZ>
__device__ __forceinline__ void fwd_col_body(const DevParams& P, const uint64_t* src, uint64_t* dst, int pr,
                                             int tile, int N2, uint64_t* lds) {
  using G = Geo<S>;
  constexpr int C = G::C, TS = G::TS, LOGS = G::LOGS, LDW = C + 1;
  const int col0 = tile * C;
  const uint64_t p = P.primes[pr].p;
  const uint64_t* T = P.tw_fwd + (size_t)pr * P.N * 2;
  src += col0;
  dst += col0;
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int sl = tid / C;
  uint64_t x[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = src[(size_t)(sl + TS * k) * N2 + c];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int half = 8 >> r;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      const tw2 t = load_tw(T, (1 << r) + (k >> (4 - r)));
      ct_bfly<LZ>(x[k], x[k + half], t, p);
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) lds[(sl + TS * k) * LDW + c] = x[k];
  __syncthreads();
  const int sh = sl;
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = lds[(16 * sh + k) * LDW + c];
#pragma unroll
  for (int r = 4; r < LOGS; ++r) {
    const int half = 1 << (LOGS - 1 - r);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      const tw2 t = load_tw(T, (1 << r) + (sh << (r - (LOGS - 4))) + (k >> (LOGS - r)));
      ct_bfly<LZ>(x[k], x[k + half], t, p);
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) dst[(size_t)(16 * sh + k) * N2 + c] = x[k];
}

// Row pass: global stages g0..g0+log2(S)-1 on contiguous rows of length S.
template <int S, bool LZ>
__device__ __forceinline__ void fwd_row_body(const DevParams& P, const uint64_t* src, uint64_t* dst, int pr,
                                             int tile, int g0, uint64_t* lds) {
  using G = Geo<S>;
  constexpr int C = G::C, TS = G::TS, LOGS = G::LOGS, PS = S + S / 16;
  const int row0 = tile * C;
  const uint64_t p = P.primes[pr].p;
  const uint64_t* T = P.tw_fwd + (size_t)pr * P.N * 2;
  // the second round's stages: entries in row order (row_twiddle_order)
  const uint64_t* TR = P.tw_row[LOGS - 6][0] + (size_t)pr * P.N * 2;
  const int tid = threadIdx.x;
  const int rl = tid / TS;
  const int sl = tid % TS;
  const int row = row0 + rl;
  src += (size_t)row * S;
  dst += (size_t)row * S;
  uint64_t* L = lds + rl * PS;
  uint64_t x[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = src[sl + TS * k];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int half = 8 >> r;
    const size_t base = ((size_t)1 << (g0 + r)) + ((size_t)row << r);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      const tw2 t = load_tw(T, base + (k >> (4 - r)));
      ct_bfly<LZ>(x[k], x[k + half], t, p);
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) L[swz(sl + TS * k)] = x[k];
  const int sh = sl;
  // The next round's twiddle is requested before the barrier, so its latency
  // overlaps the exchange.  In this round thread sh's m-th entry of a stage
  // sits at m TS + sh of the row's block.
  const tw2 t4 = load_tw(TR, ((size_t)1 << (g0 + 4)) + ((size_t)row << 4) + sh);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = L[swz(16 * sh + k)];
  {
    constexpr int half = 1 << (LOGS - 5);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      if (LOGS == 8) {
        ct_bfly<LZ>(x[k], x[k + half], t4, p);
      } else {
        const size_t base = ((size_t)1 << (g0 + 4)) + ((size_t)row << 4) + sh;
        ct_bfly<LZ>(x[k], x[k + half], load_tw(TR, base + (size_t)(k >> (LOGS - 4)) * TS), p);
      }
    }
  }
#pragma unroll
  for (int r = 5; r < LOGS; ++r) {
    const int half = 1 << (LOGS - 1 - r);
    const size_t base = ((size_t)1 << (g0 + r)) + ((size_t)row << r) + sh;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      const tw2 t = load_tw(TR, base + (size_t)(k >> (LOGS - r)) * TS);
      ct_bfly<LZ>(x[k], x[k + half], t, p);
    }
  }
  if (LZ) {
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = canon4(x[k], p);
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) L[swz(16 * sh + k)] = x[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) dst[sl + TS * k] = L[swz(sl + TS * k)];
}

// Row pass of the inverse: global stages g0+log2(S)-1 .. g0.  With a key, the
// input limb is first multiplied by the key limb (key switching:
// INTT(x * key_b), INTT(x * key_a) from one input).
// The inverse row pass after the input limb is in registers (x[k] = element
// sl + TS*k of the thread's row): stages g0+log2(S)-1 .. g0, result to dst.
// T is the prime's inverse table, TR its row-ordered copy (row_twiddle_order)
// for the first round, in which thread sh's m-th entry of a stage sits at
// m TS + sh of the row's block.
template <int S, bool LZ>
__device__ __forceinline__ void inv_row_core(const DevParams& P, uint64_t (&x)[16], uint64_t* dst, uint64_t p,
                                             const uint64_t* T, const uint64_t* TR, int row, int sl, int g0,
                                             uint64_t* L) {
  using G = Geo<S>;
  constexpr int TS = G::TS, LOGS = G::LOGS;
#pragma unroll
  for (int k = 0; k < 16; ++k) L[swz(sl + TS * k)] = x[k];
  const int sh = sl;
  // The first round's eight twiddles are requested before the barrier, so
  // their latency overlaps the exchange.
  tw2 t0[8];
  {
    const size_t base = ((size_t)1 << (g0 + LOGS - 1)) + ((size_t)row << (LOGS - 1)) + sh;
#pragma unroll
    for (int q = 0; q < 8; ++q) t0[q] = load_tw(TR, base + (size_t)q * TS);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = L[swz(16 * sh + k)];
#pragma unroll
  for (int k = 0; k < 16; k += 2) gs_bfly<LZ>(x[k], x[k + 1], t0[k >> 1], p);
#pragma unroll
  for (int r = LOGS - 2; r >= 4; --r) {
    const int half = 1 << (LOGS - 1 - r);
    const size_t base = ((size_t)1 << (g0 + r)) + ((size_t)row << r) + sh;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      const tw2 t = load_tw(TR, base + (size_t)(k >> (LOGS - r)) * TS);
      gs_bfly<LZ>(x[k], x[k + half], t, p);
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) L[swz(16 * sh + k)] = x[k];
  // likewise the eight twiddles of the round after the second exchange
  {
    const size_t base = ((size_t)1 << (g0 + 3)) + ((size_t)row << 3);
#pragma unroll
    for (int q = 0; q < 8; ++q) t0[q] = load_tw(T, base + q);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = L[swz(sl + TS * k)];
#pragma unroll
  for (int k = 0; k < 16; k += 2) gs_bfly<LZ>(x[k], x[k + 1], t0[k >> 1], p);
#pragma unroll
  for (int r = 2; r >= 0; --r) {
    const int half = 8 >> r;
    const size_t base = ((size_t)1 << (g0 + r)) + ((size_t)row << r);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      const tw2 t = load_tw(T, base + (k >> (4 - r)));
      gs_bfly<LZ>(x[k], x[k + half], t, p);
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) dst[sl + TS * k] = x[k];
}

// x * key mod p for a key-switching product (LZ: result in [0, 2p), which the
// lazy inverse butterflies accept).
template <bool LZ>
__device__ __forceinline__ uint64_t key_mul(uint64_t x, uint64_t k, const DevPrime& dp) {
  return LZ ? mul_barrett62_lazy(x, k, dp.p, dp.mu2, (int)dp.nb) : mul_barrett(x, k, dp.p, dp.mu, (int)dp.nb);
}

// Same, canonical: the operand of a further product.
template <bool LZ>
__device__ __forceinline__ uint64_t key_mul_canon(uint64_t x, uint64_t k, const DevPrime& dp) {
  return LZ ? mul_barrett62(x, k, dp.p, dp.mu2, (int)dp.nb) : mul_barrett(x, k, dp.p, dp.mu, (int)dp.nb);
}

// x (products in [0, 2p) when LZ, else canonical) plus (B mod q) times the
// NTT-domain addend of a Q limb pr, left in the same range.
template <bool LZ>
__device__ __forceinline__ void add_bscaled(uint64_t (&x)[16], const uint64_t* ap, int stride, const DevParams& P,
                                            int pr, uint64_t p) {
  const tw2 bq = reinterpret_cast<const tw2*>(P.b_mod_q)[pr];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const uint64_t t = x[k] + mul_shoup(ap[stride * k], bq.w, bq.q, p);
    // lazy: [0, 3p) -> [0, 2p); strict: canonical sum
    x[k] = LZ ? (t >= 2 * p ? t - 2 * p : t) : (t >= p ? t - p : t);
  }
}

// Row pass of the inverse: global stages g0+log2(S)-1 .. g0.  With a key, the
// input limb is first multiplied by the key limb, and then (if add is set,
// for Q limbs) (B mod q) times the add limb joins the product.
template <int S, bool LZ>
__device__ __forceinline__ void inv_row_body(const DevParams& P, const uint64_t* src, uint64_t* dst,
                                             const uint64_t* key, const uint64_t* add, int pr, int tile, int g0,
                                             uint64_t* lds) {
  using G = Geo<S>;
  constexpr int C = G::C, TS = G::TS, PS = S + S / 16;
  const DevPrime dp = P.primes[pr];
  const uint64_t* T = P.tw_inv + (size_t)pr * P.N * 2;
  const uint64_t* TR = P.tw_row[G::LOGS - 6][1] + (size_t)pr * P.N * 2;
  const int rl = threadIdx.x / TS, sl = threadIdx.x % TS;
  const int row = tile * C + rl;
  src += (size_t)row * S;
  uint64_t x[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = src[sl + TS * k];
  if (key) {
    const uint64_t* kp = key + (size_t)row * S;
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = key_mul<LZ>(x[k], kp[sl + TS * k], dp);
    if (add) add_bscaled<LZ>(x, add + (size_t)row * S + sl, TS, P, pr, dp.p);
  }
  inv_row_core<S, LZ>(P, x, dst + (size_t)row * S, dp.p, T, TR, row, sl, g0, lds + rl * PS);
}

// Key switching's two products from one input limb: dst0 = first inverse
// pass of (src * key0 + (B mod q) add0), dst1 = same with key1 and add1 (the
// adds may be null).  The workgroup reads its src tile twice; the second read
// is served by the L2 cache, so HBM sees the input once while the registers
// hold one tile at a time.  With HomMul's operands (a0 set: Q limbs of the
// batched form) the input is d2 = a1 b1 and the adds are d0 = a0 b0 and
// d1 = a0 b1 + a1 b0 (congruent to the reference's (a0+a1)(b0+b1) - d0 - d2),
// formed here from the four operand tiles, which the two key halves' workgroups
// of a tile read from HBM once between them (the second from L2).
template <int S, bool LZ>
__device__ __forceinline__ void inv_row_key2_body(const DevParams& P, const uint64_t* src, uint64_t* dst0,
                                                  uint64_t* dst1, const uint64_t* key0, const uint64_t* key1,
                                                  const uint64_t* add0, const uint64_t* add1, const uint64_t* a0,
                                                  const uint64_t* a1, const uint64_t* b0, const uint64_t* b1, int pr,
                                                  int tile, int g0, uint64_t* lds, int ysel) {
  using G = Geo<S>;
  constexpr int C = G::C, TS = G::TS, PS = S + S / 16;
  const DevPrime dp = P.primes[pr];
  const uint64_t p = dp.p;
  const uint64_t* T = P.tw_inv + (size_t)pr * P.N * 2;
  const uint64_t* TR = P.tw_row[G::LOGS - 6][1] + (size_t)pr * P.N * 2;
  const int rl = threadIdx.x / TS, sl = threadIdx.x % TS;
  const int row = tile * C + rl;
  const size_t off = (size_t)row * S;
#pragma unroll
  for (int y = 0; y < 2; ++y) {
    if (ysel >= 0 && y != ysel) continue;
    const uint64_t* kp = (y ? key1 : key0) + off;
    uint64_t x[16];
    if (a0) {
      const tw2 bq = reinterpret_cast<const tw2*>(P.b_mod_q)[pr];
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const size_t i = off + sl + TS * k;
        const uint64_t A0 = a0[i], A1 = a1[i], B0 = b0[i], B1 = b1[i];
        const uint64_t d2 = key_mul_canon<LZ>(A1, B1, dp);
        // lazy: d in [0, 4p), the sum in [0, 3p) -> [0, 2p); strict: below 2p, then canonical
        const uint64_t d = y ? key_mul<LZ>(A0, B1, dp) + key_mul<LZ>(A1, B0, dp) : key_mul<LZ>(A0, B0, dp);
        const uint64_t t = key_mul<LZ>(d2, kp[sl + TS * k], dp) + mul_shoup(d, bq.w, bq.q, p);
        x[k] = LZ ? (t >= 2 * p ? t - 2 * p : t) : (t >= p ? t - p : t);
      }
    } else {
      const uint64_t* ap = y ? add1 : add0;
#pragma unroll
      for (int k = 0; k < 16; ++k) x[k] = key_mul<LZ>(src[off + sl + TS * k], kp[sl + TS * k], dp);
      if (ap) add_bscaled<LZ>(x, ap + off + sl, TS, P, pr, p);  // set for Q limbs only
    }
    if (y && ysel < 0) __syncthreads();  // the row's LDS block is reused
    inv_row_core<S, LZ>(P, x, (y ? dst1 : dst0) + off, p, T, TR, row, sl, g0, lds + rl * PS);
  }
}

// Column pass of the inverse: global stages log2(S)-1 .. 0, N^-1 folded into
// stage 0.
// CANON = false (lazy primes only) leaves the outputs in [0, 2p) for a
// consumer that takes them so (ModUp, ModDown).  BS (with CANON = false):
// the special limbs (pr >= k) also take ModDown's factor (B/b_j)^-1 mod b_j
// in this stage (bscale), so that ModDown does not multiply them again.
template <int S, bool LZ, bool CANON = true, bool BS = false>
__device__ __forceinline__ void inv_col_body(const DevParams& P, const uint64_t* src, uint64_t* dst, int pr,
                                             int tile, int N2, uint64_t* lds) {
  using G = Geo<S>;
  constexpr int C = G::C, TS = G::TS, LOGS = G::LOGS, LDW = C + 1;
  const int col0 = tile * C;
  const DevPrime dp = P.primes[pr];
  const uint64_t p = dp.p;
  const uint64_t* T = P.tw_inv + (size_t)pr * P.N * 2;
  src += col0;
  dst += col0;
  const int tid = threadIdx.x;
  const int c = tid % C;
  const int sh = tid / C;
  uint64_t x[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = src[(size_t)(16 * sh + k) * N2 + c];
#pragma unroll
  for (int r = LOGS - 1; r >= 4; --r) {
    const int half = 1 << (LOGS - 1 - r);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      const tw2 t = load_tw(T, (1 << r) + (sh << (r - (LOGS - 4))) + (k >> (LOGS - r)));
      gs_bfly<LZ>(x[k], x[k + half], t, p);
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) lds[(16 * sh + k) * LDW + c] = x[k];
  __syncthreads();
  const int sl = sh;
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = lds[(sl + TS * k) * LDW + c];
#pragma unroll
  for (int r = 3; r >= 1; --r) {
    const int half = 8 >> r;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (k & half) continue;
      const tw2 t = load_tw(T, (1 << r) + (k >> (4 - r)));
      gs_bfly<LZ>(x[k], x[k + half], t, p);
    }
  }
  // stage 0 with N^-1: a' = (u+v) N^-1, b' = (u-v) psi^-brev(1) N^-1
  uint64_t n0 = dp.ninv, n0q = dp.ninv_q, n1 = dp.w1ninv, n1q = dp.w1ninv_q;
  if (BS && pr >= P.k) {
    const cptr<uint64_t> bs = as_const(P.bscale) + 4 * (pr - P.k);
    n0 = bs[0];
    n0q = bs[1];
    n1 = bs[2];
    n1q = bs[3];
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    uint64_t u = x[k], v = x[k + 8];
    if (LZ && !CANON) {
      x[k] = mul_shoup_lazy(u + v, n0, n0q, p);
      x[k + 8] = mul_shoup_lazy(u + 2 * p - v, n1, n1q, p);
    } else {
      x[k] = mul_shoup(u + v, dp.ninv, dp.ninv_q, p);
      x[k + 8] = mul_shoup(LZ ? u + 2 * p - v : u + (p - v), dp.w1ninv, dp.w1ninv_q, p);
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) dst[(size_t)(sl + TS * k) * N2 + c] = x[k];
}

// ------------------------------------------------- argument-driven kernels --

// Lazy butterflies need 4p < 2^63.
__device__ __forceinline__ bool lazy_prime(const DevParams& P, int pr) {
  return ld_prime(P.primes, pr).p < (1ull << 61);
}

template <int S>
__global__ __launch_bounds__(256) void k_ntt_fwd_col(DevParams P, LimbMap m, int N2) {
  __shared__ uint64_t lds[Geo<S>::COL_LDS];
  const int tiles = N2 / Geo<S>::C;
  const int li = blockIdx.x / tiles;
  const int y = blockIdx.y;
  const size_t off = (size_t)li * P.N;
  if (lazy_prime(P, limb_prime(m, P.k, li)))
    fwd_col_body<S, true>(P, m.src[y] + off, m.base[y] + off, limb_prime(m, P.k, li), blockIdx.x % tiles, N2, lds);
  else
    fwd_col_body<S, false>(P, m.src[y] + off, m.base[y] + off, limb_prime(m, P.k, li), blockIdx.x % tiles, N2, lds);
}

template <int S>
__global__ __launch_bounds__(256) void k_ntt_fwd_row(DevParams P, LimbMap m, int g0) {
  __shared__ uint64_t lds[Geo<S>::ROW_LDS];
  const int tiles = (P.N / S) / Geo<S>::C;
  const int li = blockIdx.x / tiles;
  const int y = blockIdx.y;
  const size_t off = (size_t)li * P.N;
  if (lazy_prime(P, limb_prime(m, P.k, li)))
    fwd_row_body<S, true>(P, m.src[y] + off, m.base[y] + off, limb_prime(m, P.k, li), blockIdx.x % tiles, g0, lds);
  else
    fwd_row_body<S, false>(P, m.src[y] + off, m.base[y] + off, limb_prime(m, P.k, li), blockIdx.x % tiles, g0, lds);
}

template <int S>
__global__ __launch_bounds__(256) void k_ntt_inv_row(DevParams P, LimbMap m, int g0, const uint64_t* key0,
                                                      const uint64_t* key1, const uint64_t* add0,
                                                      const uint64_t* add1) {
  __shared__ uint64_t lds[Geo<S>::ROW_LDS];
  const int tiles = (P.N / S) / Geo<S>::C;
  const int li = blockIdx.x / tiles;
  const int y = blockIdx.y;
  const size_t off = (size_t)li * P.N;
  const int pr = limb_prime(m, P.k, li);
  const uint64_t* key = key0 ? (y == 0 ? key0 : key1) + (size_t)pr * P.N : nullptr;
  const uint64_t* addy = y == 0 ? add0 : add1;
  const uint64_t* add = addy && li < m.qlimbs ? addy + off : nullptr;
  const uint64_t* src = (key0 ? m.src[0] : m.src[y]) + off;
  if (lazy_prime(P, pr))
    inv_row_body<S, true>(P, src, m.base[y] + off, key, add, pr, blockIdx.x % tiles, g0, lds);
  else
    inv_row_body<S, false>(P, src, m.base[y] + off, key, add, pr, blockIdx.x % tiles, g0, lds);
}

template <int S>
__global__ __launch_bounds__(256) void k_ntt_inv_col(DevParams P, LimbMap m, int N2) {
  __shared__ uint64_t lds[Geo<S>::COL_LDS];
  const int tiles = N2 / Geo<S>::C;
  const int li = blockIdx.x / tiles;
  const int y = blockIdx.y;
  const size_t off = (size_t)li * P.N;
  if (lazy_prime(P, limb_prime(m, P.k, li)))
    inv_col_body<S, true>(P, m.src[y] + off, m.base[y] + off, limb_prime(m, P.k, li), blockIdx.x % tiles, N2, lds);
  else
    inv_col_body<S, false>(P, m.src[y] + off, m.base[y] + off, limb_prime(m, P.k, li), blockIdx.x % tiles, N2, lds);
}

// ----------------------------------------------- descriptor-driven kernels --

// MODE: 0 decides lazy or strict butterflies per descriptor; 1 (every
// descriptor's prime is below 2^61) and 2 (none is) compile one form only,
// which keeps the register count of the lazy form (launchers split a
// prime-major descriptor list at the first lazy prime).
template <int MODE>
__device__ __forceinline__ bool use_lazy(const DevParams& P, int pr) {
  return MODE == 1 ? true : (MODE == 2 ? false : lazy_prime(P, pr));
}

// Work item of this workgroup: blockIdx.x, or with P.xcd_chunks its XCD
// chunk id (xcd_chunk_id), so that neighbouring items -- in a prime-major
// descriptor list, the tiles of one prime -- share an XCD and its L2.
__device__ __forceinline__ int work_id(const DevParams& P) {
  return P.xcd_chunks ? xcd_chunk_id(blockIdx.x, gridDim.x) : (int)blockIdx.x;
}

template <int S, int MODE>
__global__ __launch_bounds__(256) void k_ntt_fwd_col_d(DevParams P, const NttDesc* d, int N2) {
  __shared__ uint64_t lds[Geo<S>::COL_LDS];
  const int wid = work_id(P);
  const int tiles = N2 / Geo<S>::C;
  const NttDesc e = d[wid / tiles];
  if (use_lazy<MODE>(P, e.prime))
    fwd_col_body<S, true>(P, glob(e.src), glob(e.dst), e.prime, wid % tiles, N2, lds);
  else if (MODE != 1)
    fwd_col_body<S, false>(P, glob(e.src), glob(e.dst), e.prime, wid % tiles, N2, lds);
}

template <int S, int MODE>
__global__ __launch_bounds__(256) void k_ntt_fwd_row_d(DevParams P, const NttDesc* d, int g0) {
  __shared__ uint64_t lds[Geo<S>::ROW_LDS];
  const int wid = work_id(P);
  const int tiles = (P.N / S) / Geo<S>::C;
  const NttDesc e = d[wid / tiles];
  if (use_lazy<MODE>(P, e.prime))
    fwd_row_body<S, true>(P, glob(e.dst), glob(e.dst), e.prime, wid % tiles, g0, lds);
  else if (MODE != 1)
    fwd_row_body<S, false>(P, glob(e.dst), glob(e.dst), e.prime, wid % tiles, g0, lds);
}

template <int S, int MODE>
__global__ __launch_bounds__(256) void k_ntt_inv_row_d(DevParams P, const NttDesc* d, int g0) {
  __shared__ uint64_t lds[Geo<S>::ROW_LDS];
  const int wid = work_id(P);
  const int tiles = (P.N / S) / Geo<S>::C;
  const NttDesc e = d[wid / tiles];
  if (use_lazy<MODE>(P, e.prime))
    inv_row_body<S, true>(P, glob(e.src), glob(e.dst), glob(e.key), nullptr, e.prime, wid % tiles, g0, lds);
  else if (MODE != 1)
    inv_row_body<S, false>(P, glob(e.src), glob(e.dst), glob(e.key), nullptr, e.prime, wid % tiles, g0, lds);
}

// One workgroup per tile and key half: workgroups w and w + 8 take the two
// halves of one tile; as workgroups are dispatched round-robin over the 8
// XCDs, both run on one XCD and the second read of the input tile comes from
// its L2.  `ntiles` = descriptors x tiles per limb; the grid is rounded up to
// a multiple of 16.
template <int S, int MODE>
__global__ __launch_bounds__(256) void k_ntt_inv_row_key2_d(DevParams P, const KeyDesc2* d, int g0, int ntiles) {
  __shared__ uint64_t lds[Geo<S>::ROW_LDS];
  const int tiles = (P.N / S) / Geo<S>::C;
  const int w = blockIdx.x;
  const int y = (w >> 3) & 1, b = ((w >> 4) << 3) | (w & 7);
  if (b >= ntiles) return;
  const KeyDesc2& e = d[b / tiles];
  const uint64_t* src = glob(e.src);
  uint64_t *d0 = glob(e.dst[0]), *d1 = glob(e.dst[1]);
  const uint64_t *k0 = glob(e.key[0]), *k1 = glob(e.key[1]), *a0 = glob(e.add[0]), *a1 = glob(e.add[1]);
  const uint64_t *t0 = glob(e.ab[0]), *t1 = glob(e.ab[1]), *t2 = glob(e.ab[2]), *t3 = glob(e.ab[3]);
  const int pr = e.prime;
  if (use_lazy<MODE>(P, pr))
    inv_row_key2_body<S, true>(P, src, d0, d1, k0, k1, a0, a1, t0, t1, t2, t3, pr, b % tiles, g0, lds, y);
  else if (MODE != 1)
    inv_row_key2_body<S, false>(P, src, d0, d1, k0, k1, a0, a1, t0, t1, t2, t3, pr, b % tiles, g0, lds, y);
}

template <int S, int MODE, bool CANON, bool BS = false>
__device__ __forceinline__ void inv_col_d(const DevParams& P, const NttDesc* d, int N2, uint64_t* lds) {
  const int wid = work_id(P);
  const int tiles = N2 / Geo<S>::C;
  const NttDesc e = d[wid / tiles];
  if (use_lazy<MODE>(P, e.prime))
    inv_col_body<S, true, CANON, BS>(P, glob(e.dst), glob(e.dst), e.prime, wid % tiles, N2, lds);
  else if (MODE != 1)
    inv_col_body<S, false>(P, glob(e.dst), glob(e.dst), e.prime, wid % tiles, N2, lds);
}

template <int S, int MODE>
__global__ __launch_bounds__(256) void k_ntt_inv_col_d(DevParams P, const NttDesc* d, int N2) {
  __shared__ uint64_t lds[Geo<S>::COL_LDS];
  inv_col_d<S, MODE, true>(P, d, N2, lds);
}

// Same, the lazy limbs' outputs left in [0, 2p) (the q0 limbs stay canonical).
template <int S, int MODE>
__global__ __launch_bounds__(256) void k_ntt_inv_col_lazy_d(DevParams P, const NttDesc* d, int N2) {
  __shared__ uint64_t lds[Geo<S>::COL_LDS];
  inv_col_d<S, MODE, false>(P, d, N2, lds);
}

// Same, the special limbs also multiplied by (B/b_j)^-1 mod b_j for ModDown.
template <int S, int MODE>
__global__ __launch_bounds__(256) void k_ntt_inv_col_bscale_d(DevParams P, const NttDesc* d, int N2) {
  __shared__ uint64_t lds[Geo<S>::COL_LDS];
  inv_col_d<S, MODE, false, true>(P, d, N2, lds);
}

// ------------------------------------------- single-workgroup transform --
// N <= 4096: one workgroup holds a whole limb (V values per thread,
// 2^R = V) and runs every stage, exchanging through LDS between rounds of R
// stages, so each limb is read and written once.  In a round whose Rr active
// bits sit at positions [lo, lo + Rr) of the element index, value j of thread
// t is element sp_index(t, j): the active bits come from j, the rest from t
// (and, when Rr < R, from the high bits of j).  Loads and stores go through
// LDS in the order i = t + TH m, so every global access is coalesced.

template <int R>
__device__ __forceinline__ int sp_index(int t, int j, int lo, int Rr) {
  const int jj = j & ((1 << Rr) - 1), g = j >> Rr;
  const int o = (t << (R - Rr)) | g;
  const int low = o & ((1 << lo) - 1), high = o >> lo;
  return (high << (lo + Rr)) | (jj << lo) | low;
}

__device__ __forceinline__ int spz(int i) { return i + (i >> 5); }

// TH = N/4 threads (16 waves at N = 4096), four values each: a limb's
// transform is one workgroup, so the threads per limb set its latency, and
// four waves per SIMD hide the twiddle and LDS waits that one wave per SIMD
// (256 threads, 16 values each) left exposed.
template <int LOGN>
struct SGeo {
  static constexpr int N = 1 << LOGN, TH = N / 4, V = 4, R = 2, LDS = N + N / 32;
};

// Coalesced load of a limb (optionally times a key limb) into LDS.
template <int LOGN, bool LZ>
__device__ __forceinline__ void sp_load(const uint64_t* src, const uint64_t* key, const DevPrime& dp,
                                        uint64_t* lds) {
  constexpr int V = SGeo<LOGN>::V, TH = SGeo<LOGN>::TH;
  const int t = threadIdx.x;
#pragma unroll
  for (int m = 0; m < V; ++m) {
    const int i = t + TH * m;
    uint64_t v = src[i];
    if (key) v = key_mul<LZ>(v, key[i], dp);
    lds[spz(i)] = v;
  }
  __syncthreads();
}

template <int LOGN>
__device__ __forceinline__ void sp_store(uint64_t* dst, const uint64_t* lds) {
  constexpr int V = SGeo<LOGN>::V, TH = SGeo<LOGN>::TH;
  const int t = threadIdx.x;
#pragma unroll
  for (int m = 0; m < V; ++m) dst[t + TH * m] = lds[spz(t + TH * m)];
}

template <int LOGN, bool LZ>
__device__ __forceinline__ void fwd_single_body(const DevParams& P, const uint64_t* src, uint64_t* dst, int pr,
                                                uint64_t* lds) {
  using G = SGeo<LOGN>;
  constexpr int V = G::V, R = G::R;
  const DevPrime dp = P.primes[pr];
  const uint64_t p = dp.p;
  const uint64_t* T = P.tw_fwd + (size_t)pr * G::N * 2;
  const int t = threadIdx.x;
  sp_load<LOGN, LZ>(src, nullptr, dp, lds);
  uint64_t x[V];
#pragma unroll
  for (int s0 = 0; s0 < LOGN; s0 += R) {
    const int Rr = (LOGN - s0) < R ? (LOGN - s0) : R;
    const int lo = LOGN - s0 - Rr;
#pragma unroll
    for (int j = 0; j < V; ++j) x[j] = lds[spz(sp_index<R>(t, j, lo, Rr))];
#pragma unroll
    for (int r = 0; r < Rr; ++r) {
      const int s = s0 + r, half = 1 << (Rr - 1 - r);
#pragma unroll
      for (int j = 0; j < V; ++j) {
        if (j & half) continue;
        const int i = sp_index<R>(t, j, lo, Rr);
        const tw2 w = load_tw(T, ((size_t)1 << s) + (size_t)(i >> (LOGN - s)));
        ct_bfly<LZ>(x[j], x[j + half], w, p);
      }
    }
    if (LZ && s0 + Rr >= LOGN) {
#pragma unroll
      for (int j = 0; j < V; ++j) x[j] = canon4(x[j], p);
    }
    __syncthreads();  // every thread has read the previous layout
#pragma unroll
    for (int j = 0; j < V; ++j) lds[spz(sp_index<R>(t, j, lo, Rr))] = x[j];
    __syncthreads();
  }
  sp_store<LOGN>(dst, lds);
}

// Inverse: stages LOGN-1 .. 0 (Gentleman-Sande), N^-1 folded into stage 0;
// with a key the input limb is first multiplied by the key limb.
template <int LOGN, bool LZ>
__device__ __forceinline__ void inv_single_body(const DevParams& P, const uint64_t* src, uint64_t* dst,
                                                const uint64_t* key, int pr, uint64_t* lds) {
  using G = SGeo<LOGN>;
  constexpr int V = G::V, R = G::R;
  const DevPrime dp = P.primes[pr];
  const uint64_t p = dp.p;
  const uint64_t* T = P.tw_inv + (size_t)pr * G::N * 2;
  const int t = threadIdx.x;
  sp_load<LOGN, LZ>(src, key, dp, lds);
  uint64_t x[V];
#pragma unroll
  for (int done = 0; done < LOGN; done += R) {
    const int shi = LOGN - 1 - done;  // highest stage of this round
    const int Rr = (LOGN - done) < R ? (LOGN - done) : R;
    const int lo = LOGN - 1 - shi;
#pragma unroll
    for (int j = 0; j < V; ++j) x[j] = lds[spz(sp_index<R>(t, j, lo, Rr))];
#pragma unroll
    for (int r = 0; r < Rr; ++r) {
      const int s = shi - r, half = 1 << r;
#pragma unroll
      for (int j = 0; j < V; ++j) {
        if (j & half) continue;
        if (s == 0) {
          const uint64_t u = x[j], v = x[j + half];
          x[j] = mul_shoup(u + v, dp.ninv, dp.ninv_q, p);
          x[j + half] = mul_shoup(LZ ? u + 2 * p - v : u + (p - v), dp.w1ninv, dp.w1ninv_q, p);
        } else {
          const int i = sp_index<R>(t, j, lo, Rr);
          const tw2 w = load_tw(T, ((size_t)1 << s) + (size_t)(i >> (LOGN - s)));
          gs_bfly<LZ>(x[j], x[j + half], w, p);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < V; ++j) lds[spz(sp_index<R>(t, j, lo, Rr))] = x[j];
    __syncthreads();
  }
  sp_store<LOGN>(dst, lds);
}

template <int LOGN>
__global__ __launch_bounds__(SGeo<LOGN>::TH) void k_ntt_fwd_single(DevParams P, LimbMap m) {
  __shared__ uint64_t lds[SGeo<LOGN>::LDS];
  const int li = blockIdx.x, y = blockIdx.y, pr = limb_prime(m, P.k, li);
  const size_t off = (size_t)li * P.N;
  if (lazy_prime(P, pr))
    fwd_single_body<LOGN, true>(P, m.src[y] + off, m.base[y] + off, pr, lds);
  else
    fwd_single_body<LOGN, false>(P, m.src[y] + off, m.base[y] + off, pr, lds);
}

template <int LOGN>
__global__ __launch_bounds__(SGeo<LOGN>::TH) void k_ntt_inv_single(DevParams P, LimbMap m, const uint64_t* key0,
                                                         const uint64_t* key1) {
  __shared__ uint64_t lds[SGeo<LOGN>::LDS];
  const int li = blockIdx.x, y = blockIdx.y, pr = limb_prime(m, P.k, li);
  const size_t off = (size_t)li * P.N;
  const uint64_t* key = key0 ? (y == 0 ? key0 : key1) + (size_t)pr * P.N : nullptr;
  const uint64_t* src = (key0 ? m.src[0] : m.src[y]) + off;
  if (lazy_prime(P, pr))
    inv_single_body<LOGN, true>(P, src, m.base[y] + off, key, pr, lds);
  else
    inv_single_body<LOGN, false>(P, src, m.base[y] + off, key, pr, lds);
}

template <int LOGN>
__global__ __launch_bounds__(SGeo<LOGN>::TH) void k_ntt_fwd_single_d(DevParams P, const NttDesc* d) {
  __shared__ uint64_t lds[SGeo<LOGN>::LDS];
  const NttDesc& e = d[blockIdx.x];
  const uint64_t* src = glob(e.src);
  uint64_t* dst = glob(e.dst);
  const int pr = e.prime;
  if (lazy_prime(P, pr))
    fwd_single_body<LOGN, true>(P, src, dst, pr, lds);
  else
    fwd_single_body<LOGN, false>(P, src, dst, pr, lds);
}

template <int LOGN>
__global__ __launch_bounds__(SGeo<LOGN>::TH) void k_ntt_inv_single_d(DevParams P, const NttDesc* d) {
  __shared__ uint64_t lds[SGeo<LOGN>::LDS];
  const NttDesc& e = d[blockIdx.x];
  const uint64_t* src = glob(e.src);
  uint64_t* dst = glob(e.dst);
  const uint64_t* key = glob(e.key);
  const int pr = e.prime;
  if (lazy_prime(P, pr))
    inv_single_body<LOGN, true>(P, src, dst, key, pr, lds);
  else
    inv_single_body<LOGN, false>(P, src, dst, key, pr, lds);
}

// Both key products of one input limb, one workgroup per (limb, key half).
template <int LOGN>
__global__ __launch_bounds__(SGeo<LOGN>::TH) void k_ntt_inv_single_key2_d(DevParams P, const KeyDesc2* d) {
  __shared__ uint64_t lds[SGeo<LOGN>::LDS];
  const KeyDesc2& e = d[blockIdx.x >> 1];
  const int y = blockIdx.x & 1;
  const uint64_t* src = glob(e.src);
  uint64_t* dst = glob(y ? e.dst[1] : e.dst[0]);
  const uint64_t* key = glob(y ? e.key[1] : e.key[0]);
  const int pr = e.prime;
  if (lazy_prime(P, pr))
    inv_single_body<LOGN, true>(P, src, dst, key, pr, lds);
  else
    inv_single_body<LOGN, false>(P, src, dst, key, pr, lds);
}

// ---------------------------------------------------------------- launch ---

// N1 == N selects the single-workgroup transform (N <= 4096).
bool ntt_valid_split(int N, int N1) {
  if (N1 == N) return N == 2048 || N == 4096;
  if (N1 != 64 && N1 != 128 && N1 != 256) return false;
  if (N % N1) return false;
  int N2 = N / N1;
  return N2 == 64 || N2 == 128 || N2 == 256;
}

int ntt_default_n1(int N) {
  switch (N) {
    case 2048: return 2048;
    case 4096: return 4096;
    case 8192: return 128;
    case 16384: return 128;
    case 32768: return 128;
    case 65536: return 256;
    default: return 0;
  }
}

static int log2i(int n) {
  int g = 0;
  while ((1 << g) < n) ++g;
  return g;
}

#define AOA_DISPATCH_S(S_, CALL) \
  switch (S_) {                  \
    case 64: CALL(64); break;    \
    case 128: CALL(128); break;  \
    case 256: CALL(256); break;  \
    default: return -1;          \
  }

static int pick_n1(const DevParams& P, int N1) {
  if (N1 == 0) N1 = P.ntt_n1 ? P.ntt_n1 : ntt_default_n1(P.N);
  return ntt_valid_split(P.N, N1) ? N1 : 0;
}

#define AOA_DISPATCH_LOGN(N_, CALL) \
  switch (N_) {                      \
    case 2048: CALL(11); break;      \
    case 4096: CALL(12); break;      \
    default: return -1;              \
  }

int ntt_forward(const DevParams& P, LimbMap m, int N1, hipStream_t st) {
  N1 = pick_n1(P, N1);
  if (!N1) return -1;
  const int N2 = P.N / N1, g0 = log2i(N1);
  for (int y = 0; y < m.npolys; ++y)
    if (!m.src[y]) m.src[y] = m.base[y];
  if (N1 == P.N) {
    dim3 grid(m.nlimbs, m.npolys);
#define CALL1(L) k_ntt_fwd_single<L><<<grid, SGeo<L>::TH, 0, st>>>(P, m)
    AOA_DISPATCH_LOGN(P.N, CALL1)
#undef CALL1
    return 0;
  }
  {
    const int tiles = N2 / (4096 / N1);
    dim3 grid(tiles * m.nlimbs, m.npolys);
#define CALLC(S) k_ntt_fwd_col<S><<<grid, 256, 0, st>>>(P, m, N2)
    AOA_DISPATCH_S(N1, CALLC)
#undef CALLC
  }
  LimbMap m2 = m;
  for (int y = 0; y < m2.npolys; ++y) m2.src[y] = m2.base[y];
  {
    const int tiles = N1 / (4096 / N2);
    dim3 grid(tiles * m.nlimbs, m.npolys);
#define CALLR(S) k_ntt_fwd_row<S><<<grid, 256, 0, st>>>(P, m2, g0)
    AOA_DISPATCH_S(N2, CALLR)
#undef CALLR
  }
  return 0;
}

int ntt_inverse(const DevParams& P, LimbMap m, int N1, const uint64_t* key0, const uint64_t* key1,
                hipStream_t st, const uint64_t* add0, const uint64_t* add1) {
  N1 = pick_n1(P, N1);
  if (!N1) return -1;
  if ((add0 || add1) && (N1 == P.N || !key0)) return -1;  // adds: two-pass key products only
  const int N2 = P.N / N1, g0 = log2i(N1);
  for (int y = 0; y < m.npolys; ++y)
    if (!m.src[y]) m.src[y] = m.base[y];
  if (N1 == P.N) {
    dim3 grid(m.nlimbs, m.npolys);
#define CALL1(L) k_ntt_inv_single<L><<<grid, SGeo<L>::TH, 0, st>>>(P, m, key0, key1)
    AOA_DISPATCH_LOGN(P.N, CALL1)
#undef CALL1
    return 0;
  }
  {
    const int tiles = N1 / (4096 / N2);
    dim3 grid(tiles * m.nlimbs, m.npolys);
#define CALLR(S) k_ntt_inv_row<S><<<grid, 256, 0, st>>>(P, m, g0, key0, key1, add0, add1)
    AOA_DISPATCH_S(N2, CALLR)
#undef CALLR
  }
  LimbMap m2 = m;
  for (int y = 0; y < m2.npolys; ++y) m2.src[y] = m2.base[y];
  {
    const int tiles = N2 / (4096 / N1);
    dim3 grid(tiles * m.nlimbs, m.npolys);
#define CALLC(S) k_ntt_inv_col<S><<<grid, 256, 0, st>>>(P, m2, N2)
    AOA_DISPATCH_S(N1, CALLC)
#undef CALLC
  }
  return 0;
}

// Launches kernel K<S, MODE> over `n` descriptors of which the first
// `strict` use primes of 2^61 and above (strict < 0: not known, one launch
// deciding per descriptor).
#define AOA_SPLIT_LAUNCH(K, S, GRIDPER, DESC, N_, STRICT, ...)                                          \
  do {                                                                                                  \
    if ((STRICT) < 0) {                                                                                 \
      K<S, 0><<<dim3((GRIDPER) * (N_)), 256, 0, st>>>(P, DESC, __VA_ARGS__);                           \
    } else {                                                                                            \
      if ((STRICT) > 0) K<S, 2><<<dim3((GRIDPER) * (STRICT)), 256, 0, st>>>(P, DESC, __VA_ARGS__);     \
      if ((N_) > (STRICT))                                                                              \
        K<S, 1><<<dim3((GRIDPER) * ((N_) - (STRICT))), 256, 0, st>>>(P, (DESC) + (STRICT), __VA_ARGS__); \
    }                                                                                                   \
  } while (0)

int ntt_forward_desc(const DevParams& P, const NttDesc* d, int ndesc, int N1, hipStream_t st, int nstrict) {
  N1 = pick_n1(P, N1);
  if (!N1 || ndesc <= 0) return -1;
  const int N2 = P.N / N1, g0 = log2i(N1);
  if (N1 == P.N) {
#define CALL1(L) k_ntt_fwd_single_d<L><<<dim3(ndesc), SGeo<L>::TH, 0, st>>>(P, d)
    AOA_DISPATCH_LOGN(P.N, CALL1)
#undef CALL1
    return 0;
  }
  {
    const int tiles = N2 / (4096 / N1);
#define CALLC(S) AOA_SPLIT_LAUNCH(k_ntt_fwd_col_d, S, tiles, d, ndesc, nstrict, N2)
    AOA_DISPATCH_S(N1, CALLC)
#undef CALLC
  }
  {
    const int tiles = N1 / (4096 / N2);
#define CALLR(S) AOA_SPLIT_LAUNCH(k_ntt_fwd_row_d, S, tiles, d, ndesc, nstrict, g0)
    AOA_DISPATCH_S(N2, CALLR)
#undef CALLR
  }
  return 0;
}

int ntt_inverse_desc(const DevParams& P, const NttDesc* d, int ndesc, int N1, hipStream_t st, int nstrict, bool canon) {
  N1 = pick_n1(P, N1);
  if (!N1 || ndesc <= 0) return -1;
  const int N2 = P.N / N1, g0 = log2i(N1);
  if (N1 == P.N) {
#define CALL1(L) k_ntt_inv_single_d<L><<<dim3(ndesc), SGeo<L>::TH, 0, st>>>(P, d)
    AOA_DISPATCH_LOGN(P.N, CALL1)
#undef CALL1
    return 0;
  }
  {
    const int tiles = N1 / (4096 / N2);
#define CALLR(S) AOA_SPLIT_LAUNCH(k_ntt_inv_row_d, S, tiles, d, ndesc, nstrict, g0)
    AOA_DISPATCH_S(N2, CALLR)
#undef CALLR
  }
  {
    const int tiles = N2 / (4096 / N1);
#define CALLC(S)                                                              \
  if (canon) AOA_SPLIT_LAUNCH(k_ntt_inv_col_d, S, tiles, d, ndesc, nstrict, N2); \
  else AOA_SPLIT_LAUNCH(k_ntt_inv_col_lazy_d, S, tiles, d, ndesc, nstrict, N2)
    AOA_DISPATCH_S(N1, CALLC)
#undef CALLC
  }
  return 0;
}

bool ntt_key2_bscaled(const DevParams& P, int N1) {
  N1 = pick_n1(P, N1);
  return N1 != 0 && N1 != P.N;
}

int ntt_inverse_key2_desc(const DevParams& P, const KeyDesc2* d2, int n2, const NttDesc* dcol, int ncol, int N1,
                          hipStream_t st, int n2strict, int ncolstrict, bool canon) {
  N1 = pick_n1(P, N1);
  if (!N1 || n2 <= 0 || ncol <= 0) return -1;
  const int N2 = P.N / N1, g0 = log2i(N1);
  if (N1 == P.N) {
#define CALL1(L) k_ntt_inv_single_key2_d<L><<<dim3(2 * n2), SGeo<L>::TH, 0, st>>>(P, d2)
    AOA_DISPATCH_LOGN(P.N, CALL1)
#undef CALL1
    return 0;
  }
  {
    const int tiles = N1 / (4096 / N2);
    auto grid = [&](int n) { return dim3((unsigned)((2 * tiles * n + 15) / 16 * 16)); };
    const int nlazy = n2strict < 0 ? 0 : n2 - n2strict;
#define CALLR(S)                                                                                          \
  if (n2strict < 0) {                                                                                     \
    k_ntt_inv_row_key2_d<S, 0><<<grid(n2), 256, 0, st>>>(P, d2, g0, tiles * n2);                          \
  } else {                                                                                                \
    if (n2strict > 0) k_ntt_inv_row_key2_d<S, 2><<<grid(n2strict), 256, 0, st>>>(P, d2, g0, tiles * n2strict); \
    if (nlazy > 0)                                                                                        \
      k_ntt_inv_row_key2_d<S, 1><<<grid(nlazy), 256, 0, st>>>(P, d2 + n2strict, g0, tiles * nlazy);        \
  }
    AOA_DISPATCH_S(N2, CALLR)
#undef CALLR
  }
  {
    const int tiles = N2 / (4096 / N1);
#define CALLC(S)                                                                    \
  if (canon) AOA_SPLIT_LAUNCH(k_ntt_inv_col_d, S, tiles, dcol, ncol, ncolstrict, N2); \
  else AOA_SPLIT_LAUNCH(k_ntt_inv_col_bscale_d, S, tiles, dcol, ncol, ncolstrict, N2)
    AOA_DISPATCH_S(N1, CALLC)
#undef CALLC
  }
  return 0;
}

}  // namespace aoa
