// Parameter set of the CKKS evaluation core: RNS primes, NTT twiddles and the
// basis-conversion constants, computed on the host and mirrored in HBM.
//
// Mirrors CUDAManager::init(k, kl, N, prec) (aoa-ntt/src/cuda/manager.cu:762-800)
// but as an immutable object instead of a process-wide singleton, so several
// ring dimensions can live in one process.
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

#include "modarith.hpp"

namespace aoa {

struct aoa_prime_entry {
  uint64_t p;
  uint32_t g;  // smallest primitive root of p
};

// Per-prime constants used by every kernel (48 bytes, read through the scalar cache).
struct DevPrime {
  uint64_t p;
  uint64_t mu;       // floor(4^nb / p): Barrett constant (same value as the reference's MU)
  uint64_t ninv;     // N^-1 mod p
  uint64_t ninv_q;   // its Shoup quotient
  uint64_t eps;      // 2^nb - p (mod 2^64)
  uint64_t w1ninv;   // psi^-brev(1) * N^-1 mod p: last inverse stage with N^-1 folded in
  uint64_t w1ninv_q; // its Shoup quotient
  uint64_t r64;      // 2^64 mod p (reduction of 128-bit sums)
  uint64_t r64_q;    // its Shoup quotient
  // ModDown cross-base products a*c (a < max special prime, c < p): the
  // reference's Barrett result equals the exact remainder r unless
  // quirk_lo <= r < quirk_thr (see k_moddown); quirk_lo = eps and
  // quirk_thr = delta_max * eps, or [0, 2^64) when that bound is not usable.
  uint64_t quirk_lo;
  uint64_t quirk_thr;
  uint64_t mu2;      // mu * 2^(63 - nb): Barrett quotient as one mulhi (mul_barrett62)
  uint32_t nb;       // bit length of p
  // Special primes only: 1 when some Q prime after q0 exceeds this prime, so
  // ModUp must reduce those residues before multiplying (q0 is checked
  // separately).
  uint32_t modup_reduce;
};

// Pointers into one device allocation holding every table of a parameter set.
// Passed to kernels by value.
// Sampler constants: the reference's Gaussian standard deviation and
// Hamming weight (aoa-ntt/include/newckks/cuda/sampler.h:71-74), and the
// half-width of the Gaussian table (P(|e| > 31) < 2^-64 at sigma 3.2).
constexpr double kGaussSigma = 3.2;
constexpr int kGaussHalf = 32;
constexpr int kHammingWeight = 64;

struct DevParams {
  int N, logN, k, K;
  int ntt_n1;  // NTT split N = N1 x (N/N1); 0 = built-in default (autotuner sets it)
  // 0: ModDown reproduces the reference's Barrett reduction of its
  // cross-base products (bit-exact residues); 1: exact reduction.
  int moddown_exact;
  // 1: descriptor-driven NTT passes deal their work items out to the 8 XCDs
  // in contiguous chunks (work_id); 0: workgroup id order
  int xcd_chunks;
  const DevPrime* primes;        // [k+K]: Q primes then B primes
  const uint64_t* tw_fwd;        // [k+K][N][2]: psi^brev(i) and Shoup quotient
  const uint64_t* tw_inv;        // [k+K][N][2]: psi^-brev(i) and Shoup quotient
  // [log2(S) - 6][0: forward, 1: inverse]: for the row pass of the split
  // N = (N/S) x S, the twiddle table with the entries of the stages done on
  // 16 consecutive values per thread reordered (row_twiddle_order), so that
  // the S/16 threads of a row load adjacent entries; null where N/S is not a
  // valid split
  const uint64_t* tw_row[3][2];
  const uint64_t* inv_qhat;      // [k][k][2]: (l,i) -> (Q_l/q_i)^-1 mod q_i, Shoup
  const uint64_t* qhat_mod_b;    // [k][K][k]: (l,j,i) -> (Q_l/q_i) mod b_j
  const uint64_t* inv_bhat;      // [K][2]: (B/b_j)^-1 mod b_j, Shoup
  // [K][4]: N^-1 (B/b_j)^-1 and psi^-brev(1) N^-1 (B/b_j)^-1 mod b_j, each
  // with its Shoup quotient: the last inverse stage of a special limb whose
  // only consumer is ModDown, which then takes the limb as it comes
  const uint64_t* bscale;
  const uint64_t* bhat_mod_q;    // [k][K][2]: (i,j) -> (B/b_j) mod q_i, Shoup
  const uint64_t* inv_b_mod_q;   // [k][2]: B^-1 mod q_i, Shoup
  const uint64_t* b_mod_q;       // [k][2]: B mod q_i, Shoup
  const uint64_t* inv_ql_mod_qi; // [k][k][2]: (l,i) -> q_l^-1 mod q_i, Shoup
  const double* enc_w;           // [2N][2]: exp(2 pi i t / 2N), the encoder's roots of unity
  const uint32_t* rot_group;     // [N/2]: 5^j mod 2N
  const uint64_t* gauss_cdt;     // [2 kGaussHalf]: round(2^64 P(e <= j - kGaussHalf)), e = round(3.2 z)
  const uint32_t* md_terms;      // [K][k][6]: ModDown product constants (MdTerm), Q primes below 2^56
  const uint32_t* md_thr;        // [k]: ModDown screen threshold of q_i
  // [k+K][2]: per prime, floor(2^(63+nb) / p) (reduce_sums_lazy; 0 for q0)
  // and the bound of the lazy sums, 5 p (q0: q0, whose sums are exact): Q
  // primes for ModDown, special primes for ModUp
  const uint64_t* md_red;
  const uint64_t* md_fin;        // [k][k][2]: (l,i) -> B^-1 q_l^-1 mod q_i, Shoup
  // [k][4]: 28-bit halves (low, high, their sum, 0) of (-B) mod q_i, Q primes
  // below 2^57: ModDown adds z (-B) to its lazy sums as one more term
  const uint32_t* md_negb;
  // [k][k+1][kp][4], kp = K rounded up to 16: (l, i, position of j) -> 28-bit
  // halves of (Q_l/q_i) mod b_j and their sum, row i = k: of
  // (2^56 Q_l/q_0) mod b_j; positions as in build_host_params
  const uint32_t* mu_terms;
};

// The device layout of the ModDown product constants: MdConst's first six
// words, ordered [j][i] so that the limbs i, i+4, i+8, i+12 of one wave's
// pass sit at fixed offsets from one scalar pointer per term j.
struct MdTerm {
  uint32_t cl, ch, cs, w0, w1, w2;
};

// ModDown constants of one (q_i, b_j) pair for c = (B/b_j) mod q_i < 2^56:
// c split at bit 28 (cl, ch) and their sum cs for the three-product
// (Karatsuba) exact accumulation, the words of floor(c 2^96 / q_i) for
// moddown_screen, q_i's screen threshold (the same in every j) and the high
// word of moddown_screen's bias (2).
struct MdConst {
  uint32_t cl, ch, cs, w0, w1, w2, thr, pad;
};

struct HostParams {
  int N = 0, logN = 0, k = 0, K = 0, prec = 0;
  std::vector<uint64_t> primes;  // k + K
  std::vector<uint32_t> roots;
  std::vector<DevPrime> dprimes;
  std::vector<uint64_t> tw_fwd, tw_inv, inv_qhat, qhat_mod_b, inv_bhat, bhat_mod_q,
      inv_b_mod_q, b_mod_q, inv_ql_mod_qi;
  std::vector<uint64_t> bscale;     // K * 4
  std::vector<double> enc_w;        // 2N complex roots, interleaved (re, im)
  std::vector<uint32_t> rot_group;  // N/2 entries
  std::vector<uint64_t> gauss_cdt;  // 2 kGaussHalf entries
  std::vector<MdConst> md_const;    // k * K entries, [i][j]
  std::vector<MdTerm> md_terms;     // K * k entries, [j][i]
  std::vector<uint32_t> md_thr;     // k entries
  std::vector<uint64_t> md_red;     // (k + K) * 2
  std::vector<uint64_t> md_fin;     // k * k * 2
  std::vector<uint32_t> md_negb;    // k * 4
  std::vector<uint32_t> mu_terms;   // k * (k + 1) * kp * 4
};

// Builds every table.  Returns an empty string on success, else the reason
// (unsupported precision, bucket too small, N not supported by a prime).
std::string build_host_params(int N, int k, int kl, int prec, HostParams* out);
// The k Q primes then the kl special primes of the reference's layout for
// `prec` (empty string, or the reason the layout does not exist).
std::string select_rns_primes(int k, int kl, int prec, std::vector<uint64_t>* primes);

// Modular inverse of a mod p (p prime, a != 0 mod p).
uint64_t inv_mod(uint64_t a, uint64_t p);

uint32_t bit_reverse(uint32_t x, int bits);

// The row pass of an N = N1 x S split (g0 = log2 N1, S = 2^LOGS, TS = S/16
// threads per row) does global stages g0 + r, 4 <= r < LOGS, with each thread
// holding 16 consecutive values of its row: thread sh of row b uses the
// entries 2^(g0+r) + b 2^r + sh c + m, m < c = 2^r / TS.  The reordered table
// keeps entry sh c + m of each such block at m TS + sh, so that for each m
// the TS threads of a row read TS adjacent entries.  Other stages keep their
// order.  `tw` is one direction's [R][N][2] table; empty output when N/S is
// not a two-pass split (64, 128, 256 each way).
void row_twiddle_order(const std::vector<uint64_t>& tw, int N, int R, int S, std::vector<uint64_t>* out);

// The 32-bit words (w0, w1, w2) of floor(c 2^96 / p), c < p < 2^63
// (moddown_screen's constant).
void moddown_screen_words(uint64_t c, uint64_t p, uint32_t w[3]);

// moddown_screen threshold for remainders below thr < p.
uint32_t moddown_screen_thr(uint64_t thr, uint64_t p);

}  // namespace aoa
