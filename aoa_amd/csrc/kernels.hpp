// Host-side launchers of the gfx950 kernels (all stream-ordered, no sync).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_common.hpp"
#include "params.hpp"

namespace aoa {

enum EltOp { ELT_ADD = 0, ELT_SUB = 1, ELT_MUL = 2, ELT_NEG = 3, ELT_MULADD = 5 };
// Sampler distributions (same values as the C-ABI's AOA_SAMPLE_*).
enum SampleKind { SAMPLE_UNIFORM = 0, SAMPLE_GAUSSIAN = 1, SAMPLE_ZO = 2, SAMPLE_HW = 3, SAMPLE_NARROW = 4 };
constexpr int kHwMaxAttempts = 4096;

int ntt_default_n1(int N);
bool ntt_valid_split(int N, int N1);
// Forward / inverse NTT of every limb in `m`.  N1 = 0 picks the default split.
// With key0/key1 set, the inverse multiplies limb li of m.src[0] by limb
// prime_of(li) of key y before transforming into m.base[y] (y = 0, 1); with
// add0/add1 also set (two-pass splits only; -1 otherwise), (B mod q) times
// limb li of add y joins the product of each Q limb (li < m.qlimbs).
int ntt_forward(const DevParams& P, LimbMap m, int N1, hipStream_t st);
int ntt_inverse(const DevParams& P, LimbMap m, int N1, const uint64_t* key0, const uint64_t* key1,
                hipStream_t st, const uint64_t* add0 = nullptr, const uint64_t* add1 = nullptr);

// out = a op b (out = a*b + c for MULADD) over nlimbs limbs mapped to primes
// as in LimbMap (qlimbs Q limbs from prime0, then special limbs).
void launch_encrypt_combine(const DevParams& P, uint64_t* c0, uint64_t* c1, const uint64_t* pkb,
                            const uint64_t* pka, const uint64_t* u, const uint64_t* e0, const uint64_t* e1,
                            const uint64_t* m, int nlimbs, hipStream_t st);
void launch_eltwise(const DevParams& P, int op, uint64_t* out, const uint64_t* a, const uint64_t* b,
                    const uint64_t* c, int nlimbs, int qlimbs, int prime0, hipStream_t st);

// d0 = a0*b0, d1 = a0*b1 + a1*b0, d2 = a1*b1 on L Q-limbs (NTT domain).
void launch_tensor(const DevParams& P, uint64_t* d0, uint64_t* d1, uint64_t* d2, const uint64_t* a0,
                   const uint64_t* a1, const uint64_t* b0, const uint64_t* b1, int L, hipStream_t st);

// ModUp: x holds Q limbs 0..level (coefficient domain, stride N); writes K
// special limbs to outB (stride N).
void launch_eltwise_scalar(const DevParams& P, int op, uint64_t* out, const uint64_t* a, uint64_t b, int nlimbs,
                           int prime0, hipStream_t st);
void launch_conjugate(const DevParams& P, uint64_t* out, const uint64_t* in, int nlimbs, hipStream_t st);

// CKKS encoder (FP64).  vals holds n complex values (re, im interleaved);
// the special FFT runs in place (inverse: encoding direction, with the
// division by n).  encode_pack writes k limbs of N residues from slots values;
// decode_unpack reads limb 0 (coefficient domain) into slots values.
// Sampler: limbs res0..res0+nres-1 of out (poly base pointer, limb r at
// out + r*N) from distribution `kind`, keyed by (seed, nonce).
void launch_sample(const DevParams& P, uint64_t* out, int res0, int nres, int kind, uint64_t seed, uint64_t nonce,
                   hipStream_t st);
// u (ZO) at nonce, e0 and e1 (Gaussian) at nonce + 1 and nonce + 2, limbs
// 0..nres-1, in one launch.
void launch_sample_encrypt(const DevParams& P, uint64_t* u, uint64_t* e0, uint64_t* e1, int nres, uint64_t seed,
                           uint64_t nonce, hipStream_t st);

void launch_special_fft(const DevParams& P, double* vals, int n, int inverse, hipStream_t st);
void launch_encode_pack(const DevParams& P, uint64_t* out, const double* vals, int slots, double scale, int k,
                        hipStream_t st);
void launch_decode_unpack(const DevParams& P, double* vals, const uint64_t* in, int slots, double scale,
                          hipStream_t st);
void launch_modup(const DevParams& P, const uint64_t* x, uint64_t* outB, int level, hipStream_t st);

// ModDown of npolys polys.  inQ[y] holds Q limbs 0..level, inB[y] the K
// special limbs (coefficient domain).  out[y] limb i = (inQ_i - conv_i(inB)) *
// B^-1 mod q_i (+ add[y] limb i when add[y] != null).  With rescale, the
// result is further rescaled by q_level and only limbs 0..level-1 are written.
struct ModDownArgs {
  const uint64_t* inQ[2];
  const uint64_t* inB[2];
  const uint64_t* add[2];
  uint64_t* out[2];
  int npolys;
  int level;
  int rescale;
};
void launch_moddown(const DevParams& P, const ModDownArgs& a, hipStream_t st);

// Rescale c0 and c1 in place (coefficient domain) from `level`.
void launch_rescale(const DevParams& P, uint64_t* c0, uint64_t* c1, int level, hipStream_t st);
// acc += sum_i x[i] (2 (i + off) + 1) mod 2^64 over n words (acc on the device).
void launch_checksum(const uint64_t* x, size_t n, uint64_t off, uint64_t* acc, hipStream_t st);

// Galois automorphism X -> X^g (g odd) in the coefficient domain, gather form.
void launch_automorphism(const DevParams& P, uint64_t* out, const uint64_t* in, int nlimbs,
                         uint64_t galois, hipStream_t st);
uint64_t galois_inverse(uint64_t galois, int N);

// rho of a QB polynomial: Q limbs times (B mod q_i), special limbs zero.
void launch_rho(const DevParams& P, uint64_t* out, const uint64_t* in, hipStream_t st);
void launch_dot(const DevParams& P, uint64_t* c, const uint64_t* const* a, const uint64_t* const* b, int k,
                hipStream_t st);
void launch_dr2(const DevParams& P, uint64_t* c21, uint64_t* c22, uint64_t* c23, const uint64_t* a01,
                const uint64_t* a02, const uint64_t* a11, const uint64_t* a12, hipStream_t st);

// Descriptor-driven (stage-fused) forms used by batches: one launch covers
// every descriptor of the table (device memory).
// nstrict: the first nstrict descriptors use primes of 2^61 and above and
// the rest do not (a prime-major list); -1 when not known.
int ntt_forward_desc(const DevParams& P, const NttDesc* d, int ndesc, int N1, hipStream_t st, int nstrict = -1);
// canon = false: the two-pass form's lazy limbs may end in [0, 2p) (for
// ModUp and ModDown, which take them so); q0 and the single-workgroup form
// stay canonical.
int ntt_inverse_desc(const DevParams& P, const NttDesc* d, int ndesc, int N1, hipStream_t st, int nstrict = -1,
                     bool canon = true);
// Key-switching inverse NTT: first pass from KeyDesc2 entries (one input limb,
// two keys, two outputs), second pass over the NttDesc list of outputs.
// canon = false as above, and in the two-pass form the special limbs (prime
// index k + j) come out multiplied by (B/b_j)^-1 mod b_j, in [0, 2 b_j), for
// a ModDown that takes them so (ModDownDesc::bscaled).
int ntt_inverse_key2_desc(const DevParams& P, const KeyDesc2* d2, int n2, const NttDesc* dcol, int ncol, int N1,
                          hipStream_t st, int n2strict = -1, int ncolstrict = -1, bool canon = true);
// Whether ntt_inverse_key2_desc(..., N1, ..., canon = false) scales the
// special limbs (the two-pass form).
bool ntt_key2_bscaled(const DevParams& P, int N1);
void launch_tensor_desc(const DevParams& P, const TensorDesc* d, int ndesc, hipStream_t st);
void launch_modup_desc(const DevParams& P, const ModUpDesc* d, int ndesc, int max_level, hipStream_t st);
// bscaled: every descriptor's special limbs come from ntt_inverse_key2_desc's
// scaled form (already multiplied by (B/b_j)^-1 mod b_j, in [0, 2 b_j)).
void launch_moddown_desc(const DevParams& P, const ModDownDesc* d, int ndesc, bool bscaled, hipStream_t st);
void launch_automorphism_desc(const DevParams& P, const AutoDesc* d, int ndesc, hipStream_t st);

}  // namespace aoa
