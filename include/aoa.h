/*
 * aoa.h -- C-ABI of the MI355X (gfx950) CKKS evaluation core.
 *
 * Plain C: opaque handles, raw device pointers (uint64_t residues, limb-major:
 * limb r of a polynomial is at ptr + r*N), sizes and integer status codes.
 * Every call is ordered on the HIP stream passed as `stream` (a hipStream_t,
 * NULL = default stream) and never synchronises the device.  No call
 * allocates or frees a caller buffer; scratch memory is passed in as a
 * workspace of aoa_*_workspace_bytes() bytes.  Ciphertext and workspace
 * pointers must be 16-byte aligned (hipMalloc returns 256-byte alignment).
 *
 * Layouts follow the reference (aoa-ntt/include/newckks/defines.h:66-72,
 * aoa-ntt/src/arithmetic/poly_t.cpp:6-31): a Q polynomial has k limbs (primes
 * q_0..q_{k-1}); a QB polynomial (switching keys) has k+kl limbs, Q limbs
 * first.  At level l only limbs 0..l of a Q polynomial are meaningful.
 */
#ifndef AOA_H
#define AOA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aoa_params aoa_params;

/* Status codes. */
enum {
  AOA_OK = 0,
  AOA_ERR_ARG = 1,         /* bad argument (null pointer, size, level) */
  AOA_ERR_HIP = 2,         /* a HIP runtime call failed */
  AOA_ERR_LEVEL = 3,       /* level 0: no more rescaling is possible */
  AOA_ERR_UNSUPPORTED = 4, /* N / prec / k not supported */
  AOA_ERR_WORKSPACE = 5    /* workspace smaller than required */
};

/* Transform directions: transform_directions, aoa-ntt/include/newckks/defines.h:4-7. */
enum { AOA_FORWARD = 0, AOA_INVERSE = 1 };

/* Element-wise ops: supported_operations, aoa-ntt/include/newckks/defines.h:14-21. */
enum { AOA_ADD = 0, AOA_SUB = 1, AOA_MUL = 2, AOA_NEGATE = 3, AOA_MULADD = 5 };

/* Message of the last failed call on this thread ("" if none). */
const char* aoa_last_error(void);

/* Parameter set: primes, NTT twiddles and RNS constants, resident in HBM.
 * Replaces CUDAManager::init(k, kl, N, prec) (aoa-ntt/src/cuda/manager.cu:762-800,
 * prime layout :426-464, constants :566-759, NTT tables
 * aoa-ntt/src/cuda/htrans/ntt.cu:750-1052); kl <= 0 selects k+1 as the
 * reference does.  Unlike the reference this is not a process singleton:
 * handles are immutable and may be shared between threads and streams. */
int aoa_params_create(int N, int k, int kl, int prec, aoa_params** out);
/* CUDAManager::destroy (manager.cu:802-821). */
void aoa_params_destroy(aoa_params* params);
int aoa_params_info(const aoa_params* params, int* N, int* k, int* kl, int* prec);
/* The k Q primes followed by the kl special primes (CUDAManager::RNSQPrimes,
 * RNSBPrimes, manager.h:35-36). */
int aoa_params_primes(const aoa_params* params, uint64_t* out, int n);

/* Host copy of one constant table (for checking against independent math).
 * which: 0 forward twiddles [k+kl][N] (psi^brev(i)), 1 inverse twiddles,
 * 2 (Q_l/q_i)^-1 mod q_i [k][k], 3 (Q_l/q_i) mod b_j [k][kl][k] (l, j, i),
 * 4 (B/b_j)^-1 mod b_j [kl], 5 (B/b_j) mod q_i [k][kl], 6 B^-1 mod q_i [k],
 * 7 B mod q_i [k], 8 q_l^-1 mod q_i [k][k], 9 Barrett MU floor(4^nb/p) [k+kl],
 * 11 encoder roots exp(2 pi i t / 2N) [2N][re, im] (IEEE double bit patterns),
 * 12 rotation group 5^j mod 2N [N/2], 13 Gaussian table of aoa_sample [64],
 * 14 ModDown product constants [k][kl][8] for c = (B/b_j) mod q_i (c mod 2^28,
 * c >> 28, their sum, the 32-bit words w0 w1 w2 of floor(c 2^96 / q_i), q_i's
 * screen threshold, 2 (the high word of the screen's bias 2^33); all zero for
 * primes of 57 bits or more), 15 [kl][2]: N^-1 (B/b_j)^-1 and psi^-brev(1) N^-1
 * (B/b_j)^-1 mod b_j (the last inverse stage of the special limbs of a batch's
 * key products).
 * Returns the number of words written (or needed, if out is NULL). */
int64_t aoa_params_table(const aoa_params* params, int which, uint64_t* out, int64_t n);

/* Forward / inverse negacyclic NTT of limbs res0 .. res0+nres-1 of the
 * polynomial at d_data (limb r at d_data + r*N, prime r).  Output order is the
 * reference's bit-reversed order.  Replaces COMMONEngine::execute
 * (aoa-ntt/src/cuda/htrans/common.cu:397-410 -> ntt.cu:463-475). */
int aoa_ntt(const aoa_params* params, void* stream, uint64_t* d_data, int res0, int nres, int dir);

/* out = a op b (MULADD: out = a*b + c) on limbs res0 .. res0+nres-1 (limb r
 * uses prime r).  Replaces NTTEngine::execute_op / execute_seqop
 * (aoa-ntt/src/cuda/htrans/ntt.cu:511-529, :650-667). */
int aoa_eltwise(const aoa_params* params, void* stream, int op, uint64_t* out, const uint64_t* a,
                const uint64_t* b, const uint64_t* c, int res0, int nres);

/* out = a op b with a 64-bit scalar b (op AOA_ADD, AOA_SUB or AOA_MUL) on
 * limbs res0 .. res0+nres-1, b used as given, not reduced: the reference's
 * formulas, including their results when b >= q_r (one conditional
 * correction for add/sub, its Barrett reduction for mul).  Replaces
 * NTTEngine::execute_op_by_uint (aoa-ntt/src/cuda/htrans/ntt.cu:531-580), the
 * kernel under poly_add / poly_sub / poly_mul by a uint64_t
 * (aoa-ntt/src/arithmetic/poly_t.cpp:157-235). */
int aoa_eltwise_scalar(const aoa_params* params, void* stream, int op, uint64_t* out, const uint64_t* a, uint64_t b,
                       int res0, int nres);

/* Conjugation of limbs 0..nres-1 in the NTT domain: out[i] = in[N-1-i] per
 * limb, the automorphism X -> X^-1 in the reference's bit-reversed order.
 * Replaces execute_conjugation (aoa-ntt/src/ckks/encoder.cu:391-427).  out
 * and in must not overlap. */
int aoa_conjugate(const aoa_params* params, void* stream, uint64_t* out, const uint64_t* in, int nres);

/* CKKS encoder on the device (FP64).  `vals` is a device buffer of n complex
 * values (re, im interleaved doubles, 16-byte aligned), transformed in place.
 *
 * aoa_special_fft: the canonical-embedding FFT of n (a power of two <= N/2)
 * values; AOA_FORWARD is the decoding direction, fftSpecial
 * (aoa-ntt/src/ckks/encoder.cu:33-73), AOA_INVERSE the encoding direction
 * including the division by n, fftSpecialInv (encoder.cu:75-124).  Same loop
 * order, twiddles and rounding as those host loops.
 *
 * aoa_encode: ckks_encode (encoder.cu:143-221) after the caller has placed
 * the `slots` values (filled, then zeros) in `vals`: the inverse FFT in
 * place, then value*scale truncated to int64 and reduced per prime into
 * limbs 0..nres-1 of `out` at coefficients s*gap (real) and N/2 + s*gap
 * (imaginary), gap = N/2/slots, zeros elsewhere.  A negative value whose
 * magnitude is a multiple of q_r is written as 0 (the reference writes q_r,
 * the same residue).  |value*scale| must be below 2^63.
 *
 * aoa_decode: ckks_decode (encoder.cu:223-298): limb 0 of `in` (coefficient
 * domain) at s*gap and N/2 + s*gap, centred and divided by scale, then the
 * forward FFT, into `vals` (slots values). */
int aoa_special_fft(const aoa_params* params, void* stream, double* vals, int n, int dir);
int aoa_encode(const aoa_params* params, void* stream, uint64_t* out, double* vals, int slots, uint64_t scale,
               int nres);
int aoa_decode(const aoa_params* params, void* stream, double* vals, const uint64_t* in, int slots,
               uint64_t scale);

/* Sampler distributions: Sampler::sample_uniform, sample_DG, sample_Z0,
 * sample_hw and sample_narrow (aoa-ntt/include/newckks/cuda/sampler.h:109-119). */
enum { AOA_SAMPLE_UNIFORM = 0, AOA_SAMPLE_GAUSSIAN = 1, AOA_SAMPLE_ZO = 2, AOA_SAMPLE_HW = 3, AOA_SAMPLE_NARROW = 4 };

/* Public-key encryption of one plaintext (ckks_encrypt_poly,
 * aoa-ntt/src/ckks/ckks.cpp:87-105): m holds the plaintext on the k Q limbs in
 * the coefficient domain and is left in the NTT domain; u (ZO), e0 and e1
 * (Gaussian) are drawn with aoa_sample's generator at nonces nonce, nonce+1
 * and nonce+2 on limbs 0..k-1; then c0 = pk_b u + e0 + m and
 * c1 = pk_a u + e1 in the NTT domain.  The same as aoa_sample, aoa_ntt and
 * aoa_eltwise called in that order, in three launches.  ws: at least
 * 3 k N 8 bytes (u, e0, e1). */
int aoa_encrypt(const aoa_params* params, void* stream, uint64_t* c0, uint64_t* c1, uint64_t* m,
                const uint64_t* pk_b, const uint64_t* pk_a, uint64_t seed, uint64_t nonce, void* ws, size_t ws_bytes);

/* Decryption to the plaintext polynomial (ckks_decrypt_poly,
 * aoa-ntt/src/ckks/ckks.cpp:179-192): c0 and c1 (coefficient domain, limbs
 * 0..level) are transformed to the NTT domain in place, then
 * m = INTT(c1 s + c0) on limbs 0..level, s being the NTT-domain secret key. */
int aoa_decrypt(const aoa_params* params, void* stream, uint64_t* m, uint64_t* c0, uint64_t* c1,
                const uint64_t* s, int level);

/* Draw limbs res0 .. res0+nres-1 of the polynomial at `out` (limb r at
 * out + r*N, prime r) on the device.  Replaces the Sampler methods
 * (aoa-ntt/src/cuda/sampler.cu:21-225), which draw from cuRAND and host rand().
 * Here every value is a function of (seed, nonce, distribution, limb,
 * coefficient) through Philox4x32-10: the same arguments give the same
 * polynomial, and a caller takes a fresh nonce per draw.  UNIFORM draws each
 * residue uniformly in [0, q_r); GAUSSIAN draws e = round(3.2 z), z ~ N(0,1);
 * ZO draws 0, +1, -1 with probabilities 1/2, 1/4, 1/4; NARROW draws -1, 0, 1
 * uniformly; HW places exactly 64 values +-1 at distinct positions, zeros
 * elsewhere.  The small distributions draw one value per coefficient, the
 * same on every limb (negatives as q_r + v).  Coefficient domain. */
int aoa_sample(const aoa_params* params, void* stream, uint64_t* out, int res0, int nres, int kind, uint64_t seed,
               uint64_t nonce);

/* Rescale c0 and c1 in place from `level` (coefficient domain): limb i < level
 * becomes (c_i - c_level) q_level^-1 mod q_i.  Replaces
 * CUDAManager::execute_rescale (aoa-ntt/src/cuda/manager.cu:90-139). */
int aoa_rescale(const aoa_params* params, void* stream, uint64_t* c0, uint64_t* c1, int level);

/* ModUp in place on a QB polynomial: reads Q limbs 0..level (coefficient
 * domain) and writes its kl special limbs.  Replaces CUDAManager::execute_modup
 * (aoa-ntt/src/cuda/manager.cu:178-249). */
int aoa_modup(const aoa_params* params, void* stream, uint64_t* d_qb, int level);

/* ModUp from a separate input: the special limbs of the QB polynomial d_qb
 * from Q limbs 0..level of d_q (coefficient domain; d_q may equal d_qb, or be
 * any Q polynomial not overlapping those special limbs).  Replaces
 * CUDAManager::execute_modup(ctx, b, a, level) with b != a
 * (aoa-ntt/include/newckks/cuda/manager.h:147-151, manager.cu:225-249). */
int aoa_modup_from(const aoa_params* params, void* stream, uint64_t* d_qb, const uint64_t* d_q, int level);

/* c = sum_{i<k} a[i] b[i] on every residue of QB polynomials (k+kl limbs):
 * products summed in 128 bits, reduced once with the reference's Barrett
 * reduction (a sum outside its exact range gives the reference's value).
 * a and b are DEVICE arrays of k device pointers, as the reference's kernel
 * reads them.  Replaces CUDAManager::execute_dot
 * (aoa-ntt/include/newckks/cuda/manager.h:128-133, manager.cu:53-84). */
int aoa_dot(const aoa_params* params, void* stream, uint64_t* c, const uint64_t* const* a,
            const uint64_t* const* b, int k);

/* The reference's dr2 on every residue of QB polynomials: c21 = a01 a11,
 * c22 = (a01 a12)(a02 a11) (a product, as the reference computes it),
 * c23 = a02 a12, each product reduced with the reference's Barrett.
 * Replaces CUDAManager::execute_dr2 (manager.h:158-166, manager.cu:370-420). */
int aoa_dr2(const aoa_params* params, void* stream, uint64_t* c21, uint64_t* c22, uint64_t* c23,
            const uint64_t* a01, const uint64_t* a02, const uint64_t* a11, const uint64_t* a12);

/* Host only: the k Q primes followed by the kl special primes (kl <= 0: k+1)
 * of the reference's layout for prec, as CUDAManager::gen_rns_primes selects
 * them (aoa-ntt/src/cuda/manager.cu:426-530).  n >= k + kl. */
int aoa_rns_primes(int k, int kl, int prec, uint64_t* out, int n);

/* ModDown: d_q_out limbs 0..level <- ModDown of the QB polynomial d_qb_in
 * (coefficient domain).  Replaces CUDAManager::execute_moddown
 * (aoa-ntt/src/cuda/manager.cu:312-367), including its Barrett reduction of
 * the cross-base products, so residues are bit-identical. */
int aoa_moddown(const aoa_params* params, void* stream, uint64_t* d_q_out, const uint64_t* d_qb_in,
                int level);

/* Galois automorphism X -> X^(5^rot_slots mod 2N) of limbs 0..nres-1 in the
 * coefficient domain (negated on wrap): rotate_slots_left's kernel
 * (aoa-ntt/src/ckks/encoder.cu:300-353).  out and in must not overlap. */
int aoa_automorphism(const aoa_params* params, void* stream, uint64_t* out, const uint64_t* in, int nres,
                     int rot_slots);

/* Workspace of aoa_keyswitch / aoa_rotate_left / aoa_hommul_relin_rescale at
 * `level`. */
size_t aoa_workspace_bytes(const aoa_params* params, int level);

/* Key switching of d_in (Q limbs 0..level, coefficient domain) with the QB
 * switching key (key_b, key_a) (NTT domain, k+kl limbs each):
 * out0 = ModDown(ModUp(d) * key_b) (+ add0 if non-null), out1 = ModDown(ModUp(d)
 * * key_a), coefficient domain, limbs 0..level.  The key-switch half of
 * ckks_rotate_left / ckks_mul_without_rescale (aoa-ntt/src/ckks/ckks.cpp:302-315,
 * :982-1009). */
int aoa_keyswitch(const aoa_params* params, void* stream, uint64_t* out0, uint64_t* out1,
                  const uint64_t* add0, const uint64_t* d_in, const uint64_t* key_b,
                  const uint64_t* key_a, int level, void* workspace, size_t ws_bytes);

/* HomMul + relinearisation + rescale: (c3_0, c3_1) <- rescale(relin(c1 * c2))
 * at `level` (the minimum level of the inputs, >= 1).  Inputs in the NTT
 * domain (limbs 0..level); outputs in the coefficient domain, limbs
 * 0..level-1 (new level = level-1).  evk in the NTT domain.  Replaces
 * ckks_mul (aoa-ntt/src/ckks/ckks.cpp:273-358).  Outputs must not alias inputs. */
int aoa_hommul_relin_rescale(const aoa_params* params, void* stream, uint64_t* c3_0, uint64_t* c3_1,
                             const uint64_t* c1_0, const uint64_t* c1_1, const uint64_t* c2_0,
                             const uint64_t* c2_1, const uint64_t* evk_b, const uint64_t* evk_a, int level,
                             void* workspace, size_t ws_bytes);

/* aoa_hommul_relin_rescale without a workspace argument, for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  It has no counterpart in the reference.
 * The workspace is a static device array of aoa_noalloc_scratch_bytes()
 * (144 MiB) per GPU, held from library load until the process ends; shapes
 * whose aoa_workspace_bytes exceed it fail with AOA_ERR_WORKSPACE.  The call
 * itself allocates and frees nothing.  All calls on one GPU share that array:
 * calls that overlap on one device (other streams, other threads) race. */
int aoa_hommul_relin_rescale_noalloc(const aoa_params* params, void* stream, uint64_t* c3_0, uint64_t* c3_1,
                                     const uint64_t* c1_0, const uint64_t* c1_1, const uint64_t* c2_0,
                                     const uint64_t* c2_1, const uint64_t* evk_b, const uint64_t* evk_a,
                                     int level);
size_t aoa_noalloc_scratch_bytes(void);

/* HomMul + relinearisation without the rescale: outputs in the coefficient
 * domain, limbs 0..level (level unchanged).  Replaces ckks_mul_without_rescale
 * (aoa-ntt/src/ckks/ckks.cpp:273-332). */
int aoa_hommul_relin(const aoa_params* params, void* stream, uint64_t* c3_0, uint64_t* c3_1, const uint64_t* c1_0,
                     const uint64_t* c1_1, const uint64_t* c2_0, const uint64_t* c2_1, const uint64_t* evk_b,
                     const uint64_t* evk_a, int level, void* workspace, size_t ws_bytes);

/* rho of a QB polynomial (k+kl limbs): Q limbs multiplied by B mod q_i,
 * special limbs set to zero, any domain.  Replaces CUDAManager::execute_rho
 * (aoa-ntt/src/cuda/manager.cu:149-176), used by switching-key generation. */
int aoa_rho(const aoa_params* params, void* stream, uint64_t* out, const uint64_t* in);

/* Rotation by rot_slots slots to the left: automorphism of both polys
 * (coefficient domain input) + key switch of c1 with rtk (NTT domain QB key).
 * Output coefficient domain, limbs 0..level.  Replaces ckks_rotate_left
 * (aoa-ntt/src/ckks/ckks.cpp:982-1009).  Outputs must not alias inputs. */
int aoa_rotate_left(const aoa_params* params, void* stream, uint64_t* out0, uint64_t* out1,
                    const uint64_t* c0, const uint64_t* c1, const uint64_t* rtk_b, const uint64_t* rtk_a,
                    int rot_slots, int level, void* workspace, size_t ws_bytes);

/* Rotation by rot_slots slots to the right (0 <= rot_slots <= N/2): the
 * rotation to the left by N/2 - rot_slots, with the right rotation key for
 * rot_slots.  Replaces ckks_rotate_right (aoa-ntt/src/ckks/ckks.cpp:953-980,
 * rotate_slots_right at aoa-ntt/src/ckks/encoder.cu:356-363). */
int aoa_rotate_right(const aoa_params* params, void* stream, uint64_t* out0, uint64_t* out1,
                     const uint64_t* c0, const uint64_t* c1, const uint64_t* rtk_b, const uint64_t* rtk_a,
                     int rot_slots, int level, void* workspace, size_t ws_bytes);

/* ---------------------------------------------------------------- tuning --
 * NTT split N = N1 x (N/N1) of a parameter set (N1 in {64, 128, 256}, N/N1 in
 * the same set).  No counterpart in the reference, whose split is fixed per N
 * at compile time (aoa-ntt/include/newckks/cuda/htrans/ntt.h:24-44; every
 * split gives the same residues).  Set before issuing work with the handle. */
int aoa_params_set_ntt_split(aoa_params* params, int n1);
int aoa_params_get_ntt_split(const aoa_params* params);

/* ModDown reduction mode.  0 (default): reproduce the reference's Barrett
 * reduction of the cross-base products (aoa-ntt/src/cuda/manager.cu:331-334 with
 * aoa-ntt/src/cuda/htrans/common.cu:257-291), so residues are bit-identical to
 * the reference's; that reduction returns exact - (2^nb - q_i) for a small
 * fraction of products when the special primes are wider than q_i (prec 48 and
 * 52 layouts) and can make decryption fail, as it does in the reference.
 * 1: exact reduction (no counterpart in the reference; residues then differ
 * from the reference's exactly where its reduction deviates). */
int aoa_params_set_moddown_exact(aoa_params* params, int exact);
int aoa_params_get_moddown_exact(const aoa_params* params);

/* Autotuner: times forward+inverse NTTs of `nlimbs` limbs for every valid
 * split (reps repetitions each, on the given stream, allocating a scratch
 * buffer), stores the fastest split in the handle and writes up to `max`
 * (n1, microseconds) pairs to out_n1/out_us.  Returns the number of
 * candidates, or a negative status. */
int aoa_autotune_ntt(aoa_params* params, void* stream, int nlimbs, int reps, int* out_n1, double* out_us, int max);

/* ---------------------------------------------------------------- batches --
 * Many ciphertexts (mixed levels and parameter sets) evaluated with one
 * stage-fused launch per pipeline stage and per parameter set; the sequence
 * is captured as a hipGraph at the first launch and replayed afterwards.
 * No counterpart in the reference, which evaluates one ciphertext per call
 * (aoa-ntt/src/ckks/ckks.cpp:273-358, :982-1009). */
typedef struct aoa_batch aoa_batch;

/* One HomMul+relin+rescale: inputs in the NTT domain at `level` (>= 1),
 * outputs (coefficient domain, limbs 0..level-1) must not alias any input of
 * the same layer. */
typedef struct {
  const uint64_t* c1[2];
  const uint64_t* c2[2];
  uint64_t* out[2];
  int level;
} aoa_mul_item;

/* One left rotation by rot_slots with its rotation key (NTT domain QB):
 * out = rot(in) (+ in when accumulate != 0, one rotate-and-sum step).
 * Input and output in the coefficient domain; out may equal in. */
typedef struct {
  const uint64_t* in[2];
  uint64_t* out[2];
  const uint64_t* rtk_b;
  const uint64_t* rtk_a;
  int rot_slots;
  int level;
  int accumulate;
} aoa_rot_item;

/* In-place forward (dir = AOA_FORWARD) or inverse NTT of limbs 0..level of
 * both polynomials of a ciphertext: the domain changes of poly_t's state
 * machine (aoa-ntt/src/arithmetic/poly_t.cpp:201-220) inside a batch. */
typedef struct {
  uint64_t* poly[2];
  int level;
  int dir;
} aoa_ntt_item;

int aoa_batch_create(aoa_batch** out);
int aoa_batch_add_ntt(aoa_batch* batch, const aoa_params* params, const aoa_ntt_item* items, int n);
int aoa_batch_add_mul(aoa_batch* batch, const aoa_params* params, const uint64_t* evk_b, const uint64_t* evk_a,
                      const aoa_mul_item* items, int n);
int aoa_batch_add_rotate(aoa_batch* batch, const aoa_params* params, const aoa_rot_item* items, int n);
/* Operations added after a barrier start after all earlier ones finish. */
int aoa_batch_barrier(aoa_batch* batch);
/* Allocates the batch's descriptor tables and workspace (once). */
int aoa_batch_build(aoa_batch* batch);
size_t aoa_batch_workspace_bytes(const aoa_batch* batch);
/* Runs the batch on `stream`; use_graph != 0 captures/replays a hipGraph. */
int aoa_batch_launch(aoa_batch* batch, void* stream, int use_graph);
void aoa_batch_destroy(aoa_batch* batch);

/* Launch options of a batch (no counterpart in the reference, whose NTT
 * factorisation is a fixed table, aoa-ntt/include/newckks/cuda/htrans/ntt.h:13-44).
 * Every choice gives the same residues; aoa_batch_autotune times them. */
typedef struct {
  int ntt_split;    /* N1 of the two-pass NTT for the groups whose ring admits it (0: the handle's split) */
  int prime_major;  /* 1: the descriptors of a stage sorted by prime (default); 0: item after item */
  int xcd_chunks;   /* 1: the NTT passes deal their tiles out to the 8 XCDs in contiguous chunks */
  int strict_split; /* limbs of the 63-bit q0 in a launch of their own: -1 never, 0 from 256
                       descriptors per stage up (default), 1 always */
} aoa_batch_options;
void aoa_batch_default_options(aoa_batch_options* opt);
/* Sets the options; a built batch refills its descriptor tables and
 * captures a new graph at its next graph launch. */
int aoa_batch_set_options(aoa_batch* batch, const aoa_batch_options* opt);
int aoa_batch_get_options(const aoa_batch* batch, aoa_batch_options* opt);

/* One candidate of aoa_batch_autotune: its options, milliseconds per graph
 * replay, and a position-weighted checksum of every output residue. */
typedef struct {
  aoa_batch_options opt;
  double ms;
  uint64_t checksum;
  int exact; /* 1: the checksum equals the first candidate's (0 when the batch's outputs
                are not reproducible between replays, e.g. in-place rotate-and-sum) */
} aoa_tune_result;

/* Autotuner: runs every candidate (each NTT split valid for every ring of
 * the batch, x prime-major or item order, x XCD chunks off / on, x the three
 * strict-split rules), each as `reps` graph replays after one warm-up
 * replay, keeps the fastest candidate whose outputs equal the first
 * candidate's and sets it on the batch.  Writes up to max results; returns
 * the number of candidates, or -status.  Allocates a few bytes of device
 * memory for the checksums; synchronises `stream`. */
int aoa_batch_autotune(aoa_batch* batch, void* stream, int reps, aoa_tune_result* results, int max);

/* Pipeline stages of a batch, in launch order inside each layer.  NTT stages
 * are two kernels (column and row pass); the others one kernel. */
enum {
  AOA_STAGE_PRE_FWD = 0,      /* forward NTT of inputs given in the coefficient domain */
  AOA_STAGE_PRE_INV = 1,      /* inverse NTT of inputs given in the NTT domain */
  AOA_STAGE_TENSOR = 2,       /* d0, d1, d2 of HomMul (N <= 4096; from N = 8192 up inside AOA_STAGE_INTT) */
  AOA_STAGE_AUTOMORPHISM = 3, /* rotation's Galois map */
  AOA_STAGE_INTT = 4,         /* inverse NTT of d0, d1, d2 (from N = 8192 up: of d2 only, with the tensor product) */
  AOA_STAGE_MODUP = 5,
  AOA_STAGE_NTT = 6,          /* forward NTT of the ModUp output */
  AOA_STAGE_KEY_INTT = 7,     /* key product fused into the inverse NTT */
  AOA_STAGE_MODDOWN = 8,      /* ModDown (+ rescale, + the additions of rotations and of HomMul below N = 8192) */
  AOA_NSTAGES = 9
};

/* Measurement helper (no counterpart in the reference): runs the batch once
 * on `stream` without a graph, with a HIP event pair around each stage
 * launch of each layer, waits for it, and returns per stage the summed
 * milliseconds (ms[s]) and the number of timed launches (launches[s]).
 * nstages must be >= AOA_NSTAGES. */
int aoa_batch_stage_times(aoa_batch* batch, void* stream, double* ms, int* launches, int nstages);

/* ---------------------------------------------------------------- serving --
 * Continuous batching of encrypted-op requests.  No counterpart in the
 * reference (one call at a time on one stream,
 * aoa-ntt/include/newckks/tool/context.h:16-49).  A request is a ciphertext
 * and a chain of ops; each HomMul+relin+rescale multiplies the running
 * ciphertext by the request's operand ciphertext (NTT domain, top level),
 * each rotation uses a registered rotation key.  The server owns a device
 * pool (working ciphertexts, admission budget) and a step workspace; keys
 * stay resident (caller memory, registered once). */
typedef struct aoa_server aoa_server;
enum { AOA_OP_MUL = 0, AOA_OP_ROTATE = 1 };
typedef struct {
  int kind;      /* AOA_OP_MUL or AOA_OP_ROTATE */
  int rot_slots; /* AOA_OP_ROTATE: left rotation by this many slots */
} aoa_op;
/* Request states reported by aoa_server_result. */
enum { AOA_REQ_PENDING = 0, AOA_REQ_ACTIVE = 1, AOA_REQ_DONE = 2, AOA_REQ_FAILED = 3 };

int aoa_server_create(size_t pool_bytes, size_t step_ws_bytes, int max_active, aoa_server** out);
/* Registers a parameter set with its evaluation key and rotation keys (NTT
 * domain QB polynomials).  Returns the key-set id (>= 0) or -status. */
int aoa_server_register(aoa_server* server, const aoa_params* params, const uint64_t* evk_b, const uint64_t* evk_a,
                        int nrot, const int* rot_steps, const uint64_t* const* rtk_b, const uint64_t* const* rtk_a);
/* Queues a request; input (c0, c1) at `level` (in_ntt: NTT domain), result
 * copied to (out0, out1) in the coefficient domain at the final level.
 * Inputs and outputs must stay valid until the request is done.  A request
 * without ops returns its input as given. */
int aoa_server_submit(aoa_server* server, int keyset, const uint64_t* c0, const uint64_t* c1, int level, int in_ntt,
                      const uint64_t* m0, const uint64_t* m1, const aoa_op* ops, int nops, uint64_t* out0,
                      uint64_t* out1, int64_t* req_id);
/* Runs one step on `stream` (NULL: the server's own stream) when no server
 * thread is running; returns requests still queued or active, or -status. */
int aoa_server_step(aoa_server* server, void* stream);
/* Background serving thread (own stream). */
int aoa_server_start(aoa_server* server);
int aoa_server_stop(aoa_server* server);
/* State, final level (once DONE) and latency from submit to completion of a
 * request.  Finished requests stay queryable until released, or until 65 536
 * later requests have finished (the oldest records are dropped first). */
int aoa_server_result(aoa_server* server, int64_t req_id, int* state, int* level, double* latency_ms);
/* Drops the record of a finished (DONE or FAILED) request. */
int aoa_server_release(aoa_server* server, int64_t req_id);
int aoa_server_stats(aoa_server* server, int64_t* completed, int64_t* ops, int64_t* steps, double* span_ms,
                     size_t* pool_used, size_t* key_bytes);
void aoa_server_destroy(aoa_server* server);

#ifdef __cplusplus
}
#endif

#endif /* AOA_H */
