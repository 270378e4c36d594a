// CKKS scheme operations (aoa-ntt/include/newckks/ckks/ckks.h:23-319) on the
// gfx950 evaluation core.  Ownership follows the reference: functions that
// return a cipher_t* allocate it with new (release with cipher_free + delete);
// ckks_decrypt returns a malloc'd array (release with free).
#ifndef NEWCKKS_CKKS_H
#define NEWCKKS_CKKS_H

#include <complex>

#include "cipher_t.h"
#include "ckkscontext.h"
#include "ckkskeys.h"

CKKSKeychain* ckks_keygen(CKKSContext* ctx, SecretKey* sk);

cipher_t* ckks_encrypt_poly(CKKSContext* ctx, poly_t* m);
cipher_t* ckks_encrypt_poly(CKKSContext* ctx, cipher_t* ct, poly_t* m);
cipher_t* ckks_encrypt(CKKSContext* ctx, std::complex<double>* val, int slots = 1, int empty_slots = 0);
cipher_t* ckks_encrypt(CKKSContext* ctx, cipher_t* ct, std::complex<double>* val, int slots = 1,
                       int empty_slots = 0);
cipher_t* ckks_encrypt(CKKSContext* ctx, double val, int slots = 1, int empty_slots = 0);
// Extension (no reference counterpart; the reference fixes the scale at 2^prec,
// aoa-ntt/src/ckks/cipher_t.cpp:7): encrypt with scale 2^scale_bits.
cipher_t* ckks_encrypt_scaled(CKKSContext* ctx, std::complex<double>* val, int slots, int scale_bits);

poly_t* ckks_decrypt_poly(CKKSContext* ctx, poly_t* m, cipher_t* c, SecretKey* sk);
std::complex<double>* ckks_decrypt(CKKSContext* ctx, cipher_t* c, SecretKey* sk);
std::complex<double>* ckks_decrypt(CKKSContext* ctx, std::complex<double>* val, cipher_t* c, SecretKey* sk);

void ckks_rescale(CKKSContext* ctx, cipher_t* c);
cipher_t* ckks_add(CKKSContext* ctx, cipher_t* c1, cipher_t* c2);
void ckks_add(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, cipher_t* c2);
cipher_t* ckks_sub(CKKSContext* ctx, cipher_t* c1, cipher_t* c2);
void ckks_sub(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, cipher_t* c2);
void ckks_mul_without_rescale(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, cipher_t* c2);
void ckks_mul(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, cipher_t* c2);
cipher_t* ckks_mul(CKKSContext* ctx, cipher_t* c1, cipher_t* c2);
void ckks_square(CKKSContext* ctx, cipher_t* c2, cipher_t* c1);
cipher_t* ckks_square(CKKSContext* ctx, cipher_t* c1);
void ckks_rotate_left(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, int rotSlots);
void ckks_rotate_right(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, int rotSlots);
void ckks_sumslots(CKKSContext* ctx, cipher_t* c2, cipher_t* c1);
cipher_t* ckks_sumslots(CKKSContext* ctx, cipher_t* c);
cipher_t* ckks_encrypt(CKKSContext* ctx, cipher_t* ct, double val, int slots = 1, int empty_slots = 0);
cipher_t* ckks_mul_without_rescale(CKKSContext* ctx, cipher_t* c1, cipher_t* c2);

// Plaintext and constant operands (ckks.cpp:497-691).
void ckks_add(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, double val);
cipher_t* ckks_add(CKKSContext* ctx, cipher_t* c1, double val);
void ckks_mul(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, double val);
cipher_t* ckks_mul(CKKSContext* ctx, cipher_t* c1, double val);
void ckks_mul_without_rescale(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, double val);
cipher_t* ckks_mul_without_rescale(CKKSContext* ctx, cipher_t* c1, double val);
void ckks_mul(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, std::complex<double>* val);
cipher_t* ckks_mul(CKKSContext* ctx, cipher_t* c1, std::complex<double>* val);
void ckks_mul_without_rescale(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, poly_t* val);
void ckks_mul(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, poly_t* val);
cipher_t* ckks_mul(CKKSContext* ctx, cipher_t* c1, poly_t* val);

// Compositions (ckks.cpp:423-477, :693-951, :1011-1189).
void ckks_inner_prod(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, cipher_t* c2, const int size);
void ckks_inner_prod(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, double* c2, const int size);
template <typename T>
void ckks_batch_inner_prod(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, T* c2);
void ckks_power(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, uint64_t x);
cipher_t* ckks_power(CKKSContext* ctx, cipher_t* c1, uint64_t x);
void ckks_power_of_2(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, uint64_t x);
cipher_t* ckks_power_of_2(CKKSContext* ctx, cipher_t* c1, uint64_t x);
void ckks_eval_polynomial(CKKSContext* ctx, cipher_t* result, cipher_t* x, double* coeffs, int n);
cipher_t* ckks_eval_polynomial(CKKSContext* ctx, cipher_t* x, double* coeffs, int n);
void ckks_exp(CKKSContext* ctx, cipher_t* result, cipher_t* ct);
cipher_t* ckks_exp(CKKSContext* ctx, cipher_t* ct);
void ckks_sin(CKKSContext* ctx, cipher_t* result, cipher_t* ct);
cipher_t* ckks_sin(CKKSContext* ctx, cipher_t* ct);
void ckks_cos(CKKSContext* ctx, cipher_t* result, cipher_t* ct);
cipher_t* ckks_cos(CKKSContext* ctx, cipher_t* ct);
void ckks_sigmoid(CKKSContext* ctx, cipher_t* result, cipher_t* ct);
cipher_t* ckks_sigmoid(CKKSContext* ctx, cipher_t* ct);
void ckks_log1minus(CKKSContext* ctx, cipher_t* result, cipher_t* ct);
cipher_t* ckks_log1minus(CKKSContext* ctx, cipher_t* ct);
void ckks_log(CKKSContext* ctx, cipher_t* result, cipher_t* ct, int steps = 10);
cipher_t* ckks_log(CKKSContext* ctx, cipher_t* ct, int steps = 10);
void ckks_inverse(CKKSContext* ctx, cipher_t* result, cipher_t* ct, int steps = 10);
cipher_t* ckks_inverse(CKKSContext* ctx, cipher_t* ct, int steps = 10);
void ckks_conjugate(CKKSContext* ctx, cipher_t* c2, cipher_t* c1);
cipher_t* ckks_conjugate(CKKSContext* ctx, cipher_t* ct);
void ckks_merge(CKKSContext* ctx, cipher_t* c_merged, cipher_t* cts, int n);
cipher_t* ckks_merge(CKKSContext* ctx, cipher_t* cts, int n);
void ckks_discard_higher_slots(CKKSContext* ctx, cipher_t* ct);
void ckks_discard_slots_except(CKKSContext* ctx, cipher_t* ct, int idx);

#endif
