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

#endif
