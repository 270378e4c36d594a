// Ciphertext (aoa-ntt/include/newckks/ckks/cipher_t.h:17-27).
#ifndef NEWCKKS_CIPHER_T_H
#define NEWCKKS_CIPHER_T_H

#include "../arithmetic/poly_t.h"

struct ciphertext {
  poly_t c[2];
  int level;
  int slots;
  int64_t scale;
};
typedef ciphertext cipher_t;

void cipher_init(Context* ctx, cipher_t* c, int slots = 1);
void cipher_clear(Context* ctx, cipher_t* ct);
void cipher_free(Context* ctx, cipher_t* c);
void cipher_copy(Context* ctx, cipher_t* b, cipher_t* a);

#endif
