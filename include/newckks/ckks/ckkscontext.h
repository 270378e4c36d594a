// CKKSContext (aoa-ntt/include/newckks/ckks/ckkscontext.h:28-248): a Context
// with the scheme's keychain, encoding and the device workspace of the
// evaluation calls.
#ifndef NEWCKKS_CKKSCONTEXT_H
#define NEWCKKS_CKKSCONTEXT_H

#include <complex>

#include "../arithmetic/poly_t.h"
#include "cipher_t.h"
#include "ckkskeys.h"

class CKKSContext : public Context {
 public:
  CKKSKeychain keys;
  poly_t *u, *m, *e;  // encryption temporaries
  cipher_t caux[3];   // auxiliary ciphertexts (sumslots)
  void* d_ws;         // workspace of aoa_hommul / aoa_rotate at the top level
  size_t ws_bytes;

  CKKSContext();
  ~CKKSContext();

  // scale: in/out; -1 selects 2^scalingfactor.
  void encode(poly_t* a, int64_t* scale, std::complex<double>* val, int slots, int empty_slots = 0);
  void decode(std::complex<double>* val, poly_t* a, int64_t scale, int slots);
};

#endif
