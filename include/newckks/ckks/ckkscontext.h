// CKKSContext (aoa-ntt/include/newckks/ckks/ckkscontext.h:28-248): a Context
// with the scheme's keychain, encoding and the device workspace of the
// evaluation calls.
#ifndef NEWCKKS_CKKSCONTEXT_H
#define NEWCKKS_CKKSCONTEXT_H

#include <complex>

#include "../arithmetic/poly_t.h"
#include "cipher_t.h"
#include "ckkskeys.h"

#include <map>
#include <string>
#include <vector>

// Names of the precomputed Maclaurin series (ckkscontext.h:22-27).
static const std::string LN1MINUS = "Log1Minus";  // log(1-x)
static const std::string LOGARITHM = "Logarithm";  // log(x)
static const std::string EXPONENT = "Exponent";    // exp(x)
static const std::string SIN = "Sin";              // sin(x)
static const std::string COS = "Cos";              // cos(x)
static const std::string SIGMOID = "Sigmoid";      // 1/(1+exp(-x))

class CKKSContext : public Context {
 public:
  CKKSKeychain keys;
  poly_t *u, *m, *e;  // encryption temporaries
  cipher_t caux[3];   // auxiliary ciphertexts (sumslots)
  void* d_ws;         // workspace of aoa_hommul / aoa_rotate at the top level
  size_t ws_bytes;
  std::complex<double>* h_val;                       // encoding scratch, host (pinned, N values)
  std::complex<double>* d_val;                       // encoding scratch, device (N values)
  std::map<std::string, double*> maclaurin_coeffs;  // series of exp, sin, cos, sigmoid, logs

  CKKSContext();
  ~CKKSContext();

  // scale: in/out; -1 selects 2^scalingfactor.
  void encode(poly_t* a, int64_t* scale, std::complex<double>* val, int slots, int empty_slots = 0);
  void decode(std::complex<double>* val, poly_t* a, int64_t scale, int slots);

  // {"sk": {"s"}, "pk": {"a", "b"}, "evk": {"a", "b"}}, each a poly_export
  // string of the QB key (ckkscontext.h:141-158, ckkscontext.cpp:41-100).
  json export_keys();
  void load_keys(const json& k);
  void clear_keys();
  // Binary form of the same five keys (no counterpart in the reference).
  std::vector<uint8_t> export_keys_binary();
  void load_keys_binary(const uint8_t* data, size_t size);
};

#endif
