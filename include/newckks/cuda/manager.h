// CUDAManager: the reference's parameter singleton (aoa-ntt/include/newckks/cuda/manager.h:15-167)
// as a facade over one aoa_params handle (include/aoa.h).
#ifndef NEWCKKS_MANAGER_H
#define NEWCKKS_MANAGER_H

#include <stdint.h>

#include <vector>

#include "../../aoa.h"
#include "../defines.h"

class CUDAManager {
 public:
  static int is_init;
  static int N;
  static int dnum;
  static uint64_t scalingfactor;  // prec: the scale is 2^scalingfactor
  static std::vector<uint64_t> RNSQPrimes;
  static std::vector<uint64_t> RNSBPrimes;

  // kl <= 0 selects k + 1.  Throws std::runtime_error on unsupported N/prec.
  static void init(const int k, const int kl, const int N, const uint64_t scalingfactor,
                   engine_types e = NTTTrans);
  static void init(const int k, const int N, const uint64_t scalingfactor, engine_types e = NTTTrans);
  static void destroy();
  static int get_n_residues(poly_bases b) {
    switch (b) {
      case QBase: return (int)RNSQPrimes.size();
      case BBase: return (int)RNSBPrimes.size();
      default: return (int)(RNSQPrimes.size() + RNSBPrimes.size());
    }
  }
  // The underlying C-ABI handle.
  static aoa_params* params();
  // Extension (no reference counterpart): exact ModDown reduction instead of
  // the reference's Barrett reduction of its cross-base products, which makes
  // decryption fail after key switching in the prec-48 layout.  Applies to the
  // current parameter set and to later init() calls.
  static void set_exact_moddown(bool exact);
  static bool exact_moddown;
};

size_t poly_get_size(poly_bases b);

#endif
