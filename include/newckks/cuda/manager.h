// CUDAManager: the reference's parameter singleton (aoa-ntt/include/newckks/cuda/manager.h:15-167)
// as a facade over one aoa_params handle (include/aoa.h).
#ifndef NEWCKKS_MANAGER_H
#define NEWCKKS_MANAGER_H

#include <stdint.h>

#include <vector>

#include "../../aoa.h"
#include "../defines.h"

class Context;

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
  // Selects the k Q primes and kl special primes of the layout for
  // scalingfactor (manager.cu:483-564) into RNSQPrimes / RNSBPrimes.  init()
  // calls it; a later call with the initialised sizes leaves them as they are.
  static void gen_rns_primes(unsigned int k, unsigned int kl);
  // The RNS constants are computed when the parameter set is created
  // (aoa_params_create), so this only checks that init() has run.
  static void precompute();
  // Products of the Q and special primes.  The reference returns NTL::ZZ,
  // which this library does not use: here they are little-endian 64-bit words.
  static std::vector<uint64_t> get_q_product();
  static std::vector<uint64_t> get_b_product();

  // The reference's device entry points (manager.h:128-166), on raw residue
  // pointers and the context's stream.  execute_dot: c = sum_i a[i] b[i] over
  // the QB residues, a and b device arrays of k device pointers.  execute_rho:
  // b = rho(a).  execute_rescale: a, b rescaled in place from level.
  // execute_modup: special limbs of the QB polynomial a from Q limbs
  // 0..level of b (a == b in poly_modup).  execute_moddown: Q limbs 0..level
  // of a from the QB polynomial b.  execute_dr2: the reference's dr2 kernel
  // over the QB residues.
  static void execute_dot(Context* ctx, uint64_t* c, uint64_t** a, uint64_t** b, const int k);
  static void execute_rho(Context* ctx, uint64_t* b, uint64_t* a);
  static void execute_rescale(Context* ctx, uint64_t* a, uint64_t* b, const int level);
  static void execute_modup(Context* ctx, uint64_t* a, uint64_t* b, const int level);
  static void execute_moddown(Context* ctx, uint64_t* a, uint64_t* b, const int level);
  static void execute_dr2(Context* ctx, uint64_t* ct21, uint64_t* ct22, uint64_t* ct23, const uint64_t* ct01,
                          const uint64_t* ct02, const uint64_t* ct11, const uint64_t* ct12);

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
