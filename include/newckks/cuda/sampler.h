// Sampler (aoa-ntt/include/newckks/cuda/sampler.h:80-122): uniform, rounded
// Gaussian (sigma 3.2), ternary ZO, Hamming weight 64 and narrow ternary,
// drawn on the device (aoa_sample: Philox4x32-10 keyed by the seed, one
// nonce per draw).  The streams differ from the reference's cuRAND and
// rand() streams; the distributions are the same, except that ZO draws each
// coefficient independently (the reference fixes N/2 nonzeros).
#ifndef NEWCKKS_SAMPLER_H
#define NEWCKKS_SAMPLER_H

#include <stdint.h>

#include "../defines.h"

#define GAUSSIAN_STD_DEVIATION (double)3.2
#define GAUSSIAN_BOUND (double)0
#define HAMMINGWEIGHT 64

class Context;

class Sampler {
 public:
  explicit Sampler(Context* ctx, uint64_t seed = 42);
  void sample_uniform(poly_t* p, poly_bases b);
  void sample_DG(poly_t* p, poly_bases b);
  void sample_Z0(poly_t* p, poly_bases b);
  void sample_hw(poly_t* p, poly_bases b);
  void sample_narrow(poly_t* p, poly_bases b);
  // Restarts the sequence of draws.
  void set_seed(uint64_t s) {
    seed = s;
    nonce = 0;
  }

 private:
  Context* ctx;
  uint64_t seed, nonce;
  void draw(poly_t* p, poly_bases b, int kind);
};

#endif
