// Sampler (aoa-ntt/include/newckks/cuda/sampler.h:80-122): uniform, discrete
// Gaussian (sigma 3.2), ternary with N/2 nonzeros (Z0) and Hamming weight 64.
// Seeded host generator (std::mt19937_64); streams differ from the
// reference's cuRAND streams, the distributions are the same.
#ifndef NEWCKKS_SAMPLER_H
#define NEWCKKS_SAMPLER_H

#include <stdint.h>

#include <random>
#include <vector>

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
  void set_seed(uint64_t seed) { rng.seed(seed); }

 private:
  Context* ctx;
  std::mt19937_64 rng;
  void upload_small(poly_t* p, poly_bases b, const std::vector<int64_t>& v);
};

#endif
