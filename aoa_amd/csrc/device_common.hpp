// Shared device-side definitions for the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "modarith.hpp"
#include "params.hpp"

namespace aoa {

struct alignas(16) tw2 {
  uint64_t w, q;  // constant and its Shoup quotient
};

// Where the limbs of one launch live.  Limb li of poly y is at
// base[y] + li * N and uses prime index prime_of(li): the first `qlimbs`
// limbs are Q limbs prime0, prime0+1, ..., the rest are special limbs
// k, k+1, ...
struct LimbMap {
  uint64_t* base[4];
  const uint64_t* src[4];  // optional separate input (out-of-place first pass), else == base
  int nlimbs;
  int qlimbs;
  int npolys;
  int prime0;  // prime index of limb 0 when it is a Q limb
};

__device__ __forceinline__ int limb_prime(const LimbMap& m, int k, int li) {
  return li < m.qlimbs ? m.prime0 + li : k + (li - m.qlimbs);
}

__device__ __forceinline__ tw2 load_tw(const uint64_t* table, size_t idx) {
  return reinterpret_cast<const tw2*>(table)[idx];
}

}  // namespace aoa
