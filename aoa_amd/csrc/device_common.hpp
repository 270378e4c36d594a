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

// Descriptor tables of the batched (stage-fused) launches: one entry per
// limb (NTT, tensor) or per polynomial (ModUp, ModDown).
struct NttDesc {
  const uint64_t* src;  // first-pass input limb (== dst for in place)
  uint64_t* dst;        // output limb
  const uint64_t* key;  // inverse only: key limb to multiply by first (or null)
  int prime;
  int pad;
};

struct TensorDesc {
  const uint64_t* a0;
  const uint64_t* a1;
  const uint64_t* b0;
  const uint64_t* b1;
  uint64_t* d0;
  uint64_t* d1;
  uint64_t* d2;
  int prime;
  int pad;
};

struct ModUpDesc {
  const uint64_t* x;  // Q limbs 0..level, stride N
  uint64_t* outB;     // K special limbs, stride N
  int level;
  int pad;
};

struct ModDownDesc {
  const uint64_t* inQ;
  const uint64_t* inB;
  const uint64_t* add;
  uint64_t* out;
  int level;
  int rescale;
};

struct AutoDesc {
  const uint64_t* in;      // limb of the input (coefficient domain)
  uint64_t* out;           // limb of the output
  const uint64_t* addsrc;  // optional limb added to sigma(in)
  uint64_t ginv;           // inverse Galois element mod 2N
  int prime;
  int pad;
};

__device__ __forceinline__ tw2 load_tw(const uint64_t* table, size_t idx) {
  return reinterpret_cast<const tw2*>(table)[idx];
}

}  // namespace aoa
