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

// Read-only tables seen through the constant address space: loads at a
// wave-uniform index become scalar loads, which the compiler cannot prove
// safe through a generic pointer once the kernel also stores to memory.
// The tables are written only by aoa_params_create, before any kernel runs.
template <class T>
using cptr = const T __attribute__((address_space(4)))*;
template <class T>
__device__ __forceinline__ cptr<T> as_const(const T* p) {
  return (cptr<T>)p;
}

// A pointer read from a descriptor in memory, marked as pointing to global
// memory.  The compiler otherwise treats it as generic and accesses it with
// flat instructions, which count against the LDS / scalar-load counter as
// well, so that every wait for an LDS exchange or a scalar load also waits
// for them, and take 64-bit addresses in vector registers.  The cast is
// undone at once; the compiler carries the address space to every use.
template <class T>
__device__ __forceinline__ T* glob(T* p) {
  return (T*)(T __attribute__((address_space(1)))*)(uintptr_t)p;
}

// Entry i of a read-only table of tw2 / DevPrime through the constant
// address space (word by word: structs cannot be copied out of it).
__device__ __forceinline__ tw2 ld_tw2(const tw2* base, size_t i) {
  const cptr<uint64_t> w = as_const(reinterpret_cast<const uint64_t*>(base + i));
  return tw2{w[0], w[1]};
}
__device__ __forceinline__ uint64_t ld_u64(const uint64_t* base, size_t i) { return as_const(base)[i]; }
// Workgroups are dispatched round-robin over the 8 XCDs (workgroup w runs on
// XCD w % 8).  This map deals the n work items out in 8 contiguous chunks,
// one per XCD (the first n % 8 chunks one item longer): XCD x receives
// exactly the ids w < n with w % 8 == x, as many as its chunk holds, so the
// map is a bijection of [0, n).
__device__ __forceinline__ int xcd_chunk_id(unsigned w, unsigned n) {
  const unsigned q = n >> 3, r = n & 7, x = w & 7;
  return (int)(x * q + (x < r ? x : r) + (w >> 3));
}

// One field of a DevPrime at a wave-uniform index, as a scalar load.
__device__ __forceinline__ uint64_t ld_prime_p(const DevPrime* base, int i) { return as_const(&base[i].p)[0]; }
__device__ __forceinline__ int ld_prime_nb(const DevPrime* base, int i) { return (int)as_const(&base[i].nb)[0]; }
__device__ __forceinline__ DevPrime ld_prime(const DevPrime* base, int i) {
  constexpr int W = (int)(sizeof(DevPrime) / 8);
  const cptr<uint64_t> w = as_const(reinterpret_cast<const uint64_t*>(base + i));
  uint64_t t[W];
#pragma unroll
  for (int x = 0; x < W; ++x) t[x] = w[x];
  DevPrime d;
  __builtin_memcpy(&d, t, sizeof(d));
  return d;
}

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

// Key switching's inverse first pass with both key halves: dst[y] = first
// inverse pass of (src * key[y] + (B mod q) add[y]) for y = 0, 1 (the column
// pass then runs on dst[0] and dst[1] as ordinary NttDesc entries).  add is
// null, or (Q limbs of a two-pass transform only) NTT-domain limbs whose
// ModDown-scaled sum joins the product: ModDown(X + B d) = ModDown(X) + d.
// With ab[0] set (HomMul's Q limbs, two-pass transforms), ab = a0, a1, b0,
// b1 (NTT domain, canonical) stand for src and add: the input is d2 = a1 b1
// and add[y] is d0 = a0 b0 for y = 0, d1 = a0 b1 + a1 b0 for y = 1, computed
// in the pass from the operands, so that the tensor product is never stored.
struct KeyDesc2 {
  const uint64_t* src;
  uint64_t* dst[2];
  const uint64_t* key[2];
  const uint64_t* add[2];
  const uint64_t* ab[4];
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
