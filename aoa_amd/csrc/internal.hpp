// Definitions shared by the C-ABI translation units (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/aoa.h"
#include "kernels.hpp"
#include "params.hpp"

struct aoa_params {
  aoa::HostParams host;
  aoa::DevParams dev;
  void* dmem = nullptr;
};

namespace aoa {

// Records the message of a failed call for aoa_last_error and returns code.
int set_error(int code, const std::string& msg);
// AOA_OK, or AOA_ERR_HIP with the pending launch error.
int check_launch(const char* what);
// 5^rot mod 2N: the Galois element of a left rotation by rot slots.
uint64_t galois_of(int rot_slots, int N);

// HomMul + relinearisation (+ rescale when rescale != 0) of (a0, a1) and
// (b0, b1) at `level` into (c3_0, c3_1), with ws_bytes of caller workspace.
int hommul_core(const aoa_params* h, void* stream, uint64_t* c3_0, uint64_t* c3_1, const uint64_t* a0,
                const uint64_t* a1, const uint64_t* b0, const uint64_t* b1, const uint64_t* evk_b,
                const uint64_t* evk_a, int level, void* ws, size_t ws_bytes, int rescale);

// Batch planning with caller-provided memory (the serving loop rebuilds a
// batch every step from preallocated workspace and descriptor buffers).
void batch_sizes(const aoa_batch* b, size_t* ws_bytes, size_t* desc_bytes);
size_t batch_fill(aoa_batch* b, uint64_t* ws, char* host, char* dev, size_t desc_cap);
struct StageTimer;
int batch_enqueue(aoa_batch* b, hipStream_t st, StageTimer* tm = nullptr);
int batch_add_mul(aoa_batch* b, const aoa_params* p, const uint64_t* evk_b, const uint64_t* evk_a,
                  const aoa_mul_item* items, int n);
int batch_add_rotate(aoa_batch* b, const aoa_params* p, const aoa_rot_item* items, int n);
int batch_add_ntt(aoa_batch* b, const aoa_params* p, const aoa_ntt_item* items, int n);
void batch_clear(aoa_batch* b);
bool batch_empty(const aoa_batch* b);
void batch_barrier(aoa_batch* b);

}  // namespace aoa
