// HomMul + relinearisation + rescale for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  The workspace that aoa_hommul_relin_rescale takes from
// its caller lives here in a file-scope device array instead; the caller
// still passes the ciphertexts and the evaluation key.
//
// Each GPU holds its own copy of the array, at its own address, from the
// moment the library's code object is loaded on it until the process ends.
// In host code the array's name is only a stand-in for that address, so the
// address is looked up once per device with hipGetSymbolAddress (with the
// device current) and only that pointer is ever handed to kernels.
//
// Every call on one GPU shares that GPU's array: two calls that overlap on the
// same device (different streams, or different host threads) race on it and
// give wrong results.  Calls on one stream are ordered and safe.  Nothing in
// the workspace has to start at zero: every region (d2 in both domains, d0,
// d1, the special limbs, the two key products) is written before it is read
// within the call, so the previous call's contents are never observed.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.hpp"

namespace {

// aoa_workspace_bytes at the largest shape the builds here run: N = 2^16,
// k = 30, K = 31, top level ((6*30 + 3*31) limbs of 512 KiB = 136.5 MiB).
constexpr size_t kScratchBytes = (size_t)144 << 20;

static __device__ __align__(256) uint64_t scratch[kScratchBytes / sizeof(uint64_t)];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices];  // device address of `scratch`, per device

// The current device's copy of `scratch`, looked up on its first use there.
int scratch_ptr(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return aoa::set_error(AOA_ERR_HIP, "aoa_hommul_relin_rescale_noalloc: no usable current device");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return aoa::set_error(AOA_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                             hipGetErrorString(e));
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return AOA_OK;
}

}  // namespace

extern "C" size_t aoa_noalloc_scratch_bytes(void) { return kScratchBytes; }

extern "C" int aoa_hommul_relin_rescale_noalloc(const aoa_params* h, void* stream, uint64_t* c3_0,
                                                uint64_t* c3_1, const uint64_t* a0, const uint64_t* a1,
                                                const uint64_t* b0, const uint64_t* b1, const uint64_t* evk_b,
                                                const uint64_t* evk_a, int level) {
  if (!h) return aoa::set_error(AOA_ERR_ARG, "null argument");
  if (level < 0 || level >= h->host.k) return aoa::set_error(AOA_ERR_ARG, "level out of range");
  const size_t need = aoa_workspace_bytes(h, level);
  if (need > kScratchBytes)
    return aoa::set_error(AOA_ERR_WORKSPACE,
                          "aoa_hommul_relin_rescale_noalloc: this shape needs " + std::to_string(need) +
                              " bytes of workspace, more than the static scratch of " +
                              std::to_string(kScratchBytes) + " bytes");
  void* ws = nullptr;
  const int rc = scratch_ptr(&ws);
  if (rc != AOA_OK) return rc;
  return aoa::hommul_core(h, stream, c3_0, c3_1, a0, a1, b0, b1, evk_b, evk_a, level, ws, kScratchBytes, 1);
}
