// Issue-rate probe for gfx950: wave instructions per second for the integer
// vector instructions the kernels are built from, and for scalar ALU
// instructions issued beside them.  Each thread runs eight independent chains
// of one instruction (inline assembly, registers only) and writes one word at
// the end so that the work is kept; a launch holds 8192 workgroups of 256
// threads (eight waves per SIMD on 256 CUs, four rounds).
//
// Output: one line per probe, "name scalar_per_8 ms T_wave_instr_per_s".
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define CHECK(x)                                                               \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      std::printf("hip error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      return 1;                                                                \
    }                                                                          \
  } while (0)

// KIND 0: v_mad_u64_u32, 1: v_mul_lo_u32, 2: v_add_u32, 3: v_mul_hi_u32,
// 4: v_lshl_add_u64 (64-bit add).  NS scalar adds per eight vector
// instructions (0, 2, 4 or 8).
template <int KIND, int NS>
__global__ __launch_bounds__(256) void k_issue(uint64_t* out, int iters) {
  const uint32_t t = threadIdx.x;
  uint64_t a0 = t, a1 = t * 3 + 1, a2 = t * 5 + 2, a3 = t * 7 + 3;
  uint64_t a4 = t * 11, a5 = t * 13, a6 = t * 17, a7 = t * 19;
  uint32_t b0 = t, b1 = t + 1, b2 = t + 2, b3 = t + 3, b4 = t + 4, b5 = t + 5, b6 = t + 6, b7 = t + 7;
  const uint32_t m = t | 1, m2 = t * 9 + 5;
  uint32_t s0 = blockIdx.x, s1 = blockIdx.x * 3;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      if (KIND == 0) {
        asm volatile(
            "v_mad_u64_u32 %0, vcc, %8, %9, %0\n\tv_mad_u64_u32 %1, vcc, %8, %9, %1\n\t"
            "v_mad_u64_u32 %2, vcc, %8, %9, %2\n\tv_mad_u64_u32 %3, vcc, %8, %9, %3\n\t"
            "v_mad_u64_u32 %4, vcc, %8, %9, %4\n\tv_mad_u64_u32 %5, vcc, %8, %9, %5\n\t"
            "v_mad_u64_u32 %6, vcc, %8, %9, %6\n\tv_mad_u64_u32 %7, vcc, %8, %9, %7"
            : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7)
            : "v"(m), "v"(m2)
            : "vcc");
      } else if (KIND == 4) {
        asm volatile(
            "v_lshl_add_u64 %0, %0, 0, %1\n\tv_lshl_add_u64 %1, %1, 0, %2\n\t"
            "v_lshl_add_u64 %2, %2, 0, %3\n\tv_lshl_add_u64 %3, %3, 0, %4\n\t"
            "v_lshl_add_u64 %4, %4, 0, %5\n\tv_lshl_add_u64 %5, %5, 0, %6\n\t"
            "v_lshl_add_u64 %6, %6, 0, %7\n\tv_lshl_add_u64 %7, %7, 0, %0"
            : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7));
      } else {
#define AOA_P8(OP)                                                                                   \
  asm volatile(OP " %0, %8, %0\n\t" OP " %1, %8, %1\n\t" OP " %2, %8, %2\n\t" OP " %3, %8, %3\n\t" OP \
                  " %4, %8, %4\n\t" OP " %5, %8, %5\n\t" OP " %6, %8, %6\n\t" OP " %7, %8, %7"        \
               : "+v"(b0), "+v"(b1), "+v"(b2), "+v"(b3), "+v"(b4), "+v"(b5), "+v"(b6), "+v"(b7)       \
               : "v"(m))
        if (KIND == 1) AOA_P8("v_mul_lo_u32");
        if (KIND == 2) AOA_P8("v_add_u32");
        if (KIND == 3) AOA_P8("v_mul_hi_u32");
#undef AOA_P8
      }
      if (NS >= 2) asm volatile("s_add_u32 %0, %0, 1\n\ts_add_u32 %1, %1, 3" : "+s"(s0), "+s"(s1) : : "scc");
      if (NS >= 4) asm volatile("s_add_u32 %0, %0, 1\n\ts_add_u32 %1, %1, 3" : "+s"(s0), "+s"(s1) : : "scc");
      if (NS >= 8)
        asm volatile("s_add_u32 %0, %0, 1\n\ts_add_u32 %1, %1, 3\n\ts_add_u32 %0, %0, 1\n\ts_add_u32 %1, %1, 3"
                     : "+s"(s0), "+s"(s1)
                     :
                     : "scc");
    }
  }
  out[(size_t)blockIdx.x * 256 + t] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7 ^ b0 ^ b1 ^ b2 ^ b3 ^ b4 ^ b5 ^ b6 ^ b7 ^ s0 ^ s1;
}

constexpr int kBlocks = 256 * 8 * 4;
constexpr int kIters = 2000;

template <int KIND, int NS>
static int probe(const char* name, uint64_t* d) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  k_issue<KIND, NS><<<kBlocks, 256>>>(d, kIters);  // warm-up
  CHECK(hipDeviceSynchronize());
  CHECK(hipEventRecord(e0));
  k_issue<KIND, NS><<<kBlocks, 256>>>(d, kIters);
  CHECK(hipEventRecord(e1));
  CHECK(hipEventSynchronize(e1));
  float ms = 0;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  // vector wave instructions: 4 waves x iterations x 4 rounds x 8 per block
  const double winstr = (double)kBlocks * 4 * kIters * 4 * 8;
  std::printf("%s %d %.4f %.4f\n", name, NS, ms, winstr / (ms * 1e-3) / 1e12);
  std::fflush(stdout);
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  return 0;
}

int main() {
  uint64_t* d = nullptr;
  CHECK(hipMalloc(&d, (size_t)kBlocks * 256 * sizeof(uint64_t)));
  int rc = 0;
  rc |= probe<2, 0>("v_add_u32", d);
  rc |= probe<0, 0>("v_mad_u64_u32", d);
  rc |= probe<1, 0>("v_mul_lo_u32", d);
  rc |= probe<3, 0>("v_mul_hi_u32", d);
  rc |= probe<4, 0>("v_lshl_add_u64", d);
  rc |= probe<2, 4>("v_add_u32+salu", d);
  rc |= probe<2, 8>("v_add_u32+salu", d);
  rc |= probe<0, 8>("v_mad_u64_u32+salu", d);
  CHECK(hipFree(d));
  return rc;
}
