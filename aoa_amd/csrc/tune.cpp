// NTT autotuner: chooses the split N = N1 x N2 (tile shapes of the column
// and row passes) per parameter set by timing every candidate on the device.
#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.hpp"

using namespace aoa;

extern "C" int aoa_params_set_ntt_split(aoa_params* h, int n1) {
  if (!h) return set_error(AOA_ERR_ARG, "null params");
  if (n1 != 0 && !ntt_valid_split(h->host.N, n1)) return set_error(AOA_ERR_ARG, "invalid NTT split");
  h->dev.ntt_n1 = n1;
  return AOA_OK;
}

extern "C" int aoa_params_set_moddown_exact(aoa_params* h, int exact) {
  if (!h) return set_error(AOA_ERR_ARG, "null params");
  h->dev.moddown_exact = exact ? 1 : 0;
  return AOA_OK;
}

extern "C" int aoa_params_get_moddown_exact(const aoa_params* h) {
  if (!h) return -set_error(AOA_ERR_ARG, "null params");
  return h->dev.moddown_exact;
}

extern "C" int aoa_params_get_ntt_split(const aoa_params* h) {
  if (!h) return -set_error(AOA_ERR_ARG, "null params");
  return h->dev.ntt_n1 ? h->dev.ntt_n1 : ntt_default_n1(h->host.N);
}

extern "C" int aoa_autotune_ntt(aoa_params* h, void* stream, int nlimbs, int reps, int* out_n1, double* out_us,
                                int max) {
  if (!h) return -set_error(AOA_ERR_ARG, "null params");
  const int R = h->host.k + h->host.K;
  if (nlimbs <= 0 || nlimbs > R) nlimbs = R;
  if (reps <= 0) reps = 10;
  hipStream_t st = (hipStream_t)stream;
  const size_t N = (size_t)h->host.N;
  uint64_t* buf = nullptr;
  if (hipMalloc(&buf, N * nlimbs * sizeof(uint64_t)) != hipSuccess)
    return -set_error(AOA_ERR_HIP, "hipMalloc of tuning buffer failed");
  // zero residues are valid inputs (0 < p); timing does not depend on values
  if (hipMemsetAsync(buf, 0, N * nlimbs * sizeof(uint64_t), st) != hipSuccess) {
    (void)hipFree(buf);
    return -set_error(AOA_ERR_HIP, "hipMemset failed");
  }
  LimbMap m;
  std::memset(&m, 0, sizeof(m));
  m.base[0] = buf;
  m.nlimbs = nlimbs;
  m.npolys = 1;
  m.qlimbs = std::min(nlimbs, h->host.k);
  m.prime0 = 0;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  int best = 0, count = 0;
  double best_us = 1e30;
  for (int n1 : {64, 128, 256, h->host.N}) {
    if (!ntt_valid_split(h->host.N, n1)) continue;
    DevParams P = h->dev;
    P.ntt_n1 = n1;
    for (int w = 0; w < 2; ++w) {
      ntt_forward(P, m, n1, st);
      ntt_inverse(P, m, n1, nullptr, nullptr, st);
    }
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r) {
      ntt_forward(P, m, n1, st);
      ntt_inverse(P, m, n1, nullptr, nullptr, st);
    }
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double us = 1000.0 * ms / reps;
    if (count < max && out_n1 && out_us) {
      out_n1[count] = n1;
      out_us[count] = us;
    }
    ++count;
    if (us < best_us) {
      best_us = us;
      best = n1;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(buf);
  int rc = check_launch("aoa_autotune_ntt");
  if (rc) return -rc;
  h->dev.ntt_n1 = best;
  return count;
}
