// Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the
// parameter construction of the product (params.cpp), its modular arithmetic
// (modarith.hpp, through the checks of modarith_check.cpp) and the oracle,
// run over small and large shapes.  Built and run by
// tests/test_sanitize.py; any sanitizer report aborts with a non-zero status,
// and a failed check returns 1.
#include <stdint.h>

#include <cstdio>
#include <random>
#include <vector>

#include "params.hpp"

extern "C" {
int64_t chk_exact_mults(int N, int k, int kl, int prec, int64_t trials, uint64_t seed);
int chk_moddown_terms(int N, int k, int kl, int prec, int64_t trials, uint64_t seed, int64_t* out);
int64_t chk_barrett_delta(int N, int k, int kl, int prec, int64_t trials, uint64_t seed);
int64_t chk_moddown_screen_words(int64_t trials, uint64_t seed);
int64_t chk_moddown_accumulate(int N, int k, int kl, int prec, int64_t trials, uint64_t seed);
int64_t chk_moddown_reduce(int N, int k, int kl, int prec, int64_t trials, uint64_t seed);
int64_t chk_moddown_finish(int N, int k, int kl, int prec, int64_t trials, uint64_t seed);
int64_t chk_table(int N, int k, int kl, int prec, int which, uint64_t* out);
void* orc_create(int N, int k, int kl, int prec);
void orc_destroy(void* h);
int orc_set_threads(int n);
int orc_primes(void* h, uint64_t* out);
void orc_ntt(void* h, uint64_t* d, int res0, int nres, int dir);
void orc_modup(void* h, uint64_t* qb, int level);
void orc_moddown(void* h, uint64_t* out, const uint64_t* qb, int level);
void orc_rescale(void* h, uint64_t* a, uint64_t* b, int level);
void orc_hommul(void* h, uint64_t* c30, uint64_t* c31, const uint64_t* c10, const uint64_t* c11, const uint64_t* c20,
                const uint64_t* c21, int in_state, const uint64_t* evkb, const uint64_t* evka, int level, int rescale);
void orc_rotate_left(void* h, uint64_t* o0, uint64_t* o1, const uint64_t* c0, const uint64_t* c1, const uint64_t* rtkb,
                     const uint64_t* rtka, int rot_slots, int level);
void orc_fft_special(double* v, long size, long ringN, int inverse);
void orc_encode(void* h, uint64_t* out, double* vals, int slots, uint64_t scale, int nres);
void orc_decode(void* h, double* vals, const uint64_t* in, int slots, uint64_t scale);
}

static int failures = 0;
static void expect(bool ok, const char* what) {
  if (!ok) {
    ++failures;
    std::printf("FAIL %s\n", what);
  }
}

struct Shape {
  int N, k, kl, prec;
};

int main() {
  orc_set_threads(2);
  const Shape shapes[] = {{2048, 3, 3, 55}, {4096, 4, 5, 52}, {8192, 6, 0, 48}, {16384, 5, 5, 55},
                          {4096, 10, 11, 45}, {65536, 30, 31, 48}, {65536, 30, 64, 48}};
  for (const Shape& s : shapes) {
    aoa::HostParams H;
    expect(aoa::build_host_params(s.N, s.k, s.kl, s.prec, &H).empty(), "build_host_params");
    for (int t = 0; t <= 14; ++t) {
      if (t == 11 && s.N > 16384) continue;  // encoder roots: 2N doubles, skipped at the largest rings
      const int64_t n = chk_table(s.N, s.k, s.kl, s.prec, t, nullptr);
      if (n <= 0) continue;
      std::vector<uint64_t> v((size_t)n);
      expect(chk_table(s.N, s.k, s.kl, s.prec, t, v.data()) == n, "chk_table");
    }
    expect(chk_exact_mults(s.N, s.k, s.kl, s.prec, 200, 1) == 0, "chk_exact_mults");
    int64_t out[4];
    expect(chk_moddown_terms(s.N, s.k, s.kl, s.prec, 2000, 2, out) == 0 && out[0] == 0, "chk_moddown_terms");
    expect(chk_barrett_delta(s.N, s.k, s.kl, s.prec, 500, 3) == 0, "chk_barrett_delta");
    expect(chk_moddown_accumulate(s.N, s.k, s.kl, s.prec, 20, 4) == 0, "chk_moddown_accumulate");
    expect(chk_moddown_reduce(s.N, s.k, s.kl, s.prec, 100, 5) == 0, "chk_moddown_reduce");
    expect(chk_moddown_finish(s.N, s.k, s.kl > 0 ? s.kl : s.k + 1, s.prec, 5, 6) == 0, "chk_moddown_finish");
  }
  expect(chk_moddown_screen_words(2000, 7) == 0, "chk_moddown_screen_words");

  // the oracle over the path at two small shapes
  for (const Shape& s : {Shape{2048, 3, 3, 55}, Shape{8192, 4, 0, 48}}) {
    void* o = orc_create(s.N, s.k, s.kl, s.prec);
    expect(o != nullptr, "orc_create");
    if (!o) continue;
    const int kl = s.kl > 0 ? s.kl : s.k + 1, R = s.k + kl;
    std::vector<uint64_t> primes((size_t)R);
    orc_primes(o, primes.data());
    std::mt19937_64 rng(11);
    auto poly = [&](int nl) {
      std::vector<uint64_t> p((size_t)nl * s.N);
      for (int r = 0; r < nl; ++r)
        for (int i = 0; i < s.N; ++i) p[(size_t)r * s.N + i] = rng() % primes[r];
      return p;
    };
    const int level = s.k - 1;
    std::vector<uint64_t> x = poly(R), y = x;
    orc_ntt(o, y.data(), 0, R, 0);
    orc_ntt(o, y.data(), 0, R, 1);
    expect(x == y, "oracle NTT round trip");
    std::vector<uint64_t> qb = poly(R), out((size_t)s.k * s.N);
    orc_modup(o, qb.data(), level);
    orc_moddown(o, out.data(), qb.data(), level);
    std::vector<uint64_t> a = poly(s.k), b = poly(s.k);
    orc_rescale(o, a.data(), b.data(), level);
    std::vector<uint64_t> c10 = poly(s.k), c11 = poly(s.k), c20 = poly(s.k), c21 = poly(s.k), ek0 = poly(R),
                          ek1 = poly(R), o0((size_t)s.k * s.N), o1((size_t)s.k * s.N);
    orc_hommul(o, o0.data(), o1.data(), c10.data(), c11.data(), c20.data(), c21.data(), 2, ek0.data(), ek1.data(),
               level, 1);
    orc_rotate_left(o, o0.data(), o1.data(), c10.data(), c11.data(), ek0.data(), ek1.data(), 3, level);
    const int slots = s.N / 2;
    std::vector<double> v((size_t)2 * slots), w;
    for (auto& d : v) d = std::uniform_real_distribution<double>(-1, 1)(rng);
    w = v;
    orc_fft_special(w.data(), slots, s.N, 1);
    orc_fft_special(w.data(), slots, s.N, 0);
    double err = 0;
    for (size_t i = 0; i < v.size(); ++i) err = std::max(err, std::abs(v[i] - w[i]));
    expect(err < 1e-9, "oracle special FFT round trip");
    std::vector<uint64_t> m((size_t)s.k * s.N);
    w = v;
    orc_encode(o, m.data(), w.data(), slots, (uint64_t)1 << 40, s.k);
    std::vector<double> d((size_t)2 * slots);
    orc_decode(o, d.data(), m.data(), slots, (uint64_t)1 << 40);
    err = 0;
    for (size_t i = 0; i < v.size(); ++i) err = std::max(err, std::abs(v[i] - d[i]));
    expect(err < 1e-6, "oracle encode/decode");
    orc_destroy(o);
  }
  std::printf("sanitized host checks: %d failure(s)\n", failures);
  return failures ? 1 : 0;
}
