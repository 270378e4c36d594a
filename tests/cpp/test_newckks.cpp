// Functional tests of the newckks-compatible C++ API on the GPU, written in
// the style of the reference's own tests (aoa-ntt/tests/tests.cpp:862-1760):
// random slot values, encrypt -> op -> decrypt, compared with the FP64 result
// within the reference's bounds (ERRBOUND 0.005, tests.cpp:34; rotations 0.01,
// tests.cpp:1578, :1635).  Prints one line per case; exit status = failures.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include <newckks/newckks.h>

typedef std::complex<double> cd;
static const double ERRBOUND = 0.005;
static int failures = 0;

static double get_range(uint64_t prec) {  // tests.cpp:36-38
  return (double)(CUDAManager::RNSQPrimes[0] - 1) / (double)((uint64_t)1 << (prec + 1));
}

static void report(const std::string& name, double err, double bound) {
  bool ok = err <= bound && std::isfinite(err);
  if (!ok) ++failures;
  std::printf("%s %s max_err=%.3e bound=%.3e\n", ok ? "PASS" : "FAIL", name.c_str(), err, bound);
  std::fflush(stdout);
}

static double maxerr(const cd* a, const cd* b, int n) {
  double e = 0;
  for (int i = 0; i < n; ++i) e = std::max(e, std::max(std::abs(a[i].real() - b[i].real()), std::abs(a[i].imag() - b[i].imag())));
  return e;
}

static void run_params(int k, int kl, int N, int prec, bool heavy, bool exact = false) {
  CUDAManager::set_exact_moddown(exact);
  CUDAManager::init(k, kl, N, prec);
  CKKSContext* ctx = new CKKSContext();
  SecretKey* sk = ckks_new_sk(ctx);
  CKKSKeychain* keys = new CKKSKeychain;
  keys->pk = ckks_new_pk(ctx, sk);
  keys->evk = ckks_new_mtk(ctx, sk);
  std::vector<int> steps;
  for (int r = 1; r < (N >> 1); r *= 2) steps.push_back(r);
  keys->rtk_left = ckks_new_rtk_left(ctx, sk, steps);
  ctx->keys = *keys;
  const int slots = N >> 1;
  const std::string tag = "N" + std::to_string(N) + "_k" + std::to_string(k) + "_kl" + std::to_string(kl) + "_prec" +
                          std::to_string(prec) + (exact ? "_exactmoddown" : "");
  std::default_random_engine gen;
  const double R = get_range(prec);
  std::uniform_real_distribution<double> dist(-R, R), dsq(-std::sqrt(R), std::sqrt(R));
  std::vector<cd> v1(slots), v2(slots), exp(slots);

  // EncodeDecode (tests.cpp:862-918)
  {
    for (auto& x : v1) x = cd(dist(gen), dist(gen));
    poly_t m;
    int64_t scale = -1;
    ctx->encode(&m, &scale, v1.data(), slots);
    std::vector<cd> out(slots);
    ctx->decode(out.data(), &m, scale, slots);
    report("EncodeDecode_" + tag, maxerr(v1.data(), out.data(), slots), ERRBOUND);
    poly_free(ctx, &m);
  }
  // EncryptDecrypt (tests.cpp:920-967)
  {
    for (auto& x : v1) x = cd(dist(gen), dist(gen));
    cipher_t* ct = ckks_encrypt(ctx, v1.data(), slots);
    cd* out = ckks_decrypt(ctx, ct, sk);
    report("EncryptDecrypt_" + tag, maxerr(v1.data(), out, slots), ERRBOUND);
    free(out);
    cipher_free(ctx, ct);
    delete ct;
  }
  // Add / Sub (tests.cpp:969-1097)
  {
    std::uniform_real_distribution<double> dh(-R / 2, R / 2);
    for (int i = 0; i < slots; ++i) {
      v1[i] = cd(dh(gen), dh(gen));
      v2[i] = cd(dh(gen), dh(gen));
    }
    cipher_t* a = ckks_encrypt(ctx, v1.data(), slots);
    cipher_t* b = ckks_encrypt(ctx, v2.data(), slots);
    cipher_t* c = ckks_add(ctx, a, b);
    cd* out = ckks_decrypt(ctx, c, sk);
    for (int i = 0; i < slots; ++i) exp[i] = v1[i] + v2[i];
    report("Add_" + tag, maxerr(exp.data(), out, slots), ERRBOUND);
    free(out);
    cipher_t* d = ckks_sub(ctx, a, b);
    out = ckks_decrypt(ctx, d, sk);
    for (int i = 0; i < slots; ++i) exp[i] = v1[i] - v2[i];
    report("Sub_" + tag, maxerr(exp.data(), out, slots), ERRBOUND);
    free(out);
    for (cipher_t* x : {a, b, c, d}) {
      cipher_free(ctx, x);
      delete x;
    }
  }
  // Mul, Square, MulWithoutRescale + Rescale (tests.cpp:1099-1222)
  {
    for (int i = 0; i < slots; ++i) {
      v1[i] = cd(dsq(gen), dsq(gen));
      v2[i] = cd(dsq(gen), dsq(gen));
    }
    cipher_t* a = ckks_encrypt(ctx, v1.data(), slots);
    cipher_t* b = ckks_encrypt(ctx, v2.data(), slots);
    cipher_t* c = ckks_mul(ctx, a, b);
    cd* out = ckks_decrypt(ctx, c, sk);
    for (int i = 0; i < slots; ++i) exp[i] = v1[i] * v2[i];
    report("Mul_" + tag, maxerr(exp.data(), out, slots), ERRBOUND);
    free(out);
    cipher_t* s = ckks_square(ctx, a);
    out = ckks_decrypt(ctx, s, sk);
    for (int i = 0; i < slots; ++i) exp[i] = v1[i] * v1[i];
    report("Square_" + tag, maxerr(exp.data(), out, slots), ERRBOUND);
    free(out);
    cipher_t w;
    cipher_init(ctx, &w);
    ckks_mul_without_rescale(ctx, &w, a, b);
    ckks_rescale(ctx, &w);
    out = ckks_decrypt(ctx, &w, sk);
    for (int i = 0; i < slots; ++i) exp[i] = v1[i] * v2[i];
    report("MulWithoutRescale_Rescale_" + tag, maxerr(exp.data(), out, slots), ERRBOUND);
    free(out);
    cipher_free(ctx, &w);
    for (cipher_t* x : {a, b, c, s}) {
      cipher_free(ctx, x);
      delete x;
    }
  }
  // MulSequence (tests.cpp:1701-1760): all-ones vectors to depth k-1
  if (heavy) {
    std::vector<cd> ones(slots, cd(1.0, 0.0));
    cipher_t* a = ckks_encrypt(ctx, ones.data(), slots);
    cipher_t* b = ckks_encrypt(ctx, ones.data(), slots);
    double worst = 0;
    for (int d = 0; d < k - 1; ++d) {
      ckks_mul(ctx, b, b, a);
      cd* out = ckks_decrypt(ctx, b, sk);
      worst = std::max(worst, maxerr(ones.data(), out, slots));
      free(out);
    }
    report("MulSequence_" + tag, worst, ERRBOUND);
    for (cipher_t* x : {a, b}) {
      cipher_free(ctx, x);
      delete x;
    }
  }
  // RotateLeft (tests.cpp:1594-1649) and sumslots
  {
    for (auto& x : v1) x = cd(dist(gen), dist(gen));
    cipher_t* a = ckks_encrypt(ctx, v1.data(), slots);
    double worst = 0;
    for (int r : {1, 2, 4, 64}) {
      if (r >= slots) continue;
      cipher_t c;
      cipher_init(ctx, &c);
      ckks_rotate_left(ctx, &c, a, r);
      cd* out = ckks_decrypt(ctx, &c, sk);
      for (int i = 0; i < slots; ++i) exp[i] = v1[(i + r) % slots];
      worst = std::max(worst, maxerr(exp.data(), out, slots));
      free(out);
      cipher_free(ctx, &c);
    }
    report("RotateLeft_" + tag, worst, 0.01);
    // sumslots over 16 slots (rotate-and-add, ckks.cpp:479-495)
    const int ns = 16;
    std::vector<cd> sv(ns);
    std::uniform_real_distribution<double> du(-1, 1);
    cd total = 0;
    for (auto& x : sv) {
      x = cd(du(gen), du(gen));
      total += x;
    }
    cipher_t* b = ckks_encrypt(ctx, sv.data(), ns);
    cipher_t* s = ckks_sumslots(ctx, b);
    cd* out = ckks_decrypt(ctx, s, sk);
    std::vector<cd> e2(ns, total);
    report("SumSlots_" + tag, maxerr(e2.data(), out, ns), 10 * ERRBOUND);
    free(out);
    for (cipher_t* x : {a, b, s}) {
      cipher_free(ctx, x);
      delete x;
    }
  }
  // level 0 cannot be rescaled (ckks.cpp:224-227)
  {
    cipher_t* a = ckks_encrypt(ctx, 1.0, slots);
    a->level = 0;
    bool threw = false;
    try {
      ckks_rescale(ctx, a);
    } catch (const std::runtime_error&) {
      threw = true;
    }
    report("RescaleLevel0Throws_" + tag, threw ? 0.0 : 1.0, 0.0);
    cipher_free(ctx, a);
    delete a;
  }
  ckkskeychain_free(ctx, keys);
  delete keys;
  ckks_free_sk(ctx, sk);
  delete sk;
  delete ctx;
  CUDAManager::destroy();
  CUDAManager::set_exact_moddown(false);
}

// The reference's own ModDown reduction at prec 48 (special primes 55-bit,
// wider than q_i): reported, not checked -- it makes decryption after a
// HomMul fail once a deviating product reaches the rescaled limbs, and the
// bit-exact default reproduces that.
static void reference_mode_prec48() {
  CUDAManager::set_exact_moddown(false);
  CUDAManager::init(30, 30, 4096, 48);
  CKKSContext* ctx = new CKKSContext();
  SecretKey* sk = ckks_new_sk(ctx);
  CKKSKeychain keys;
  keys.pk = ckks_new_pk(ctx, sk);
  keys.evk = ckks_new_mtk(ctx, sk);
  ctx->keys = keys;
  const int slots = 2048;
  std::vector<cd> v(slots, cd(1.0, 0.5)), e(slots, cd(1.0, 0.5) * cd(1.0, 0.5));
  cipher_t* a = ckks_encrypt(ctx, v.data(), slots);
  cipher_t* c = ckks_mul(ctx, a, a);
  cd* out = ckks_decrypt(ctx, c, sk);
  std::printf("INFO ReferenceModDown_Mul_N4096_k30_kl30_prec48 max_err=%.3e (not checked)\n", maxerr(e.data(), out, slots));
  free(out);
  for (cipher_t* x : {a, c}) {
    cipher_free(ctx, x);
    delete x;
  }
  ckkskeychain_free(ctx, &keys);
  ckks_free_sk(ctx, sk);
  delete sk;
  delete ctx;
  CUDAManager::destroy();
}

// N=4096, 4 limbs (prec 52), K=5; 2048 slots uniform in [-1,1];
// encode -> encrypt -> HomMul + relin + rescale -> decrypt -> decode, against
// the FP64 product, at the reference's scale 2^52 and at the explicit 2^40.
static void mul_n4096_k4() {
  CUDAManager::init(4, 5, 4096, 52);
  CKKSContext* ctx = new CKKSContext();
  SecretKey* sk = ckks_new_sk(ctx);
  CKKSKeychain keys;
  keys.pk = ckks_new_pk(ctx, sk);
  keys.evk = ckks_new_mtk(ctx, sk);
  ctx->keys = keys;
  const int slots = 2048;
  std::mt19937_64 rng(1);
  std::uniform_real_distribution<double> d(-1, 1);
  std::vector<cd> v1(slots), v2(slots), exp(slots);
  for (int i = 0; i < slots; ++i) {
    v1[i] = cd(d(rng), d(rng));
    v2[i] = cd(d(rng), d(rng));
    exp[i] = v1[i] * v2[i];
  }
  for (int bits : {52, 40}) {
    cipher_t* a = bits == 52 ? ckks_encrypt(ctx, v1.data(), slots) : ckks_encrypt_scaled(ctx, v1.data(), slots, bits);
    cipher_t* b = bits == 52 ? ckks_encrypt(ctx, v2.data(), slots) : ckks_encrypt_scaled(ctx, v2.data(), slots, bits);
    cipher_t* c = ckks_mul(ctx, a, b);
    cd* out = ckks_decrypt(ctx, c, sk);
    report("MulN4096K4_scale2^" + std::to_string(bits), maxerr(exp.data(), out, slots), ERRBOUND);
    free(out);
    for (cipher_t* x : {a, b, c}) {
      cipher_free(ctx, x);
      delete x;
    }
  }
  ckkskeychain_free(ctx, &keys);
  ckks_free_sk(ctx, sk);
  delete sk;
  delete ctx;
  CUDAManager::destroy();
}

int main(int argc, char** argv) {
  const bool quick = argc > 1 && std::string(argv[1]) == "quick";
  mul_n4096_k4();
  run_params(3, 3, 4096, 55, true);
  run_params(5, 5, 8192, 55, true);
  if (!quick) {
    run_params(5, 5, 16384, 55, true);
    run_params(30, 30, 4096, 48, true, true);
    run_params(30, 30, 4096, 52, true);
    run_params(20, 0, 32768, 48, false, true);
    run_params(30, 0, 65536, 48, false, true);
    reference_mode_prec48();
  }
  std::printf("failures=%d\n", failures);
  return failures;
}
