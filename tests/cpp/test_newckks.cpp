// Functional tests of the newckks-compatible C++ API on the GPU, written in
// the style of the reference's own tests (aoa-ntt/tests/tests.cpp:862-1760):
// random slot values, encrypt -> op -> decrypt, compared with the FP64 result
// within the reference's bounds (ERRBOUND 0.005, tests.cpp:34; rotations 0.01,
// tests.cpp:1578, :1635).  Prints one line per case; exit status = failures.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include <newckks/newckks.h>

typedef std::complex<double> cd;
static const double ERRBOUND = 0.005;
static int failures = 0;
static bool info_only = false;  // report without checking (reference-mode runs)

static double get_range(uint64_t prec) {  // tests.cpp:36-38
  return (double)(CUDAManager::RNSQPrimes[0] - 1) / (double)((uint64_t)1 << (prec + 1));
}

static void report(const std::string& name, double err, double bound) {
  bool ok = err <= bound && std::isfinite(err);
  if (info_only) {
    std::printf("INFO %s_referencemoddown max_err=%.3e (not checked)\n", name.c_str(), err);
    std::fflush(stdout);
    return;
  }
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
  // Serialisation: the reference's text and JSON layouts and the binary form
  // round-trip bit for bit (poly_t.cpp:445-473, cipher_t.cpp:42-82)
  {
    for (auto& x : v1) x = cd(dist(gen), dist(gen));
    cipher_t* a = ckks_encrypt(ctx, v1.data(), slots);
    poly_t* p = poly_import(ctx, poly_export(ctx, &a->c[0]));
    bool same = poly_are_equal(ctx, p, &a->c[0]);
    poly_free(ctx, p);
    delete p;
    cipher_t* j = cipher_import_json(ctx, cipher_export_json(ctx, a));
    same = same && poly_are_equal(ctx, &j->c[0], &a->c[0]) && poly_are_equal(ctx, &j->c[1], &a->c[1]) &&
           j->level == a->level && j->scale == a->scale && j->slots == a->slots;
    std::vector<uint8_t> bin = cipher_export_binary(ctx, a);
    cipher_t* b = cipher_import_binary(ctx, bin.data(), bin.size());
    same = same && poly_are_equal(ctx, &b->c[0], &a->c[0]) && poly_are_equal(ctx, &b->c[1], &a->c[1]);
    cd* out = ckks_decrypt(ctx, b, sk);
    report("Serialisation_" + tag, same ? maxerr(v1.data(), out, slots) : 1.0, ERRBOUND);
    free(out);
    for (cipher_t* x : {a, j, b}) {
      cipher_free(ctx, x);
      delete x;
    }
  }
  // Keys through export_keys / load_keys (ckkscontext.cpp:50-100) and the
  // binary form: a second context with the loaded keys encrypts, multiplies
  // and decrypts within the bound
  if (heavy) {
    ctx->keys.sk = sk;
    const std::string doc = ctx->export_keys().dump();
    const std::vector<uint8_t> kb = ctx->export_keys_binary();
    for (int form = 0; form < 2; ++form) {
      CKKSContext* c2 = new CKKSContext();
      if (form == 0)
        c2->load_keys(json::parse(doc));
      else
        c2->load_keys_binary(kb.data(), kb.size());
      bool same = poly_are_equal(c2, &c2->keys.evk->b, &ctx->keys.evk->b) &&
                  poly_are_equal(c2, &c2->keys.sk->s, &sk->s) && poly_are_equal(c2, &c2->keys.pk->a, &ctx->keys.pk->a);
      for (int i = 0; i < slots; ++i) {
        v1[i] = cd(dsq(gen), dsq(gen));
        v2[i] = cd(dsq(gen), dsq(gen));
        exp[i] = v1[i] * v2[i];
      }
      cipher_t* a = ckks_encrypt(c2, v1.data(), slots);
      cipher_t* b = ckks_encrypt(c2, v2.data(), slots);
      cipher_t* m = ckks_mul(c2, a, b);
      cd* out = ckks_decrypt(c2, m, c2->keys.sk);
      report(std::string(form ? "KeysBinary_" : "KeysJson_") + tag, same ? maxerr(exp.data(), out, slots) : 1.0,
             ERRBOUND);
      free(out);
      for (cipher_t* x : {a, b, m}) {
        cipher_free(c2, x);
        delete x;
      }
      SecretKey* s2 = c2->keys.sk;
      ckkskeychain_free(c2, &c2->keys);
      ckks_free_sk(c2, s2);
      delete s2;
      delete c2;
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

// The reference's own ModDown reduction at prec 48 and 52 (special primes
// 55-bit, wider than q_i): reported, not checked.  A cross-base product whose
// remainder falls in the Barrett deviation window is reduced to r - eps, and
// once such a product reaches a kept limb the coefficient decrypts to
// garbage; the bit-exact default reproduces that.  At prec 48 the window
// covers ~1e-4 of the remainders and one HomMul fails; at prec 52 ~2.5e-7,
// and with the draws of the prec-52 parameter set the 29-step MulSequence
// meets it.  The same runs in exact mode are checked by run_params.
static void reference_mode(int prec, int depth) {
  CUDAManager::set_exact_moddown(false);
  CUDAManager::init(30, 30, 4096, prec);
  CKKSContext* ctx = new CKKSContext();
  SecretKey* sk = ckks_new_sk(ctx);
  CKKSKeychain keys;
  keys.pk = ckks_new_pk(ctx, sk);
  keys.evk = ckks_new_mtk(ctx, sk);
  ctx->keys = keys;
  const int slots = 2048;
  std::vector<cd> v(slots, cd(1.0, 0.0));
  cipher_t* a = ckks_encrypt(ctx, v.data(), slots);
  cipher_t* b = ckks_encrypt(ctx, v.data(), slots);
  double worst = 0;
  for (int d = 0; d < depth; ++d) {
    ckks_mul(ctx, b, b, a);
    cd* out = ckks_decrypt(ctx, b, sk);
    worst = std::max(worst, maxerr(v.data(), out, slots));
    free(out);
  }
  std::printf("INFO ReferenceModDown_MulSequence%d_N4096_k30_kl30_prec%d max_err=%.3e (not checked)\n", depth, prec,
              worst);
  for (cipher_t* x : {a, b}) {
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

// Plaintext / constant operands, conjugation and the compositions
// (tests.cpp:1224-1535, :1651-1700, :1762-2830) on one parameter set with
// enough levels (the reference's advanced tests use k=30, N=4096).
static void run_adv(int k, int kl, int N, int prec) {
  CUDAManager::init(k, kl, N, prec);
  CKKSContext* ctx = new CKKSContext();
  SecretKey* sk = ckks_new_sk(ctx);
  CKKSKeychain* keys = new CKKSKeychain;
  keys->pk = ckks_new_pk(ctx, sk);
  keys->evk = ckks_new_mtk(ctx, sk);
  std::vector<int> steps;
  for (int r = 1; r < (N >> 1); r *= 2) steps.push_back(r);
  keys->rtk_left = ckks_new_rtk_left(ctx, sk, steps);
  keys->rtk_right = ckks_new_rtk_right(ctx, sk);
  keys->cjk = ckks_new_cjk(ctx, sk);
  ctx->keys = *keys;
  const int slots = N >> 1;
  const std::string tag = "N" + std::to_string(N) + "_k" + std::to_string(k) + "_prec" + std::to_string(prec);
  std::default_random_engine gen;
  const double R = get_range(prec);
  std::uniform_real_distribution<double> dh(-R / 2, R / 2), dsq(-std::sqrt(R / 2), std::sqrt(R / 2)), d01(0, 1),
      dm1(-1, 1);
  std::vector<cd> v1(slots), v2(slots), exp(slots);
  auto done = [&](std::initializer_list<cipher_t*> cs) {
    for (cipher_t* x : cs) {
      cipher_free(ctx, x);
      delete x;
    }
  };
  // AddPlaintextDouble (tests.cpp:1224-1282)
  {
    for (auto& x : v1) x = cd(dh(gen), dh(gen));
    const double c = dh(gen);
    cipher_t* a = ckks_encrypt(ctx, v1.data(), slots);
    cipher_t* b = ckks_add(ctx, a, c);
    cd* out = ckks_decrypt(ctx, b, sk);
    for (int i = 0; i < slots; ++i) exp[i] = v1[i] + c;
    report("AddPlaintextDouble_" + tag, maxerr(exp.data(), out, slots), ERRBOUND);
    free(out);
    done({a, b});
  }
  // MulByConstantDouble, both signs, and DivByConstant (tests.cpp:1284-1344, :1475-1535)
  {
    for (auto& x : v1) x = cd(dsq(gen), dsq(gen));
    double worst = 0;
    for (double c : {std::abs(dsq(gen)), -std::abs(dsq(gen))}) {
      cipher_t* a = ckks_encrypt(ctx, v1.data(), slots);
      cipher_t* b = ckks_mul(ctx, a, c);
      cd* out = ckks_decrypt(ctx, b, sk);
      for (int i = 0; i < slots; ++i) exp[i] = v1[i] * c;
      worst = std::max(worst, maxerr(exp.data(), out, slots));
      free(out);
      done({a, b});
    }
    report("MulByConstantDouble_" + tag, worst, ERRBOUND);
    double c = dsq(gen);
    c += 0.1 * (std::abs(1.0 / c) >= R / 2);
    for (auto& x : v1) x = cd(dsq(gen), 0);
    cipher_t* a = ckks_encrypt(ctx, v1.data(), slots);
    cipher_t* b = ckks_mul(ctx, a, 1.0 / c);
    cd* out = ckks_decrypt(ctx, b, sk);
    for (int i = 0; i < slots; ++i) exp[i] = v1[i] / c;
    report("DivByConstant_" + tag, maxerr(exp.data(), out, slots), ERRBOUND);
    free(out);
    done({a, b});
  }
  // MulByConstantArray and MulByConstantPoly (tests.cpp:1346-1473)
  {
    for (int i = 0; i < slots; ++i) {
      v1[i] = cd(dsq(gen), dsq(gen));
      v2[i] = cd(dsq(gen), dsq(gen));
      exp[i] = v1[i] * v2[i];
    }
    cipher_t* a = ckks_encrypt(ctx, v1.data(), slots);
    cipher_t* b = ckks_mul(ctx, a, v2.data());
    cd* out = ckks_decrypt(ctx, b, sk);
    report("MulByConstantArray_" + tag, maxerr(exp.data(), out, slots), ERRBOUND);
    free(out);
    poly_t* p = new poly_t;
    poly_init(ctx, p);
    int64_t scale = -1;
    ctx->encode(p, &scale, v2.data(), slots, 0);
    cipher_t* a2 = ckks_encrypt(ctx, v1.data(), slots);
    cipher_t* c = ckks_mul(ctx, a2, p);
    out = ckks_decrypt(ctx, c, sk);
    report("MulByConstantPoly_" + tag, maxerr(exp.data(), out, slots), ERRBOUND);
    free(out);
    poly_free(ctx, p);
    delete p;
    done({a, b, a2, c});
  }
  // RotateRight (tests.cpp:1537-1592) and Conjugate (tests.cpp:1651-1700)
  {
    for (auto& x : v1) x = cd(d01(gen), d01(gen));
    cipher_t* a = ckks_encrypt(ctx, v1.data(), slots);
    double worst = 0;
    for (int r : {1, 8, 256}) {
      if (r >= slots) continue;
      cipher_t c;
      cipher_init(ctx, &c);
      ckks_rotate_right(ctx, &c, a, r);
      cd* out = ckks_decrypt(ctx, &c, sk);
      for (int i = 0; i < slots; ++i) exp[i] = v1[(i - r + slots) % slots];
      worst = std::max(worst, maxerr(exp.data(), out, slots));
      free(out);
      cipher_free(ctx, &c);
    }
    report("RotateRight_" + tag, worst, 0.01);
    cipher_t* b = ckks_conjugate(ctx, a);
    cd* out = ckks_decrypt(ctx, b, sk);
    for (int i = 0; i < slots; ++i) exp[i] = std::conj(v1[i]);
    report("Conjugate_" + tag, maxerr(exp.data(), out, slots), ERRBOUND);
    free(out);
    done({a, b});
  }
  // InnerProductSingle (tests.cpp:1762-1831): length-4 vectors of 1-slot ciphertexts
  {
    const int length = 4;
    double worst = 0;
    for (int t = 0; t < 3; ++t) {
      cd va[length], vb[length], ip = 0;
      do {
        ip = 0;
        for (int i = 0; i < length; ++i) {
          va[i] = cd(dsq(gen), dsq(gen));
          vb[i] = cd(dsq(gen), dsq(gen));
          ip += va[i] * vb[i];
        }
      } while (std::abs(ip.real()) >= R || std::abs(ip.imag()) >= R);
      cipher_t* ca = new cipher_t[length];
      cipher_t* cb = new cipher_t[length];
      for (int i = 0; i < length; ++i) {
        cipher_init(ctx, &ca[i]);
        cipher_init(ctx, &cb[i]);
        ckks_encrypt(ctx, &ca[i], &va[i]);
        ckks_encrypt(ctx, &cb[i], &vb[i]);
      }
      cipher_t c3;
      ckks_inner_prod(ctx, &c3, ca, cb, length);
      cd* out = ckks_decrypt(ctx, &c3, sk);
      worst = std::max(worst, std::max(std::abs(out[0].real() - ip.real()), std::abs(out[0].imag() - ip.imag())));
      free(out);
      for (int i = 0; i < length; ++i) {
        cipher_free(ctx, &ca[i]);
        cipher_free(ctx, &cb[i]);
      }
      cipher_free(ctx, &c3);
      delete[] ca;
      delete[] cb;
    }
    report("InnerProductSingle_" + tag, worst, 10 * ERRBOUND);
  }
  // InnerProductBatch (tests.cpp:1833-1908): ckks_batch_inner_prod over 16 slots
  {
    const int ns = 16;
    std::vector<cd> a1(ns), a2(ns);
    cd ip = 0;
    for (int i = 0; i < ns; ++i) {
      a1[i] = cd(dm1(gen), dm1(gen));
      a2[i] = cd(dm1(gen), dm1(gen));
      ip += a1[i] * a2[i];
    }
    cipher_t* x = ckks_encrypt(ctx, a1.data(), ns);
    cipher_t* y = ckks_encrypt(ctx, a2.data(), ns);
    cipher_t c3;
    cipher_init(ctx, &c3);
    ckks_batch_inner_prod(ctx, &c3, x, y);
    cd* out = ckks_decrypt(ctx, &c3, sk);
    report("InnerProductBatch_" + tag, std::max(std::abs(out[0].real() - ip.real()), std::abs(out[0].imag() - ip.imag())),
           10 * ERRBOUND);
    free(out);
    cipher_free(ctx, &c3);
    done({x, y});
  }
  // Power (tests.cpp:2052-2102)
  {
    const int bound = 5;
    std::uniform_real_distribution<double> dp(-std::pow(R, 1.0 / (bound + 1)), std::pow(R, 1.0 / (bound + 1)));
    double worst = 0;
    for (int x : {0, 1, 2, 3, 4}) {
      for (auto& v : v1) v = cd(dp(gen), dp(gen));
      cipher_t* a = ckks_encrypt(ctx, v1.data(), slots);
      cipher_t* b = ckks_power(ctx, a, (uint64_t)x);
      cd* out = ckks_decrypt(ctx, b, sk);
      for (int i = 0; i < slots; ++i) exp[i] = std::pow(v1[i], x);
      worst = std::max(worst, maxerr(exp.data(), out, slots));
      free(out);
      done({a, b});
    }
    report("Power_" + tag, worst, ERRBOUND);
  }
  // evalpoly with the sine series (tests.cpp:2153-2204)
  {
    const int degree = 11;
    for (auto& v : v1) v = cd(dm1(gen), 0);
    double* coeffs = ctx->maclaurin_coeffs[SIN];
    cipher_t* a = ckks_encrypt(ctx, v1.data(), slots);
    cipher_t* b = ckks_eval_polynomial(ctx, a, coeffs, degree);
    cd* out = ckks_decrypt(ctx, b, sk);
    for (int j = 0; j < slots; ++j) {
      cd e = 0;
      for (int i = degree - 1; i >= 0; --i) e = coeffs[i] + v1[j] * e;
      exp[j] = e;
    }
    report("EvalPoly_" + tag, maxerr(exp.data(), out, slots), ERRBOUND);
    free(out);
    done({a, b});
  }
  // exp, sin, cos, sigmoid, log, log1minus, inverse (tests.cpp:2206-2782):
  // real inputs over the reference's ranges, with its bounds.  (Its cosine
  // series has the x^3 term where x^4 belongs, so it is only accurate on
  // [0, 1], the range its test uses.)
  if (prec > 52) {
    struct Fn {
      const char* name;
      cipher_t* (*f)(CKKSContext*, cipher_t*);
      double (*ref)(double);
      double lo, hi, bound;
    };
    const Fn fns[] = {
        {"Exp", ckks_exp, [](double x) { return std::exp(x); }, 0, 1, 2 * ERRBOUND},
        {"Sin", ckks_sin, [](double x) { return std::sin(x); }, 0, 1, ERRBOUND},
        {"Cos", ckks_cos, [](double x) { return std::cos(x); }, 0, 1, ERRBOUND},
        {"Sigmoid", ckks_sigmoid, [](double x) { return 1.0 / (1.0 + std::exp(-x)); }, 0, 1, ERRBOUND},
        {"Log", [](CKKSContext* c, cipher_t* x) { return ckks_log(c, x); }, [](double x) { return std::log(x); },
         0.1, 1, 0.5},
        // the 10-step series diverges from log(1-x) beyond 0.5 near x = 1,
        // so the range stops at 0.9
        {"Log1Minus", ckks_log1minus, [](double x) { return std::log(1 - x); }, 0, 0.9, 0.5},
        {"Inverse", [](CKKSContext* c, cipher_t* x) { return ckks_inverse(c, x); },
         [](double x) { return 1.0 / x; }, 0.1, 1, 10},
    };
    for (const Fn& fn : fns) {
      std::uniform_real_distribution<double> dd(fn.lo, fn.hi);
      for (auto& v : v1) v = cd(dd(gen), 0);
      cipher_t* a = ckks_encrypt(ctx, v1.data(), slots);
      cipher_t* b = fn.f(ctx, a);
      cd* out = ckks_decrypt(ctx, b, sk);
      double worst = 0;
      for (int i = 0; i < slots; ++i) worst = std::max(worst, std::abs(out[i].real() - fn.ref(v1[i].real())));
      report(std::string(fn.name) + "_" + tag, worst, fn.bound);
      free(out);
      done({a, b});
    }
  }
  // DiscardSlots (tests.cpp:2784-2829): values in [0.1, 10], bound 1e-3
  {
    std::uniform_real_distribution<double> dv(0.1, 10);
    for (auto& v : v1) v = cd(dv(gen), 0);
    cipher_t* a = ckks_encrypt(ctx, v1.data(), slots);
    const int idx = 3;
    ckks_discard_slots_except(ctx, a, idx);
    cd* out = ckks_decrypt(ctx, a, sk);
    double worst = 0;
    for (int i = 0; i < slots; ++i) worst = std::max(worst, std::abs(out[i].real() - (i == idx ? v1[i].real() : 0.0)));
    report("DiscardSlots_" + tag, worst, 1e-3);
    free(out);
    done({a});
  }
  ckkskeychain_free(ctx, keys);
  delete keys;
  ckks_free_sk(ctx, sk);
  delete sk;
  delete ctx;
  CUDAManager::destroy();
}

// CUDAManager's device entry points on raw residue pointers
// (aoa-ntt/include/newckks/cuda/manager.h:128-166) against the poly_t layer
// that the reference builds on them, and the prime helpers.
static void manager_statics() {
  const int k = 4, kl = 5, N = 4096, prec = 52;
  CUDAManager::init(k, kl, N, prec);
  Context* ctx = new Context();
  const int R = k + kl;
  std::mt19937_64 gen(11);
  auto primes = [&]() {
    std::vector<uint64_t> p(CUDAManager::RNSQPrimes);
    p.insert(p.end(), CUDAManager::RNSBPrimes.begin(), CUDAManager::RNSBPrimes.end());
    return p;
  }();
  auto random_poly = [&](poly_t* p, poly_bases b) {
    poly_init(ctx, p, b);
    const int n = CUDAManager::get_n_residues(b);
    std::vector<uint64_t> h((size_t)n * N);
    for (int r = 0; r < n; ++r)
      for (int i = 0; i < N; ++i) h[(size_t)r * N + i] = gen() % primes[r];
    poly_copy_to_device(ctx, p, h.data());
  };
  auto host = [&](poly_t* p) {
    uint64_t* h = poly_copy_to_host(ctx, p);
    std::vector<uint64_t> v(h, h + (size_t)CUDAManager::get_n_residues(p->base) * N);
    free(h);
    return v;
  };
  auto same = [&](const std::vector<uint64_t>& a, const std::vector<uint64_t>& b, int r0, int r1) {
    return std::equal(a.begin() + (size_t)r0 * N, a.begin() + (size_t)r1 * N, b.begin() + (size_t)r0 * N);
  };
  const int level = k - 2;
  // execute_modup(ctx, out QB, in Q, level) with separate buffers == poly_modup
  {
    poly_t q, ref, raw;
    random_poly(&q, QBase);
    poly_modup(ctx, &ref, &q, level);
    poly_init(ctx, &raw, QBBase);
    CUDAManager::execute_modup(ctx, raw.d_coefs, q.d_coefs, level);
    report("ManagerExecuteModUp", same(host(&ref), host(&raw), k, R) ? 0.0 : 1.0, 0.0);
    for (poly_t* p : {&q, &ref, &raw}) poly_free(ctx, p);
  }
  // execute_moddown(ctx, out Q, in QB, level) == poly_moddown
  {
    poly_t qb, ref, raw;
    random_poly(&qb, QBBase);
    poly_moddown(ctx, &ref, &qb, level);
    poly_init(ctx, &raw, QBase);
    CUDAManager::execute_moddown(ctx, raw.d_coefs, qb.d_coefs, level);
    report("ManagerExecuteModDown", same(host(&ref), host(&raw), 0, level + 1) ? 0.0 : 1.0, 0.0);
    for (poly_t* p : {&qb, &ref, &raw}) poly_free(ctx, p);
  }
  // execute_rescale and execute_rho == poly_rescale, poly_rho
  {
    poly_t a, b, a2, b2, x, rho1, rho2;
    random_poly(&a, QBase);
    random_poly(&b, QBase);
    poly_copy(ctx, &a2, &a);
    poly_copy(ctx, &b2, &b);
    poly_rescale(ctx, &a, &b, level);
    CUDAManager::execute_rescale(ctx, a2.d_coefs, b2.d_coefs, level);
    bool ok = same(host(&a), host(&a2), 0, level) && same(host(&b), host(&b2), 0, level);
    random_poly(&x, QBBase);
    poly_rho(ctx, &rho1, &x);
    poly_init(ctx, &rho2, QBBase);
    CUDAManager::execute_rho(ctx, rho2.d_coefs, x.d_coefs);
    ok = ok && same(host(&rho1), host(&rho2), 0, R);
    report("ManagerExecuteRescaleRho", ok ? 0.0 : 1.0, 0.0);
    for (poly_t* p : {&a, &b, &a2, &b2, &x, &rho1, &rho2}) poly_free(ctx, p);
  }
  // execute_dot with one term == poly_mul; poly_dr2 == its poly_mul composition
  {
    poly_t a0, a1, b0, b1, c, m, m2, m3, t21, t22, t23;
    for (poly_t* p : {&a0, &a1, &b0, &b1}) random_poly(p, QBBase);
    poly_init(ctx, &c, QBBase);
    uint64_t* pa[1] = {a0.d_coefs};
    uint64_t* pb[1] = {b0.d_coefs};
    uint64_t** da = nullptr;
    uint64_t** db = nullptr;
    (void)hipMalloc((void**)&da, sizeof(pa));
    (void)hipMalloc((void**)&db, sizeof(pb));
    (void)hipMemcpy(da, pa, sizeof(pa), hipMemcpyHostToDevice);
    (void)hipMemcpy(db, pb, sizeof(pb), hipMemcpyHostToDevice);
    CUDAManager::execute_dot(ctx, c.d_coefs, da, db, 1);
    c.state = RNSSTATE;  // the product of raw residues, read back as stored
    poly_init(ctx, &m, QBBase);
    (void)hipStreamSynchronize(ctx->get_stream());
    std::vector<uint64_t> hc((size_t)R * N), ha(hc.size()), hb(hc.size());
    (void)hipMemcpy(hc.data(), c.d_coefs, hc.size() * 8, hipMemcpyDeviceToHost);
    (void)hipMemcpy(ha.data(), a0.d_coefs, ha.size() * 8, hipMemcpyDeviceToHost);
    (void)hipMemcpy(hb.data(), b0.d_coefs, hb.size() * 8, hipMemcpyDeviceToHost);
    bool ok = true;
    for (size_t t = 0; t < hc.size(); ++t)
      ok = ok && hc[t] == (uint64_t)((unsigned __int128)ha[t] * hb[t] % primes[t / N]);
    (void)hipFree(da);
    (void)hipFree(db);
    poly_dr2(ctx, &t21, &t22, &t23, &a0, &a1, &b0, &b1);
    poly_mul(ctx, &m, &a0, &b0);
    ok = ok && same(host(&t21), host(&m), 0, R);
    poly_mul(ctx, &m2, &a0, &b1);
    poly_mul(ctx, &m3, &a1, &b0);
    poly_mul(ctx, &m, &m2, &m3);
    ok = ok && same(host(&t22), host(&m), 0, R);
    poly_mul(ctx, &m, &a1, &b1);
    ok = ok && same(host(&t23), host(&m), 0, R);
    report("ManagerExecuteDotDr2", ok ? 0.0 : 1.0, 0.0);
    for (poly_t* p : {&a0, &a1, &b0, &b1, &c, &m, &m2, &m3, &t21, &t22, &t23}) poly_free(ctx, p);
  }
  // gen_rns_primes keeps the initialised bases; the products are the primes'
  {
    const std::vector<uint64_t> q0 = CUDAManager::RNSQPrimes, b0 = CUDAManager::RNSBPrimes;
    CUDAManager::gen_rns_primes(k, kl);
    bool ok = CUDAManager::RNSQPrimes == q0 && CUDAManager::RNSBPrimes == b0;
    CUDAManager::precompute();
    for (int which = 0; which < 2; ++which) {
      const std::vector<uint64_t> w = which ? CUDAManager::get_b_product() : CUDAManager::get_q_product();
      const std::vector<uint64_t>& ps = which ? b0 : q0;
      int bits = 0;
      for (uint64_t p : ps) bits += 64 - __builtin_clzll(p);
      const int wbits = 64 * ((int)w.size() - 1) + (64 - __builtin_clzll(w.back()));
      ok = ok && wbits <= bits && wbits > bits - (int)ps.size();
      for (uint64_t p : ps) {  // every prime divides the product
        unsigned __int128 r = 0;
        for (size_t i = w.size(); i-- > 0;) r = ((r << 64) | w[i]) % p;
        ok = ok && r == 0;
      }
    }
    report("ManagerPrimesAndProducts", ok ? 0.0 : 1.0, 0.0);
  }
  delete ctx;
  CUDAManager::destroy();
}

static std::string read_file(const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  std::string s;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) s.append(buf, n);
  std::fclose(f);
  return s;
}

static void write_file(const std::string& path, const void* data, size_t n) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(data, 1, n, f) != n) throw std::runtime_error("cannot write " + path);
  std::fclose(f);
}

static void write_poly(FILE* f, Context* ctx, poly_t* p) {
  uint64_t* h = poly_copy_to_host(ctx, p);
  std::fwrite(h, 1, poly_get_size(p->base), f);
  free(h);
}

// Reads a ciphertext document and a key document in the reference's layouts
// (N=2048, k=2, kl=1, prec 55), writes their residues (coefficient domain,
// uint64) and properties as imported, their exports as compact text, and
// checks the binary forms round-trip.  Returns 0 on success.
static int serial_mode(const char* ct_path, const char* keys_path, const std::string& out) {
  CUDAManager::init(2, 1, 2048, 55);
  CKKSContext* ctx = new CKKSContext();
  cipher_t* ct = cipher_import(ctx, json::parse(read_file(ct_path)));
  FILE* f = std::fopen((out + "/ct.bin").c_str(), "wb");
  write_poly(f, ctx, &ct->c[0]);
  write_poly(f, ctx, &ct->c[1]);
  std::fclose(f);
  const std::string props = std::to_string(ct->level) + " " + std::to_string(ct->scale) + " " + std::to_string(ct->slots);
  write_file(out + "/ct_props.txt", props.data(), props.size());
  const std::string ce = cipher_export(ctx, ct).dump();
  write_file(out + "/ct_export.json", ce.data(), ce.size());
  ctx->load_keys(json::parse(read_file(keys_path)));
  f = std::fopen((out + "/keys.bin").c_str(), "wb");
  for (poly_t* p : {&ctx->keys.sk->s, &ctx->keys.pk->a, &ctx->keys.pk->b, &ctx->keys.evk->a, &ctx->keys.evk->b})
    write_poly(f, ctx, p);
  std::fclose(f);
  const std::string ke = ctx->export_keys().dump();
  write_file(out + "/keys_export.json", ke.data(), ke.size());
  // binary forms
  std::vector<uint8_t> cb = cipher_export_binary(ctx, ct);
  cipher_t* c2 = cipher_import_binary(ctx, cb.data(), cb.size());
  bool ok = poly_are_equal(ctx, &c2->c[0], &ct->c[0]) && poly_are_equal(ctx, &c2->c[1], &ct->c[1]) &&
            c2->level == ct->level && c2->scale == ct->scale && c2->slots == ct->slots;
  std::vector<uint8_t> kb = ctx->export_keys_binary();
  CKKSContext* k2 = new CKKSContext();
  k2->load_keys_binary(kb.data(), kb.size());
  ok = ok && poly_are_equal(k2, &k2->keys.sk->s, &ctx->keys.sk->s) && poly_are_equal(k2, &k2->keys.evk->a, &ctx->keys.evk->a) &&
       poly_are_equal(k2, &k2->keys.evk->b, &ctx->keys.evk->b) && poly_are_equal(k2, &k2->keys.pk->b, &ctx->keys.pk->b);
  // a truncated or foreign document is refused
  bool refused = false;
  try {
    cipher_import(ctx, json::parse("{\"ct\":{\"0\":\"1,2,\",\"1\":\"3,\"},\"props\":{\"level\":1,\"scale\":2,\"slots\":4}}"));
  } catch (const std::runtime_error&) {
    refused = true;
  }
  std::printf("%s SerialBinaryRoundTrip\n%s SerialShortDocumentRefused\n", ok ? "PASS" : "FAIL", refused ? "PASS" : "FAIL");
  return ok && refused ? 0 : 1;
}

int main(int argc, char** argv) {
  if (argc == 5 && std::string(argv[1]) == "serial") return serial_mode(argv[2], argv[3], argv[4]);
  const bool quick = argc > 1 && std::string(argv[1]) == "quick";
  manager_statics();
  mul_n4096_k4();
  run_params(2, 2, 2048, 55, true);
  run_params(3, 3, 4096, 55, true);
  run_params(5, 5, 8192, 55, true);
  run_adv(30, 30, 4096, 55);
  run_adv(20, 20, 2048, 55);
  if (!quick) {
    run_params(5, 5, 16384, 55, true);
    run_params(30, 30, 4096, 48, true, true);
    run_params(30, 30, 4096, 52, true, true);
    run_params(20, 0, 32768, 48, false, true);
    run_params(30, 0, 65536, 48, false, true);
    reference_mode(48, 1);
    info_only = true;  // the prec-52 set again, with the reference's reduction
    run_params(30, 30, 4096, 52, true);
    info_only = false;
  }
  std::printf("failures=%d\n", failures);
  return failures;
}
