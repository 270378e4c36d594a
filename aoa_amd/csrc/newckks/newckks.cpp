// newckks-compatible C++ API over the C-ABI of the gfx950 core (include/aoa.h).
//
// Mirrors the reference's host layers: CUDAManager (aoa-ntt/src/cuda/manager.cu:762-821),
// Context (aoa-ntt/src/tool/context.cpp), the poly_t state machine
// (aoa-ntt/src/arithmetic/poly_t.cpp), the sampler (aoa-ntt/src/cuda/sampler.cu),
// keys (aoa-ntt/src/ckks/ckkskeys.cpp), the encoder (aoa-ntt/src/ckks/encoder.cu),
// and the scheme (aoa-ntt/src/ckks/ckks.cpp).  Every residue computation runs
// in libaoa.so; this file only sequences calls, moves host data and keeps
// the bookkeeping (levels, scales, domains).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/newckks/newckks.h"

// ------------------------------------------------------------ helpers ------

static void ok(int rc) {
  if (rc != AOA_OK) throw std::runtime_error(aoa_last_error());
}
static void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------ CUDAManager --

int CUDAManager::is_init = 0;
int CUDAManager::N = 0;
int CUDAManager::dnum = 1;
uint64_t CUDAManager::scalingfactor = 0;
std::vector<uint64_t> CUDAManager::RNSQPrimes;
std::vector<uint64_t> CUDAManager::RNSBPrimes;
bool CUDAManager::exact_moddown = false;
static aoa_params* g_params = nullptr;

void CUDAManager::set_exact_moddown(bool exact) {
  exact_moddown = exact;
  if (g_params) ok(aoa_params_set_moddown_exact(g_params, exact ? 1 : 0));
}

aoa_params* CUDAManager::params() {
  if (!g_params) throw std::runtime_error("CUDAManager is not initialised");
  return g_params;
}

void CUDAManager::init(const int k, const int kl, const int N_, const uint64_t prec, engine_types e) {
  if (e != NTTTrans) throw std::runtime_error("Unknown engine mode");
  destroy();
  aoa_params* p = nullptr;
  int rc = aoa_params_create(N_, k, kl, (int)prec, &p);
  if (rc != AOA_OK) throw std::runtime_error(aoa_last_error());
  g_params = p;
  ok(aoa_params_set_moddown_exact(p, exact_moddown ? 1 : 0));
  int n, kk, kkl, pr;
  ok(aoa_params_info(p, &n, &kk, &kkl, &pr));
  std::vector<uint64_t> all(kk + kkl);
  ok(aoa_params_primes(p, all.data(), kk + kkl));
  RNSQPrimes.assign(all.begin(), all.begin() + kk);
  RNSBPrimes.assign(all.begin() + kk, all.end());
  N = n;
  scalingfactor = prec;
  dnum = 1;
  is_init = 1;
}

void CUDAManager::init(const int k, const int N_, const uint64_t prec, engine_types e) { init(k, k + 1, N_, prec, e); }

void CUDAManager::destroy() {
  if (!is_init) return;
  (void)hipDeviceSynchronize();
  aoa_params_destroy(g_params);
  g_params = nullptr;
  RNSQPrimes.clear();
  RNSBPrimes.clear();
  is_init = 0;
}

void CUDAManager::gen_rns_primes(unsigned int k, unsigned int kl) {  // manager.cu:483-564
  if (is_init && RNSQPrimes.size() == k && RNSBPrimes.size() == kl) return;
  std::vector<uint64_t> all(k + kl);
  ok(aoa_rns_primes((int)k, (int)kl, (int)scalingfactor, all.data(), (int)all.size()));
  RNSQPrimes.assign(all.begin(), all.begin() + k);
  RNSBPrimes.assign(all.begin() + k, all.end());
}

void CUDAManager::precompute() {  // manager.cu:566-759: done by aoa_params_create
  if (!is_init) throw std::runtime_error("CUDAManager is not initialised");
}

// Product of the primes as little-endian 64-bit words.
static std::vector<uint64_t> prime_product(const std::vector<uint64_t>& primes) {
  std::vector<uint64_t> w(1, 1);
  for (uint64_t p : primes) {
    unsigned __int128 carry = 0;
    for (auto& x : w) {
      const unsigned __int128 t = (unsigned __int128)x * p + carry;
      x = (uint64_t)t;
      carry = t >> 64;
    }
    if (carry) w.push_back((uint64_t)carry);
  }
  return w;
}
std::vector<uint64_t> CUDAManager::get_q_product() { return prime_product(RNSQPrimes); }
std::vector<uint64_t> CUDAManager::get_b_product() { return prime_product(RNSBPrimes); }

void CUDAManager::execute_dot(Context* ctx, uint64_t* c, uint64_t** a, uint64_t** b, const int k) {
  ok(aoa_dot(params(), ctx->get_stream(), c, a, b, k));
}
void CUDAManager::execute_rho(Context* ctx, uint64_t* b, uint64_t* a) {
  ok(aoa_rho(params(), ctx->get_stream(), b, a));
}
void CUDAManager::execute_rescale(Context* ctx, uint64_t* a, uint64_t* b, const int level) {
  ok(aoa_rescale(params(), ctx->get_stream(), a, b, level));
}
void CUDAManager::execute_modup(Context* ctx, uint64_t* a, uint64_t* b, const int level) {
  ok(aoa_modup_from(params(), ctx->get_stream(), a, b, level));
}
void CUDAManager::execute_moddown(Context* ctx, uint64_t* a, uint64_t* b, const int level) {
  ok(aoa_moddown(params(), ctx->get_stream(), a, b, level));
}
void CUDAManager::execute_dr2(Context* ctx, uint64_t* ct21, uint64_t* ct22, uint64_t* ct23, const uint64_t* ct01,
                              const uint64_t* ct02, const uint64_t* ct11, const uint64_t* ct12) {
  ok(aoa_dr2(params(), ctx->get_stream(), ct21, ct22, ct23, ct01, ct02, ct11, ct12));
}

size_t poly_get_size(poly_bases b) { return (size_t)CUDAManager::N * CUDAManager::get_n_residues(b) * sizeof(uint64_t); }

// ------------------------------------------------------------ Context ------

Context::Context() {
  hip_ok(hipStreamCreate(&stream), "hipStreamCreate");
  d_tmp = nullptr;
  if (CUDAManager::is_init) hip_ok(hipMalloc((void**)&d_tmp, poly_get_size(QBBase)), "hipMalloc");
  s = new Sampler(this);
}

Context::~Context() {
  (void)hipStreamSynchronize(stream);
  if (d_tmp) (void)hipFree(d_tmp);
  delete s;
  (void)hipStreamDestroy(stream);
}

// ------------------------------------------------------------ poly layer ---

// Polynomials come from the device's stream-ordered memory pool where the
// device has one (poly_t.cpp:10-31 takes cudaMallocAsync when supported), so
// allocating and freeing the temporaries of every ckks_* call neither blocks
// the host nor synchronises the stream.  The pool keeps freed memory for
// reuse instead of returning it at each synchronisation.
static bool stream_alloc() {
  static int supported = -1;
  if (supported < 0) {
    int dev = 0, attr = 0;
    supported = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&attr, hipDeviceAttributeMemoryPoolsSupported, dev) == hipSuccess && attr) {
      hipMemPool_t pool;
      if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        supported = 1;
      }
    }
  }
  return supported == 1;
}

void poly_init(Context* ctx, poly_t* a, poly_bases b) {
  if (a->state != NONINITIALIZED) return;
  if (!CUDAManager::is_init) throw std::runtime_error("CUDAManager is not initialised");
  if (stream_alloc())
    hip_ok(hipMallocAsync((void**)&a->d_coefs, poly_get_size(b), ctx->get_stream()), "hipMallocAsync");
  else
    hip_ok(hipMalloc((void**)&a->d_coefs, poly_get_size(b)), "hipMalloc");
  hip_ok(hipMemsetAsync(a->d_coefs, 0, poly_get_size(b), ctx->get_stream()), "hipMemsetAsync");
  a->state = RNSSTATE;
  a->base = b;
}

void poly_free(Context* ctx, poly_t* a) {
  if (!a || a->state == NONINITIALIZED) return;
  if (stream_alloc()) {
    (void)hipFreeAsync(a->d_coefs, ctx->get_stream());
  } else {
    (void)hipStreamSynchronize(ctx->get_stream());
    (void)hipFree(a->d_coefs);
  }
  a->d_coefs = nullptr;
  a->state = NONINITIALIZED;
}

void poly_clear(Context* ctx, poly_t* a) {
  poly_init(ctx, a);
  hip_ok(hipMemsetAsync(a->d_coefs, 0, poly_get_size(a->base), ctx->get_stream()), "hipMemsetAsync");
  a->state = RNSSTATE;
}

void poly_copy(Context* ctx, poly_t* b, poly_t* a) {
  if (a->state == NONINITIALIZED) throw std::runtime_error("poly_copy of an uninitialised polynomial");
  poly_init(ctx, b, a->base);
  if (b == a) return;
  hip_ok(hipMemcpyAsync(b->d_coefs, a->d_coefs, poly_get_size(a->base), hipMemcpyDeviceToDevice, ctx->get_stream()),
         "hipMemcpyAsync");
  b->state = a->state;
}

void poly_copy_to_device(Context* ctx, poly_t* a, uint64_t* h) {
  poly_init(ctx, a);
  hip_ok(hipMemcpyAsync(a->d_coefs, h, poly_get_size(a->base), hipMemcpyHostToDevice, ctx->get_stream()),
         "hipMemcpyAsync");
  hip_ok(hipStreamSynchronize(ctx->get_stream()), "hipStreamSynchronize");
  a->state = RNSSTATE;
}

void COMMONEngine::execute(Context* ctx, poly_t* p, const transform_directions dir) {
  if ((p->state == TRANSSTATE && dir == FORWARD) || (p->state == RNSSTATE && dir == INVERSE)) return;
  ok(aoa_ntt(CUDAManager::params(), ctx->get_stream(), p->d_coefs, 0, CUDAManager::get_n_residues(p->base),
             dir == FORWARD ? AOA_FORWARD : AOA_INVERSE));
  p->state = dir == FORWARD ? TRANSSTATE : RNSSTATE;
}

uint64_t* poly_copy_to_host(Context* ctx, poly_t* a) {
  poly_init(ctx, a);
  poly_t b;
  poly_copy(ctx, &b, a);
  COMMONEngine::execute(ctx, &b, INVERSE);
  uint64_t* h = (uint64_t*)malloc(poly_get_size(a->base));
  hip_ok(hipMemcpyAsync(h, b.d_coefs, poly_get_size(a->base), hipMemcpyDeviceToHost, ctx->get_stream()),
         "hipMemcpyAsync");
  hip_ok(hipStreamSynchronize(ctx->get_stream()), "hipStreamSynchronize");
  poly_free(ctx, &b);
  return h;
}

bool poly_are_equal(Context* ctx, poly_t* a, poly_t* b) {
  poly_init(ctx, a);
  poly_init(ctx, b);
  if (a->base != b->base) return false;
  uint64_t* ha = poly_copy_to_host(ctx, a);
  uint64_t* hb = poly_copy_to_host(ctx, b);
  bool eq = std::memcmp(ha, hb, poly_get_size(a->base)) == 0;
  free(ha);
  free(hb);
  return eq;
}

static void eltwise(Context* ctx, int op, poly_t* c, poly_t* a, poly_t* b, poly_t* d, poly_bases base) {
  ok(aoa_eltwise(CUDAManager::params(), ctx->get_stream(), op, c->d_coefs, a->d_coefs, b ? b->d_coefs : nullptr,
                 d ? d->d_coefs : nullptr, 0, CUDAManager::get_n_residues(base)));
}

// poly_add / poly_sub (poly_t.cpp:137-191): a state mismatch brings both
// operands to the coefficient domain first.
static void poly_addsub(Context* ctx, int op, poly_t* c, poly_t* a, poly_t* b) {
  poly_bases base = std::min(a->base, b->base);
  poly_init(ctx, a, base);
  poly_init(ctx, b, base);
  poly_init(ctx, c, base);
  if (a->state != b->state) {
    COMMONEngine::execute(ctx, a, INVERSE);
    COMMONEngine::execute(ctx, b, INVERSE);
  }
  eltwise(ctx, op, c, a, b, nullptr, base);
  c->state = a->state;
}

void poly_add(Context* ctx, poly_t* c, poly_t* a, poly_t* b) { poly_addsub(ctx, AOA_ADD, c, a, b); }
void poly_sub(Context* ctx, poly_t* c, poly_t* a, poly_t* b) { poly_addsub(ctx, AOA_SUB, c, a, b); }

void poly_mul(Context* ctx, poly_t* c, poly_t* a, poly_t* b) {  // poly_t.cpp:201-220
  poly_bases base = std::min(a->base, b->base);
  poly_init(ctx, a, base);
  poly_init(ctx, b, base);
  poly_init(ctx, c, base);
  COMMONEngine::execute(ctx, a, FORWARD);
  COMMONEngine::execute(ctx, b, FORWARD);
  eltwise(ctx, AOA_MUL, c, a, b, nullptr, base);
  c->state = a->state;
}

void poly_add_add(Context* ctx, poly_t* c1, poly_t* a1, poly_t* b1, poly_t* c2, poly_t* a2, poly_t* b2) {
  poly_add(ctx, c1, a1, b1);
  poly_add(ctx, c2, a2, b2);
}

void poly_mul_add(Context* ctx, poly_t* d, poly_t* a, poly_t* b, poly_t* c) {  // poly_t.cpp:274-300
  poly_bases base = std::min(std::min(a->base, b->base), c->base);
  poly_init(ctx, d, base);
  poly_init(ctx, a, base);
  poly_init(ctx, b, base);
  poly_init(ctx, c, base);
  COMMONEngine::execute(ctx, a, FORWARD);
  COMMONEngine::execute(ctx, b, FORWARD);
  COMMONEngine::execute(ctx, c, FORWARD);
  eltwise(ctx, AOA_MULADD, d, a, b, c, base);
  d->state = a->state;
}

// poly_t.cpp:302-342.  The reference's kernel covers every QB residue
// whatever the operands' base, so QB-base operands are required here.
void poly_dr2(Context* ctx, poly_t* ct21, poly_t* ct22, poly_t* ct23, poly_t* ct01, poly_t* ct02, poly_t* ct11,
              poly_t* ct12) {
  const poly_bases base = ct01->base;
  for (poly_t* p : {ct21, ct22, ct23, ct01, ct02, ct11, ct12}) {
    poly_init(ctx, p, base);
    if (p->base != QBBase) throw std::runtime_error("poly_dr2 needs QB-base polynomials");
  }
  for (poly_t* p : {ct01, ct02, ct11, ct12}) COMMONEngine::execute(ctx, p, FORWARD);
  CUDAManager::execute_dr2(ctx, ct21->d_coefs, ct22->d_coefs, ct23->d_coefs, ct01->d_coefs, ct02->d_coefs,
                           ct11->d_coefs, ct12->d_coefs);
  for (poly_t* p : {ct21, ct22, ct23}) {
    p->state = TRANSSTATE;
    p->base = base;
  }
}

void poly_negate(Context* ctx, poly_t* a) {
  poly_init(ctx, a);
  eltwise(ctx, AOA_NEGATE, a, a, nullptr, nullptr, a->base);
}

void poly_rho(Context* ctx, poly_t* b, poly_t* a) {
  poly_init(ctx, a, QBBase);
  poly_init(ctx, b, QBBase);
  ok(aoa_rho(CUDAManager::params(), ctx->get_stream(), b->d_coefs, a->d_coefs));
  b->state = a->state;
}

void poly_modup(Context* ctx, poly_t* a, poly_t* b, int level) {  // poly_t.cpp:345-361
  poly_init(ctx, a, QBBase);
  poly_init(ctx, b);
  hip_ok(hipMemcpyAsync(a->d_coefs, b->d_coefs, poly_get_size(QBase), hipMemcpyDeviceToDevice, ctx->get_stream()),
         "hipMemcpyAsync");
  a->state = b->state;
  COMMONEngine::execute(ctx, a, INVERSE);
  ok(aoa_modup(CUDAManager::params(), ctx->get_stream(), a->d_coefs, level));
}

void poly_moddown(Context* ctx, poly_t* a, poly_t* b, int level) {  // poly_t.cpp:363-378
  poly_init(ctx, a, QBase);
  poly_init(ctx, b, QBBase);
  COMMONEngine::execute(ctx, b, INVERSE);
  ok(aoa_moddown(CUDAManager::params(), ctx->get_stream(), a->d_coefs, b->d_coefs, level));
  a->state = RNSSTATE;
}

void poly_rescale(Context* ctx, poly_t* a, poly_t* b, int level) {  // poly_t.cpp:395-411
  poly_init(ctx, a);
  poly_init(ctx, b);
  COMMONEngine::execute(ctx, a, INVERSE);
  COMMONEngine::execute(ctx, b, INVERSE);
  ok(aoa_rescale(CUDAManager::params(), ctx->get_stream(), a->d_coefs, b->d_coefs, level));
  a->state = RNSSTATE;
  b->state = RNSSTATE;
}

std::string poly_residue_to_string(Context* ctx, poly_t* a, int id) {
  uint64_t* h = poly_copy_to_host(ctx, a);
  const int N = CUDAManager::N;
  const int lo = id >= 0 ? id * N : 0;
  const int hi = id >= 0 ? (id + 1) * N : CUDAManager::get_n_residues(a->base) * N;
  std::string s;
  for (int i = lo; i < hi; ++i) s += std::to_string(h[i]) + ",";
  free(h);
  return s;
}

// ------------------------------------------------------------ sampler ------

Sampler::Sampler(Context* c, uint64_t s) : ctx(c), seed(s), nonce(0) {}

// Every draw takes the next nonce, so a context's sequence of polynomials is
// fixed by its seed.
void Sampler::draw(poly_t* p, poly_bases b, int kind) {
  poly_init(ctx, p, b);
  ok(aoa_sample(CUDAManager::params(), ctx->get_stream(), p->d_coefs, 0, CUDAManager::get_n_residues(b), kind, seed,
                nonce++));
  p->state = RNSSTATE;
}

void Sampler::sample_uniform(poly_t* p, poly_bases b) { draw(p, b, AOA_SAMPLE_UNIFORM); }  // sampler.cu:21-45

void Sampler::sample_DG(poly_t* p, poly_bases b) { draw(p, b, AOA_SAMPLE_GAUSSIAN); }  // sampler.cu:72-117

void Sampler::sample_Z0(poly_t* p, poly_bases b) { draw(p, b, AOA_SAMPLE_ZO); }  // sampler.cu:119-155

void Sampler::sample_hw(poly_t* p, poly_bases b) { draw(p, b, AOA_SAMPLE_HW); }  // sampler.cu:157-192

void Sampler::sample_narrow(poly_t* p, poly_bases b) { draw(p, b, AOA_SAMPLE_NARROW); }  // sampler.cu:194-225

// ------------------------------------------------------------ ciphertexts --

void cipher_init(Context* ctx, cipher_t* c, int slots) {  // cipher_t.cpp:3-9
  poly_init(ctx, &c->c[0]);
  poly_init(ctx, &c->c[1]);
  c->level = (int)CUDAManager::RNSQPrimes.size() - 1;
  c->scale = (int64_t)((uint64_t)1 << CUDAManager::scalingfactor);
  c->slots = slots;
}

void cipher_clear(Context* ctx, cipher_t* c) {
  if (c->c[0].state != NONINITIALIZED && c->c[1].state != NONINITIALIZED) {
    poly_clear(ctx, &c->c[0]);
    poly_clear(ctx, &c->c[1]);
    c->level = (int)CUDAManager::RNSQPrimes.size() - 1;
    c->scale = (int64_t)((uint64_t)1 << CUDAManager::scalingfactor);
  } else {
    cipher_init(ctx, c);
  }
}

void cipher_free(Context* ctx, cipher_t* c) {
  poly_free(ctx, &c->c[0]);
  poly_free(ctx, &c->c[1]);
}

void cipher_copy(Context* ctx, cipher_t* b, cipher_t* a) {
  poly_copy(ctx, &b->c[0], &a->c[0]);
  poly_copy(ctx, &b->c[1], &a->c[1]);
  b->level = a->level;
  b->scale = a->scale;
  b->slots = a->slots;
}

// ------------------------------------------------------------ encoder ------

// The FFTs run on the device (aoa_special_fft): same loops, twiddles and
// rounding as the reference's host fftSpecial / fftSpecialInv.
static void special_fft_host(std::complex<double>* vals, long size, int dir) {
  void* d = nullptr;
  const size_t bytes = (size_t)size * sizeof(std::complex<double>);
  hip_ok(hipMalloc(&d, bytes), "hipMalloc");
  hip_ok(hipMemcpy(d, vals, bytes, hipMemcpyHostToDevice), "hipMemcpy");
  const int rc = aoa_special_fft(CUDAManager::params(), nullptr, (double*)d, (int)size, dir);
  hipError_t e = rc == AOA_OK ? hipMemcpy(vals, d, bytes, hipMemcpyDeviceToHost) : hipSuccess;
  (void)hipFree(d);
  ok(rc);
  hip_ok(e, "hipMemcpy");
}

void fftSpecial(std::complex<double>* vals, const long size) { special_fft_host(vals, size, AOA_FORWARD); }

void fftSpecialInv(std::complex<double>* vals, const long size) { special_fft_host(vals, size, AOA_INVERSE); }

// encoder.cu:143-221: the values (filled, then zeros) go to the device, where
// aoa_encode runs the inverse FFT and writes every Q residue.
void ckks_encode(CKKSContext* ctx, uint64_t* d_out, std::complex<double>* in, int filled, int empty,
                 uint64_t scale) {
  const int slots = filled + empty;
  const int N = CUDAManager::N;
  if (filled < 0 || empty < 0 || slots <= 0 || slots > N / 2 || (slots & (slots - 1)))
    throw std::runtime_error("slots must be a power of two <= N/2");
  if (in != ctx->h_val) std::memmove(ctx->h_val, in, (size_t)filled * sizeof(std::complex<double>));
  for (int i = filled; i < slots; ++i) ctx->h_val[i] = 0.0;
  hip_ok(hipMemcpyAsync(ctx->d_val, ctx->h_val, (size_t)slots * sizeof(std::complex<double>), hipMemcpyHostToDevice,
                        ctx->get_stream()),
         "hipMemcpyAsync");
  ok(aoa_encode(CUDAManager::params(), ctx->get_stream(), d_out, (double*)ctx->d_val, slots, scale,
                (int)CUDAManager::RNSQPrimes.size()));
  hip_ok(hipStreamSynchronize(ctx->get_stream()), "hipStreamSynchronize");
}

// encoder.cu:223-298: residue 0 centred and divided by the scale, then the
// forward FFT, on the device; the slot values come back to the host.
void ckks_decode(CKKSContext* ctx, std::complex<double>* out, uint64_t* d_a, int slots, uint64_t scale) {
  ok(aoa_decode(CUDAManager::params(), ctx->get_stream(), (double*)ctx->d_val, d_a, slots, scale));
  hip_ok(hipMemcpyAsync(out, ctx->d_val, (size_t)slots * sizeof(std::complex<double>), hipMemcpyDeviceToHost,
                        ctx->get_stream()),
         "hipMemcpyAsync");
  hip_ok(hipStreamSynchronize(ctx->get_stream()), "hipStreamSynchronize");
}

// encoder.cu:325-363: the input is brought to the coefficient domain in place,
// then X -> X^(5^r mod 2N) on every residue of its base.
void rotate_slots_left(Context* ctx, poly_t* b, poly_t* a, int rotSlots) {
  poly_init(ctx, b, a->base);
  COMMONEngine::execute(ctx, a, INVERSE);
  ok(aoa_automorphism(CUDAManager::params(), ctx->get_stream(), b->d_coefs, a->d_coefs,
                      CUDAManager::get_n_residues(a->base), rotSlots));
  b->state = RNSSTATE;
}

void rotate_slots_right(Context* ctx, poly_t* b, poly_t* a, int rotSlots) {
  rotate_slots_left(ctx, b, a, (CUDAManager::N >> 1) - rotSlots);
}

void conjugate_slots(Context* ctx, poly_t* a) {  // encoder.cu:404-427
  COMMONEngine::execute(ctx, a, FORWARD);
  poly_t t;
  poly_init(ctx, &t, a->base);
  ok(aoa_conjugate(CUDAManager::params(), ctx->get_stream(), t.d_coefs, a->d_coefs,
                   CUDAManager::get_n_residues(a->base)));
  hip_ok(hipMemcpyAsync(a->d_coefs, t.d_coefs, poly_get_size(a->base), hipMemcpyDeviceToDevice, ctx->get_stream()),
         "hipMemcpyAsync");
  poly_free(ctx, &t);
}

// ------------------------------------------------------------ CKKSContext --

CKKSContext::CKKSContext() : Context() {
  u = new poly_t;
  m = new poly_t;
  e = new poly_t;
  poly_init(this, u);
  poly_init(this, m);
  poly_init(this, e);
  for (auto& c : caux) cipher_init(this, &c);
  hip_ok(hipHostMalloc((void**)&h_val, (size_t)CUDAManager::N * sizeof(std::complex<double>), 0), "hipHostMalloc");
  hip_ok(hipMalloc((void**)&d_val, (size_t)CUDAManager::N * sizeof(std::complex<double>)), "hipMalloc");
  // Maclaurin series (ckkscontext.h:101-135), the reference's values as they
  // are (its cosine row has 1/24 at x^3, which matches cos only on [0, 1])
  maclaurin_coeffs[EXPONENT] = new double[11]{1, 1, 1. / 2, 1. / 6, 1. / 24, 1. / 120, 1. / 720, 1. / 5040,
                                              1. / 40320, 1. / 362880, 1. / 3628800};
  maclaurin_coeffs[SIN] = new double[12]{0, 1, 0, -1. / 6, 0, 1. / 120, 0, -1. / 5040, 0, 1. / 362880, 0,
                                         -1. / 39916800};
  maclaurin_coeffs[COS] = new double[12]{1, 0, -0.5, 1. / 24, 0, -1. / 720, 0, 1. / 40320, 0, -1. / 362880, 0,
                                         1. / 362880};
  maclaurin_coeffs[SIGMOID] = new double[10]{1. / 2, 1. / 4, 0, -1. / 48, 0, 1. / 480, 0, -17. / 80640, 0,
                                             31. / 1451520};
  {
    double* l1 = new double[21];
    l1[0] = 0;
    for (int i = 1; i < 21; ++i) l1[i] = -1. / i;
    maclaurin_coeffs[LN1MINUS] = l1;
    double* lg = new double[40];
    for (int i = 0; i < 40; ++i) lg[i] = (i % 2) ? 1. / i : 0;
    maclaurin_coeffs[LOGARITHM] = lg;
  }
  const int top = (int)CUDAManager::RNSQPrimes.size() - 1;
  ws_bytes = aoa_workspace_bytes(CUDAManager::params(), top);
  hip_ok(hipMalloc(&d_ws, ws_bytes), "hipMalloc");
}

CKKSContext::~CKKSContext() {
  (void)hipStreamSynchronize(get_stream());
  for (auto& c : caux) cipher_free(this, &c);
  (void)hipHostFree(h_val);
  (void)hipFree(d_val);
  for (auto& kv : maclaurin_coeffs) delete[] kv.second;
  poly_free(this, u);
  poly_free(this, m);
  poly_free(this, e);
  delete u;
  delete m;
  delete e;
  (void)hipFree(d_ws);
}

void CKKSContext::encode(poly_t* a, int64_t* scale, std::complex<double>* val, int slots, int empty_slots) {
  poly_init(this, a);
  if (*scale == -1) *scale = (int64_t)((uint64_t)1 << CUDAManager::scalingfactor);
  ckks_encode(this, a->d_coefs, val, slots, empty_slots, (uint64_t)*scale);
  a->state = RNSSTATE;
}

void CKKSContext::decode(std::complex<double>* val, poly_t* a, int64_t scale, int slots) {
  poly_init(this, a);
  COMMONEngine::execute(this, a, INVERSE);
  ckks_decode(this, val, a->d_coefs, slots, (uint64_t)scale);
}

// ------------------------------------------------------------ keys ---------

SecretKey* ckks_new_sk(Context* ctx) {  // ckkskeys.cpp:61-72
  SecretKey* sk = new SecretKey;
  poly_init(ctx, &sk->s, QBBase);
  ctx->get_sampler()->sample_hw(&sk->s, QBBase);
  return sk;
}

void ckks_free_sk(Context* ctx, SecretKey* sk) { poly_free(ctx, &sk->s); }

PublicKey* ckks_new_pk(Context* ctx, SecretKey* sk) {  // ckkskeys.cpp:79-99: b = e - a s
  PublicKey* pk = new PublicKey;
  poly_init(ctx, &pk->a, QBBase);
  poly_init(ctx, &pk->b, QBBase);
  poly_t e;
  poly_init(ctx, &e, QBBase);
  ctx->get_sampler()->sample_uniform(&pk->a, QBBase);
  ctx->get_sampler()->sample_DG(&pk->b, QBBase);
  poly_mul(ctx, &e, &pk->a, &sk->s);
  poly_sub(ctx, &pk->b, &pk->b, &e);
  COMMONEngine::execute(ctx, &pk->a, FORWARD);
  COMMONEngine::execute(ctx, &pk->b, FORWARD);
  poly_free(ctx, &e);
  return pk;
}

EvaluationKey* ckks_new_evk(Context* ctx, poly_t* s1, poly_t* s2) {  // ckkskeys.cpp:101-131
  EvaluationKey* evk = new EvaluationKey;
  poly_init(ctx, &evk->a, QBBase);
  poly_init(ctx, &evk->b, QBBase);
  poly_t a, e;
  poly_init(ctx, &a, QBBase);
  poly_init(ctx, &e, QBBase);
  poly_rho(ctx, &evk->b, s1);
  ctx->get_sampler()->sample_uniform(&evk->a, QBBase);
  ctx->get_sampler()->sample_DG(&e, QBBase);
  poly_add(ctx, &evk->b, &evk->b, &e);
  poly_mul(ctx, &a, &evk->a, s2);
  poly_sub(ctx, &evk->b, &evk->b, &a);
  // keys are kept in the NTT domain (the reference transforms b on first use)
  COMMONEngine::execute(ctx, &evk->a, FORWARD);
  COMMONEngine::execute(ctx, &evk->b, FORWARD);
  poly_free(ctx, &a);
  poly_free(ctx, &e);
  return evk;
}

EvaluationKey* ckks_new_mtk(Context* ctx, SecretKey* sk) {  // ckkskeys.cpp:133-144
  poly_t s2;
  poly_init(ctx, &s2, QBBase);
  poly_mul(ctx, &s2, &sk->s, &sk->s);
  EvaluationKey* evk = ckks_new_evk(ctx, &s2, &sk->s);
  poly_free(ctx, &s2);
  return evk;
}

std::map<int, RotationKey*> ckks_new_rtk_left(Context* ctx, SecretKey* sk, const std::vector<int>& steps) {
  std::map<int, RotationKey*> r;
  poly_t s2;
  poly_init(ctx, &s2, QBBase);
  for (int i : steps) {
    rotate_slots_left(ctx, &s2, &sk->s, i);
    r[i] = ckks_new_evk(ctx, &s2, &sk->s);
  }
  poly_free(ctx, &s2);
  return r;
}

std::map<int, RotationKey*> ckks_new_rtk_left(Context* ctx, SecretKey* sk) {  // ckkskeys.cpp:146-163
  std::vector<int> steps;
  for (int i = 1; i < CUDAManager::N; i *= 2) steps.push_back(i);
  return ckks_new_rtk_left(ctx, sk, steps);
}

std::map<int, RotationKey*> ckks_new_rtk_right(Context* ctx, SecretKey* sk) {  // ckkskeys.cpp:165-181
  std::map<int, RotationKey*> r;
  poly_t s2;
  poly_init(ctx, &s2, QBBase);
  for (int i = 1; i < CUDAManager::N; i *= 2) {
    rotate_slots_right(ctx, &s2, &sk->s, i);
    r[i] = ckks_new_evk(ctx, &s2, &sk->s);
  }
  poly_free(ctx, &s2);
  return r;
}

ConjugationKey* ckks_new_cjk(Context* ctx, SecretKey* sk) {  // ckkskeys.cpp:183-193
  poly_t s2;
  poly_init(ctx, &s2, QBBase);
  poly_copy(ctx, &s2, &sk->s);
  conjugate_slots(ctx, &s2);
  ConjugationKey* cjk = ckks_new_evk(ctx, &s2, &sk->s);
  poly_free(ctx, &s2);
  return cjk;
}

static void free_swk(Context* ctx, SwitchKey* k) {
  if (!k) return;
  poly_free(ctx, &k->a);
  poly_free(ctx, &k->b);
  delete k;
}

void ckkskeychain_free(Context* ctx, CKKSKeychain* k) {
  free_swk(ctx, k->pk);
  free_swk(ctx, k->evk);
  free_swk(ctx, k->cjk);
  for (auto& kv : k->rtk_left) free_swk(ctx, kv.second);
  for (auto& kv : k->rtk_right) free_swk(ctx, kv.second);
  k->pk = k->evk = k->cjk = nullptr;
  k->rtk_left.clear();
  k->rtk_right.clear();
}

// ------------------------------------------------------------ scheme -------

CKKSKeychain* ckks_keygen(CKKSContext* ctx, SecretKey* sk) {  // ckks.cpp:46-75
  CKKSKeychain* keys = new CKKSKeychain;
  keys->pk = ckks_new_pk(ctx, sk);
  keys->evk = ckks_new_mtk(ctx, sk);
  keys->rtk_left = ckks_new_rtk_left(ctx, sk);
  keys->rtk_right = ckks_new_rtk_right(ctx, sk);
  keys->cjk = ckks_new_cjk(ctx, sk);
  keys->sk = sk;
  ctx->keys = *keys;
  return keys;
}

cipher_t* ckks_encrypt_poly(CKKSContext* ctx, cipher_t* ct, poly_t* m) {  // ckks.cpp:87-105
  cipher_init(ctx, ct);
  if (!ctx->keys.pk) throw std::runtime_error("no public key: call ckks_keygen first");
  ctx->get_sampler()->sample_Z0(ctx->u, QBase);
  ctx->get_sampler()->sample_DG(ctx->e, QBase);
  poly_add(ctx, ctx->e, ctx->e, m);
  poly_mul_add(ctx, &ct->c[0], ctx->u, &ctx->keys.pk->b, ctx->e);
  ctx->get_sampler()->sample_DG(ctx->e, QBase);
  poly_mul_add(ctx, &ct->c[1], ctx->u, &ctx->keys.pk->a, ctx->e);
  poly_clear(ctx, ctx->u);
  poly_clear(ctx, ctx->e);
  return ct;
}

cipher_t* ckks_encrypt_poly(CKKSContext* ctx, poly_t* m) {
  cipher_t* ct = new cipher_t;
  cipher_init(ctx, ct);
  return ckks_encrypt_poly(ctx, ct, m);
}

// Encodes at ct's current scale (ckks.cpp:131-148); ckks_encrypt_poly then
// resets level and scale, as the reference does.  A never-initialised ct is
// initialised first.
cipher_t* ckks_encrypt(CKKSContext* ctx, cipher_t* ct, std::complex<double>* val, int slots, int empty_slots) {
  if (ct->c[0].state == NONINITIALIZED || ct->c[1].state == NONINITIALIZED) cipher_init(ctx, ct, slots + empty_slots);
  ctx->encode(ctx->m, &ct->scale, val, slots, empty_slots);
  ckks_encrypt_poly(ctx, ct, ctx->m);
  ct->slots = slots + empty_slots;
  poly_clear(ctx, ctx->m);
  return ct;
}

cipher_t* ckks_encrypt(CKKSContext* ctx, std::complex<double>* val, int slots, int empty_slots) {
  cipher_t* ct = new cipher_t;
  return ckks_encrypt(ctx, ct, val, slots, empty_slots);
}

cipher_t* ckks_encrypt(CKKSContext* ctx, double val, int slots, int empty_slots) {
  std::vector<std::complex<double>> v(slots, std::complex<double>(val, 0));
  return ckks_encrypt(ctx, v.data(), slots, empty_slots);
}

cipher_t* ckks_encrypt_scaled(CKKSContext* ctx, std::complex<double>* val, int slots, int scale_bits) {
  if (scale_bits <= 0 || scale_bits > 62) throw std::runtime_error("scale must be 2^1 .. 2^62");
  cipher_t* ct = new cipher_t;
  cipher_init(ctx, ct, slots);
  ct->scale = (int64_t)((uint64_t)1 << scale_bits);
  ctx->encode(ctx->m, &ct->scale, val, slots, 0);
  ckks_encrypt_poly(ctx, ct, ctx->m);
  ct->scale = (int64_t)((uint64_t)1 << scale_bits);
  ct->slots = slots;
  poly_clear(ctx, ctx->m);
  return ct;
}

poly_t* ckks_decrypt_poly(CKKSContext* ctx, poly_t* m, cipher_t* c, SecretKey* sk) {  // ckks.cpp:179-192
  poly_mul_add(ctx, m, &c->c[1], &sk->s, &c->c[0]);
  return m;
}

std::complex<double>* ckks_decrypt(CKKSContext* ctx, std::complex<double>* val, cipher_t* c, SecretKey* sk) {
  ckks_decrypt_poly(ctx, ctx->m, c, sk);
  ctx->decode(val, ctx->m, c->scale, c->slots);
  return val;
}

std::complex<double>* ckks_decrypt(CKKSContext* ctx, cipher_t* c, SecretKey* sk) {
  auto* val = (std::complex<double>*)malloc(sizeof(std::complex<double>) * c->slots);
  for (int i = 0; i < c->slots; ++i) new (&val[i]) std::complex<double>(0, 0);
  return ckks_decrypt(ctx, val, c, sk);
}

void ckks_rescale(CKKSContext* ctx, cipher_t* c) {  // ckks.cpp:220-230
  if (c->level == 0) throw std::runtime_error("The ciphertext level is 0. No more rescaling is possible.");
  poly_rescale(ctx, &c->c[0], &c->c[1], c->level);
  c->level--;
}

void ckks_add(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, cipher_t* c2) {  // ckks.cpp:232-250
  if (c1->slots != c2->slots) throw std::runtime_error("Ciphertexts must have the same number of slots.");
  poly_add_add(ctx, &c3->c[0], &c1->c[0], &c2->c[0], &c3->c[1], &c1->c[1], &c2->c[1]);
  c3->level = std::min(c1->level, c2->level);
  c3->slots = c1->slots;
  c3->scale = c1->scale;
}

cipher_t* ckks_add(CKKSContext* ctx, cipher_t* c1, cipher_t* c2) {
  cipher_t* c3 = new cipher_t;
  cipher_init(ctx, c3);
  ckks_add(ctx, c3, c1, c2);
  return c3;
}

void ckks_sub(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, cipher_t* c2) {
  if (c1->slots != c2->slots) throw std::runtime_error("Ciphertexts must have the same number of slots.");
  poly_sub(ctx, &c3->c[0], &c1->c[0], &c2->c[0]);
  poly_sub(ctx, &c3->c[1], &c1->c[1], &c2->c[1]);
  c3->level = std::min(c1->level, c2->level);
  c3->slots = c1->slots;
  c3->scale = c1->scale;
}

cipher_t* ckks_sub(CKKSContext* ctx, cipher_t* c1, cipher_t* c2) {
  cipher_t* c3 = new cipher_t;
  cipher_init(ctx, c3);
  ckks_sub(ctx, c3, c1, c2);
  return c3;
}

// HomMul + relinearisation (ckks.cpp:273-358) through the fused C-ABI path.
static void mul_core(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, cipher_t* c2, bool rescale) {
  if (c1->c[0].state == NONINITIALIZED || c2->c[0].state == NONINITIALIZED)
    throw std::runtime_error("uninitialised ciphertext");
  if (c1->level == 0 || c2->level == 0)
    throw std::runtime_error("The ciphertext level is 0. No more rescaling is possible.");
  if (c1->slots != c2->slots) throw std::runtime_error("Ciphertexts must have the same number of slots.");
  if (!ctx->keys.evk) throw std::runtime_error("no evaluation key: call ckks_keygen first");
  const int level = std::min(c1->level, c2->level);
  for (poly_t* p : {&c1->c[0], &c1->c[1], &c2->c[0], &c2->c[1]}) COMMONEngine::execute(ctx, p, FORWARD);
  poly_t o[2];
  poly_init(ctx, &o[0]);
  poly_init(ctx, &o[1]);
  aoa_params* P = CUDAManager::params();
  auto fn = rescale ? aoa_hommul_relin_rescale : aoa_hommul_relin;
  ok(fn(P, ctx->get_stream(), o[0].d_coefs, o[1].d_coefs, c1->c[0].d_coefs, c1->c[1].d_coefs, c2->c[0].d_coefs,
        c2->c[1].d_coefs, ctx->keys.evk->b.d_coefs, ctx->keys.evk->a.d_coefs, level, ctx->d_ws, ctx->ws_bytes));
  const unsigned __int128 s = (unsigned __int128)(uint64_t)c1->scale * (uint64_t)c2->scale;
  const int64_t scale = (int64_t)(s / CUDAManager::RNSQPrimes[level]);
  const int slots = c1->slots;
  poly_init(ctx, &c3->c[0]);
  poly_init(ctx, &c3->c[1]);
  poly_copy(ctx, &c3->c[0], &o[0]);
  poly_copy(ctx, &c3->c[1], &o[1]);
  c3->c[0].state = c3->c[1].state = RNSSTATE;
  poly_free(ctx, &o[0]);
  poly_free(ctx, &o[1]);
  c3->level = rescale ? level - 1 : level;
  c3->slots = slots;
  c3->scale = scale;
}

void ckks_mul_without_rescale(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, cipher_t* c2) {
  mul_core(ctx, c3, c1, c2, false);
}

void ckks_mul(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, cipher_t* c2) { mul_core(ctx, c3, c1, c2, true); }

cipher_t* ckks_mul(CKKSContext* ctx, cipher_t* c1, cipher_t* c2) {
  cipher_t* c3 = new cipher_t;
  cipher_init(ctx, c3);
  ckks_mul(ctx, c3, c1, c2);
  return c3;
}

void ckks_square(CKKSContext* ctx, cipher_t* c2, cipher_t* c1) { mul_core(ctx, c2, c1, c1, true); }

cipher_t* ckks_square(CKKSContext* ctx, cipher_t* c1) {
  cipher_t* c2 = new cipher_t;
  cipher_init(ctx, c2);
  ckks_square(ctx, c2, c1);
  return c2;
}

static void rotate_core(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, int galois_slots, RotationKey* key) {
  COMMONEngine::execute(ctx, &c1->c[0], INVERSE);
  COMMONEngine::execute(ctx, &c1->c[1], INVERSE);
  poly_t o[2];
  poly_init(ctx, &o[0]);
  poly_init(ctx, &o[1]);
  ok(aoa_rotate_left(CUDAManager::params(), ctx->get_stream(), o[0].d_coefs, o[1].d_coefs, c1->c[0].d_coefs,
                     c1->c[1].d_coefs, key->b.d_coefs, key->a.d_coefs, galois_slots, c1->level, ctx->d_ws,
                     ctx->ws_bytes));
  const int level = c1->level, slots = c1->slots;
  const int64_t scale = c1->scale;
  poly_init(ctx, &c2->c[0]);
  poly_init(ctx, &c2->c[1]);
  poly_copy(ctx, &c2->c[0], &o[0]);
  poly_copy(ctx, &c2->c[1], &o[1]);
  c2->c[0].state = c2->c[1].state = RNSSTATE;
  poly_free(ctx, &o[0]);
  poly_free(ctx, &o[1]);
  c2->level = level;
  c2->slots = slots;
  c2->scale = scale;
}

void ckks_rotate_left(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, int rotSlots) {  // ckks.cpp:982-1009
  auto it = ctx->keys.rtk_left.find(rotSlots);
  if (it == ctx->keys.rtk_left.end()) throw std::runtime_error("no left rotation key for this step");
  rotate_core(ctx, c2, c1, rotSlots, it->second);
}

void ckks_rotate_right(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, int rotSlots) {  // ckks.cpp:953-980
  auto it = ctx->keys.rtk_right.find(rotSlots);
  if (it == ctx->keys.rtk_right.end()) throw std::runtime_error("no right rotation key for this step");
  rotate_core(ctx, c2, c1, (CUDAManager::N >> 1) - rotSlots, it->second);
}

void ckks_sumslots(CKKSContext* ctx, cipher_t* c2, cipher_t* c1) {  // ckks.cpp:479-495
  if (c1 != c2) cipher_copy(ctx, c2, c1);
  for (int i = (c2->slots >> 1); i >= 1; i /= 2) {
    ckks_rotate_left(ctx, &ctx->caux[0], c2, i);
    ckks_add(ctx, c2, c2, &ctx->caux[0]);
  }
}

cipher_t* ckks_sumslots(CKKSContext* ctx, cipher_t* c) {
  cipher_t* c2 = new cipher_t;
  cipher_init(ctx, c2);
  ckks_sumslots(ctx, c2, c);
  return c2;
}
