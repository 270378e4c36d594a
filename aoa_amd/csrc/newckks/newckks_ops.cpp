// newckks operations with plaintext / constant operands and the compositions
// built on them (aoa-ntt/src/ckks/ckks.cpp:423-1189), over the C-ABI of the
// gfx950 core.  The compositions keep the reference's sequence of operations
// and its level / scale bookkeeping, so that they produce the same
// ciphertexts from the same inputs; every residue operation runs in libaoa.so.
//
// The compositions follow aoa-ntt's ckks.cpp, distributed under the MIT License:
//
// Copyright (c) 2021 Pedro Alves
//
// Permission is hereby granted, free of charge, to any person obtaining a copy
// of this software and associated documentation files (the "Software"), to deal
// in the Software without restriction, including without limitation the rights
// to use, copy, modify, merge, publish, distribute, sublicense, and/or sell
// copies of the Software, and to permit persons to whom the Software is
// furnished to do so, subject to the following conditions:
//
// The above copyright notice and this permission notice shall be included in all
// copies or substantial portions of the Software.
//
// THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS OR
// IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF MERCHANTABILITY,
// FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT. IN NO EVENT SHALL THE
// AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY CLAIM, DAMAGES OR OTHER
// LIABILITY, WHETHER IN AN ACTION OF CONTRACT, TORT OR OTHERWISE, ARISING FROM,
// OUT OF OR IN CONNECTION WITH THE SOFTWARE OR THE USE OR OTHER DEALINGS IN THE
// SOFTWARE.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstring>
#include <stdexcept>
#include <string>

#include "../../../include/newckks/newckks.h"

static void ok(int rc) {
  if (rc != AOA_OK) throw std::runtime_error(aoa_last_error());
}

// ------------------------------------------------------------ poly layer ---

static void scalar_op(Context* ctx, int op, poly_t* c, poly_t* a, uint64_t b) {
  ok(aoa_eltwise_scalar(CUDAManager::params(), ctx->get_stream(), op, c->d_coefs, a->d_coefs, b, 0,
                        CUDAManager::get_n_residues(a->base)));
  c->state = a->state;
  c->base = a->base;
}

void poly_add(Context* ctx, poly_t* c, poly_t* a, uint64_t b) {  // poly_t.cpp:157-167
  poly_init(ctx, a);
  poly_init(ctx, c, a->base);
  COMMONEngine::execute(ctx, a, FORWARD);
  scalar_op(ctx, AOA_ADD, c, a, b);
}

void poly_sub(Context* ctx, poly_t* c, poly_t* a, uint64_t b) {  // poly_t.cpp:189-199
  poly_init(ctx, a);
  poly_init(ctx, c);
  COMMONEngine::execute(ctx, a, FORWARD);
  scalar_op(ctx, AOA_SUB, c, a, b);
}

void poly_mul(Context* ctx, poly_t* c, poly_t* a, uint64_t b) {  // poly_t.cpp:222-235
  poly_init(ctx, a);
  poly_init(ctx, c);
  COMMONEngine::execute(ctx, a, INVERSE);
  scalar_op(ctx, AOA_MUL, c, a, b);
}

// floor(s1 * s2 / q_level), the scale after a product (ckks.cpp:328-330).
static int64_t product_scale(int64_t s1, int64_t s2, int level) {
  const unsigned __int128 s = (unsigned __int128)(uint64_t)s1 * (uint64_t)s2;
  return (int64_t)(s / CUDAManager::RNSQPrimes[level]);
}

static cipher_t* fresh(CKKSContext* ctx) {
  cipher_t* c = new cipher_t;
  cipher_init(ctx, c);
  return c;
}

// ------------------------------------------------------------ constants ----

cipher_t* ckks_encrypt(CKKSContext* ctx, cipher_t* ct, double val, int slots, int empty_slots) {  // ckks.cpp:164-177
  std::complex<double>* cval = new std::complex<double>[slots];
  for (int i = 0; i < slots; ++i) cval[i] = {val, 0};
  cipher_t* r = ckks_encrypt(ctx, ct, cval, slots, empty_slots);
  delete[] cval;
  return r;
}

void ckks_add(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, double val) {  // ckks.cpp:497-518
  const uint64_t cnst = (uint64_t)(std::abs(val) * (double)c1->scale);
  if (val >= 0)
    poly_add(ctx, &c2->c[0], &c1->c[0], cnst);
  else
    poly_sub(ctx, &c2->c[0], &c1->c[0], cnst);
  poly_copy(ctx, &c2->c[1], &c1->c[1]);
  c2->level = c1->level;
  c2->slots = c1->slots;
  c2->scale = c1->scale;
}

cipher_t* ckks_add(CKKSContext* ctx, cipher_t* c1, double val) {
  cipher_t* c2 = fresh(ctx);
  ckks_add(ctx, c2, c1, val);
  return c2;
}

void ckks_mul_without_rescale(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, double val) {  // ckks.cpp:556-581
  const uint64_t x = (uint64_t)lrint(std::abs(val) * (double)c1->scale);
  poly_mul(ctx, &c2->c[0], &c1->c[0], x);
  poly_mul(ctx, &c2->c[1], &c1->c[1], x);
  if (val < 0) {
    poly_negate(ctx, &c2->c[0]);
    poly_negate(ctx, &c2->c[1]);
  }
  c2->level = c1->level;
  c2->slots = c1->slots;
  c2->scale = product_scale(c1->scale, c1->scale, c1->level);
}

cipher_t* ckks_mul_without_rescale(CKKSContext* ctx, cipher_t* c1, double val) {
  cipher_t* c2 = fresh(ctx);
  ckks_mul_without_rescale(ctx, c2, c1, val);
  return c2;
}

void ckks_mul(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, double val) {  // ckks.cpp:533-541
  ckks_mul_without_rescale(ctx, c2, c1, val);
  ckks_rescale(ctx, c2);
}

cipher_t* ckks_mul(CKKSContext* ctx, cipher_t* c1, double val) {
  cipher_t* c2 = fresh(ctx);
  ckks_mul(ctx, c2, c1, val);
  return c2;
}

void ckks_mul(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, std::complex<double>* val) {  // ckks.cpp:597-619
  ctx->encode(ctx->m, &c1->scale, val, c1->slots);
  poly_mul(ctx, &c2->c[0], &c1->c[0], ctx->m);
  poly_mul(ctx, &c2->c[1], &c1->c[1], ctx->m);
  c2->level = c1->level;
  c2->slots = c1->slots;
  c2->scale = product_scale(c1->scale, c1->scale, c1->level);
  ckks_rescale(ctx, c2);
}

cipher_t* ckks_mul(CKKSContext* ctx, cipher_t* c1, std::complex<double>* val) {
  cipher_t* c2 = fresh(ctx);
  ckks_mul(ctx, c2, c1, val);
  return c2;
}

void ckks_mul_without_rescale(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, poly_t* val) {  // ckks.cpp:635-652
  poly_mul(ctx, &c2->c[0], &c1->c[0], val);
  poly_mul(ctx, &c2->c[1], &c1->c[1], val);
  c2->level = c1->level;
  c2->slots = c1->slots;
  c2->scale = product_scale(c1->scale, c1->scale, c1->level);
}

void ckks_mul(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, poly_t* val) {  // ckks.cpp:654-662
  ckks_mul_without_rescale(ctx, c2, c1, val);
  ckks_rescale(ctx, c2);
}

cipher_t* ckks_mul(CKKSContext* ctx, cipher_t* c1, poly_t* val) {
  cipher_t* c2 = fresh(ctx);
  ckks_mul(ctx, c2, c1, val);
  return c2;
}

cipher_t* ckks_mul_without_rescale(CKKSContext* ctx, cipher_t* c1, cipher_t* c2) {
  cipher_t* c3 = fresh(ctx);
  ckks_mul_without_rescale(ctx, c3, c1, c2);
  return c3;
}

// ------------------------------------------------------------ compositions -

void ckks_inner_prod(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, cipher_t* c2, const int size) {  // :423-441
  cipher_init(ctx, c3);
  c3->slots = c1->slots;
  for (int i = 0; i < size; ++i) {
    ckks_mul_without_rescale(ctx, &ctx->caux[0], &c1[i], &c2[i]);
    ckks_add(ctx, c3, c3, &ctx->caux[0]);
  }
  ckks_rescale(ctx, c3);
}

void ckks_inner_prod(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, double* c2, const int size) {  // :443-460
  cipher_init(ctx, c3);
  c3->slots = c1->slots;
  for (int i = 0; i < size; ++i) {
    ckks_mul_without_rescale(ctx, &ctx->caux[0], &c1[i], c2[i]);
    ckks_add(ctx, c3, c3, &ctx->caux[0]);
  }
  ckks_rescale(ctx, c3);
}

template <typename T>
void ckks_batch_inner_prod(CKKSContext* ctx, cipher_t* c3, cipher_t* c1, T* c2) {  // :462-473
  ckks_mul(ctx, c3, c1, c2);
  ckks_sumslots(ctx, c3, c3);
}

template void ckks_batch_inner_prod<cipher_t>(CKKSContext*, cipher_t*, cipher_t*, cipher_t*);
template void ckks_batch_inner_prod<std::complex<double>>(CKKSContext*, cipher_t*, cipher_t*, std::complex<double>*);
template void ckks_batch_inner_prod<poly_t>(CKKSContext*, cipher_t*, cipher_t*, poly_t*);

// As in the reference, the squares go to the auxiliary ciphertext and c2 is
// left as initialised (a fresh encryption of 1 for x == 0); ckks_power
// computes the power.
void ckks_power_of_2(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, uint64_t x) {  // :693-715
  cipher_init(ctx, c2);
  if (x == 0) {
    ckks_encrypt(ctx, c2, 1.0, c1->slots);
    return;
  }
  cipher_clear(ctx, &ctx->caux[0]);
  cipher_copy(ctx, &ctx->caux[0], c1);
  while (x > 0) {
    ckks_square(ctx, &ctx->caux[0], &ctx->caux[0]);
    x >>= 1;
  }
}

cipher_t* ckks_power_of_2(CKKSContext* ctx, cipher_t* c1, uint64_t x) {
  cipher_t* c2 = fresh(ctx);
  ckks_power_of_2(ctx, c2, c1, x);
  return c2;
}

void ckks_power(CKKSContext* ctx, cipher_t* c2, cipher_t* c1, uint64_t x) {  // :731-761
  cipher_init(ctx, c2, c1->slots);
  if (x == 0) {
    ckks_encrypt(ctx, c2, 1.0, c1->slots);
    return;
  }
  cipher_clear(ctx, &ctx->caux[0]);
  cipher_copy(ctx, &ctx->caux[0], c1);
  bool first_op = true;
  while (x > 0) {
    if (x % 2 == 1) {
      if (first_op) {
        ckks_add(ctx, c2, c2, &ctx->caux[0]);
        first_op = false;
      } else {
        ckks_mul(ctx, c2, c2, &ctx->caux[0]);
      }
    }
    x >>= 1;
    ckks_square(ctx, &ctx->caux[0], &ctx->caux[0]);
  }
}

cipher_t* ckks_power(CKKSContext* ctx, cipher_t* c1, uint64_t x) {
  cipher_t* c2 = fresh(ctx);
  ckks_power(ctx, c2, c1, x);
  return c2;
}

void ckks_eval_polynomial(CKKSContext* ctx, cipher_t* result, cipher_t* val, double* coeffs, int n) {  // :778-796
  if (val == result) throw std::runtime_error("ckks_eval_polynomial: result must not alias the argument");
  cipher_clear(ctx, result);
  result->slots = val->slots;
  for (int i = n - 1; i >= 0; --i) {
    ckks_mul(ctx, result, result, val);
    ckks_add(ctx, result, result, coeffs[i]);
  }
}

cipher_t* ckks_eval_polynomial(CKKSContext* ctx, cipher_t* x, double* coeffs, int n) {
  cipher_t* c2 = fresh(ctx);
  ckks_eval_polynomial(ctx, c2, x, coeffs, n);
  return c2;
}

static void series(CKKSContext* ctx, cipher_t* result, cipher_t* ct, const std::string& name, int n) {
  ckks_eval_polynomial(ctx, result, ct, ctx->maclaurin_coeffs[name], n);
}

void ckks_exp(CKKSContext* ctx, cipher_t* result, cipher_t* ct) { series(ctx, result, ct, EXPONENT, 11); }
void ckks_sin(CKKSContext* ctx, cipher_t* result, cipher_t* ct) { series(ctx, result, ct, SIN, 12); }
void ckks_cos(CKKSContext* ctx, cipher_t* result, cipher_t* ct) { series(ctx, result, ct, COS, 12); }
void ckks_sigmoid(CKKSContext* ctx, cipher_t* result, cipher_t* ct) { series(ctx, result, ct, SIGMOID, 8); }

#define AOA_RETURNING(fn)                            \
  cipher_t* fn(CKKSContext* ctx, cipher_t* ct) {     \
    cipher_t* result = fresh(ctx);                   \
    fn(ctx, result, ct);                             \
    return result;                                   \
  }
AOA_RETURNING(ckks_exp)
AOA_RETURNING(ckks_sin)
AOA_RETURNING(ckks_cos)
AOA_RETURNING(ckks_sigmoid)
AOA_RETURNING(ckks_log1minus)
AOA_RETURNING(ckks_conjugate)
#undef AOA_RETURNING

void ckks_log1minus(CKKSContext* ctx, cipher_t* result, cipher_t* ct) {  // :923-938
  cipher_copy(ctx, result, ct);
  ckks_add(ctx, result, result, -1);
  ckks_mul(ctx, result, result, -1);
  ckks_log(ctx, result, result);
}

void ckks_log(CKKSContext* ctx, cipher_t* result, cipher_t* ct, int steps) {  // :1150-1174
  ckks_add(ctx, &ctx->caux[0], ct, -1);
  cipher_copy(ctx, &ctx->caux[1], &ctx->caux[0]);
  cipher_clear(ctx, result);
  result->slots = ct->slots;
  for (int i = 0; i < steps; ++i) {
    ckks_mul(ctx, &ctx->caux[2], &ctx->caux[1], 1.0 / (i + 1));
    if (i % 2 == 0)
      ckks_add(ctx, result, result, &ctx->caux[2]);
    else
      ckks_sub(ctx, result, result, &ctx->caux[2]);
    if (i < steps - 1) ckks_mul(ctx, &ctx->caux[1], &ctx->caux[1], &ctx->caux[0]);
  }
}

cipher_t* ckks_log(CKKSContext* ctx, cipher_t* ct, int steps) {
  cipher_t* result = fresh(ctx);
  ckks_log(ctx, result, ct, steps);
  return result;
}

void ckks_inverse(CKKSContext* ctx, cipher_t* result, cipher_t* ct, int steps) {  // :1101-1134
  ckks_add(ctx, &ctx->caux[0], ct, -1);
  ckks_mul(ctx, &ctx->caux[0], &ctx->caux[0], -1);
  ckks_add(ctx, result, &ctx->caux[0], 1);
  for (int i = 1; i < steps; ++i) {
    ckks_square(ctx, &ctx->caux[0], &ctx->caux[0]);
    ckks_add(ctx, &ctx->caux[1], &ctx->caux[0], 1);
    ckks_mul(ctx, result, &ctx->caux[1], result);
  }
}

cipher_t* ckks_inverse(CKKSContext* ctx, cipher_t* ct, int steps) {
  cipher_t* result = fresh(ctx);
  ckks_inverse(ctx, result, ct, steps);
  return result;
}

// ckks.cpp:1011-1040: conjugate both polynomials in the NTT domain, then key
// switch c1 with the conjugation key; c0's conjugate is the addend.
void ckks_conjugate(CKKSContext* ctx, cipher_t* c2, cipher_t* c1) {
  if (!ctx->keys.cjk) throw std::runtime_error("no conjugation key: call ckks_keygen first");
  aoa_params* P = CUDAManager::params();
  const int k = (int)CUDAManager::RNSQPrimes.size();
  COMMONEngine::execute(ctx, &c1->c[0], FORWARD);
  COMMONEngine::execute(ctx, &c1->c[1], FORWARD);
  poly_t t[2], o[2];
  for (int y = 0; y < 2; ++y) {
    poly_init(ctx, &t[y]);
    poly_init(ctx, &o[y]);
    ok(aoa_conjugate(P, ctx->get_stream(), t[y].d_coefs, c1->c[y].d_coefs, k));
    t[y].state = TRANSSTATE;
    COMMONEngine::execute(ctx, &t[y], INVERSE);
  }
  ok(aoa_keyswitch(P, ctx->get_stream(), o[0].d_coefs, o[1].d_coefs, t[0].d_coefs, t[1].d_coefs,
                   ctx->keys.cjk->b.d_coefs, ctx->keys.cjk->a.d_coefs, c1->level, ctx->d_ws, ctx->ws_bytes));
  const int level = c1->level, slots = c1->slots;
  const int64_t scale = c1->scale;
  poly_init(ctx, &c2->c[0]);
  poly_init(ctx, &c2->c[1]);
  poly_copy(ctx, &c2->c[0], &o[0]);
  poly_copy(ctx, &c2->c[1], &o[1]);
  c2->c[0].state = c2->c[1].state = RNSSTATE;
  for (int y = 0; y < 2; ++y) {
    poly_free(ctx, &t[y]);
    poly_free(ctx, &o[y]);
  }
  c2->level = level;
  c2->slots = slots;
  c2->scale = scale;
}

void ckks_discard_higher_slots(CKKSContext* ctx, cipher_t* ct) {  // :1082-1088
  ckks_encrypt(ctx, &ctx->caux[0], 1.0, 1, ct->slots - 1);
  ckks_mul(ctx, ct, ct, &ctx->caux[0]);
}

void ckks_discard_slots_except(CKKSContext* ctx, cipher_t* ct, int idx) {  // :1090-1101
  if (idx < 0 || idx >= ct->slots) throw std::runtime_error("ckks_discard_slots_except: slot out of range");
  for (int i = 0; i < ct->slots; ++i) ctx->h_val[i] = {0, 0};
  ctx->h_val[idx] = {1, 0};
  ckks_encrypt(ctx, &ctx->caux[0], ctx->h_val, ct->slots);
  ckks_mul(ctx, ct, ct, &ctx->caux[0]);
}

// ckks.cpp:1042-1066.  The reference takes the merged scale and level from
// cts[0] before discarding its other slots (a product that lowers the level
// and changes the scale) and then asserts that they equal those of every
// discarded input; here they are taken from the first discarded input, so
// the same sequence of operations runs and the assertion holds.
void ckks_merge(CKKSContext* ctx, cipher_t* c_merged, cipher_t* cts, int n) {
  if (n < 1 || n > CUDAManager::N) throw std::runtime_error("ckks_merge: bad number of ciphertexts");
  cipher_clear(ctx, c_merged);
  c_merged->slots = 0;
  for (int i = 0; i < n; ++i) {
    ckks_discard_slots_except(ctx, &cts[i], 0);
    if (i == 0) {
      c_merged->scale = cts[0].scale;
      c_merged->level = cts[0].level;
    }
    if (c_merged->scale != cts[i].scale || c_merged->level != cts[i].level)
      throw std::runtime_error("ckks_merge: ciphertexts must share scale and level");
    poly_add_add(ctx, &c_merged->c[0], &c_merged->c[0], &cts[i].c[0], &c_merged->c[1], &c_merged->c[1],
                 &cts[i].c[1]);
    ckks_rotate_right(ctx, c_merged, c_merged, 1);
    c_merged->slots++;
  }
}

cipher_t* ckks_merge(CKKSContext* ctx, cipher_t* cts, int n) {
  cipher_t* result = fresh(ctx);
  ckks_merge(ctx, result, cts, n);
  return result;
}
