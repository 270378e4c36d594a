// Keys (aoa-ntt/include/newckks/ckks/ckkskeys.h:10-30).  All switching keys
// are QB polynomials kept in the NTT domain.
#ifndef NEWCKKS_CKKSKEYS_H
#define NEWCKKS_CKKSKEYS_H

#include <map>
#include <vector>

#include "../tool/context.h"
#include "../defines.h"

typedef struct {
  PublicKey* pk = nullptr;
  EvaluationKey* evk = nullptr;
  std::map<int, RotationKey*> rtk_left;
  std::map<int, RotationKey*> rtk_right;
  ConjugationKey* cjk = nullptr;
  SecretKey* sk = nullptr;
} CKKSKeychain;

SecretKey* ckks_new_sk(Context* ctx);
void ckks_free_sk(Context* ctx, SecretKey* sk);
PublicKey* ckks_new_pk(Context* ctx, SecretKey* sk);
// Switching key from s1 to s2: b = rho(s1) + e - a*s2, a uniform (QB).
EvaluationKey* ckks_new_evk(Context* ctx, poly_t* s1, poly_t* s2);
EvaluationKey* ckks_new_mtk(Context* ctx, SecretKey* sk);
std::map<int, RotationKey*> ckks_new_rtk_left(Context* ctx, SecretKey* sk);
std::map<int, RotationKey*> ckks_new_rtk_right(Context* ctx, SecretKey* sk);
// Rotation keys for the given steps only (the reference builds every power of two < N).
std::map<int, RotationKey*> ckks_new_rtk_left(Context* ctx, SecretKey* sk, const std::vector<int>& steps);
// Conjugation key: switching key from conj(s) to s (ckkskeys.cpp:183-193).
ConjugationKey* ckks_new_cjk(Context* ctx, SecretKey* sk);
void ckkskeychain_free(Context* ctx, CKKSKeychain* k);

#endif
