// Polynomial layer (aoa-ntt/include/newckks/arithmetic/poly_t.h:17-230):
// lazy device allocation on the context stream and the coefficient/NTT state
// machine that transforms operands in place before each operation.
#ifndef NEWCKKS_POLY_T_H
#define NEWCKKS_POLY_T_H

#include <string>

#include "../defines.h"
#include "../tool/context.h"

void poly_init(Context* ctx, poly_t* a, poly_bases b = QBase);
void poly_free(Context* ctx, poly_t* a);
void poly_clear(Context* ctx, poly_t* a);
void poly_copy(Context* ctx, poly_t* b, poly_t* a);
void poly_copy_to_device(Context* ctx, poly_t* a, uint64_t* h_coefs);
// Host copy of every residue in the coefficient domain (malloc'd; caller frees).
uint64_t* poly_copy_to_host(Context* ctx, poly_t* a);
bool poly_are_equal(Context* ctx, poly_t* a, poly_t* b);

void poly_add(Context* ctx, poly_t* c, poly_t* a, poly_t* b);
void poly_sub(Context* ctx, poly_t* c, poly_t* a, poly_t* b);
void poly_mul(Context* ctx, poly_t* c, poly_t* a, poly_t* b);
void poly_add_add(Context* ctx, poly_t* c1, poly_t* a1, poly_t* b1, poly_t* c2, poly_t* a2, poly_t* b2);
void poly_mul_add(Context* ctx, poly_t* d, poly_t* a, poly_t* b, poly_t* c);
// The reference's dr2 (poly_t.cpp:302-342) on QB-base polynomials: the four
// operands are brought to the NTT domain, then CUDAManager::execute_dr2.
void poly_dr2(Context* ctx, poly_t* ct21, poly_t* ct22, poly_t* ct23, poly_t* ct01, poly_t* ct02, poly_t* ct11,
              poly_t* ct12);
void poly_negate(Context* ctx, poly_t* a);
// By a 64-bit scalar used as given (poly_t.cpp:157-235): add/sub in the NTT
// domain (the scalar lands on the constant coefficient), mul in the
// coefficient domain.
void poly_add(Context* ctx, poly_t* c, poly_t* a, uint64_t b);
void poly_sub(Context* ctx, poly_t* c, poly_t* a, uint64_t b);
void poly_mul(Context* ctx, poly_t* c, poly_t* a, uint64_t b);
void poly_rho(Context* ctx, poly_t* b, poly_t* a);
void poly_modup(Context* ctx, poly_t* a, poly_t* b, int level);
void poly_moddown(Context* ctx, poly_t* a, poly_t* b, int level);
void poly_rescale(Context* ctx, poly_t* a, poly_t* b, int level = 1);

std::string poly_residue_to_string(Context* ctx, poly_t* a, int id = -1);
// Text form of every residue of the base, coefficient domain, each followed by
// a comma (poly_t.cpp:445-473).  poly_import returns a new Q-base polynomial.
std::string poly_export(Context* ctx, poly_t* p);
poly_t* poly_import(Context* ctx, std::string s);

#endif
