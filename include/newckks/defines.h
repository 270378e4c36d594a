// newckks public types, kept name-compatible with the reference
// (aoa-ntt/include/newckks/defines.h:4-101) so callers switch by relinking.
#ifndef NEWCKKS_DEFINES_H
#define NEWCKKS_DEFINES_H

#include <stdint.h>

enum transform_directions { FORWARD, INVERSE };

enum engine_types { NTTTrans, DGTTrans };

enum supported_operations { ADDOP, SUBOP, MULOP, NEGATEOP, ADDADDOP, MULandADDOP };

// Domain of a polynomial's residues: coefficient (RNSSTATE) or NTT (TRANSSTATE).
enum poly_states { NONINITIALIZED, RNSSTATE, TRANSSTATE };

enum poly_bases { QBase, BBase, QBBase };

// Residue-major device array d_coefs[rid * N + cid]; Q residues first, then B.
struct polynomial {
  uint64_t* d_coefs = nullptr;
  poly_states state = NONINITIALIZED;
  poly_bases base = QBase;
};
typedef polynomial poly_t;

typedef struct {
  poly_t s;
} SecretKey;

typedef struct {
  poly_t a;
  poly_t b;
} SwitchKey;

typedef SwitchKey PublicKey;
typedef SwitchKey EvaluationKey;
typedef SwitchKey RotationKey;
typedef SwitchKey ConjugationKey;

#endif
