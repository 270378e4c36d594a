// Transform dispatch (aoa-ntt/include/newckks/cuda/htrans/common.h:111-167).
#ifndef NEWCKKS_COMMON_H
#define NEWCKKS_COMMON_H

#include "../../defines.h"
#include "../../tool/context.h"

class COMMONEngine {
 public:
  // Forward / inverse NTT of every residue of the polynomial's base; no-op
  // if it is already in the requested domain.
  static void execute(Context* ctx, poly_t* data, const transform_directions direction);
};

#endif
