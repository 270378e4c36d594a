// Umbrella header of the newckks-compatible C++ API.
#ifndef NEWCKKS_H
#define NEWCKKS_H
#include "ckks/ckks.h"
#include "ckks/encoder.h"
#include "cuda/htrans/common.h"
#include "cuda/manager.h"
#include "cuda/sampler.h"
#endif
