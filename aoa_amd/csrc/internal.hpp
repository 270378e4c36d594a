// Definitions shared by the C-ABI translation units (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/aoa.h"
#include "kernels.hpp"
#include "params.hpp"

struct aoa_params {
  aoa::HostParams host;
  aoa::DevParams dev;
  void* dmem = nullptr;
};

namespace aoa {

// Records the message of a failed call for aoa_last_error and returns code.
int set_error(int code, const std::string& msg);
// AOA_OK, or AOA_ERR_HIP with the pending launch error.
int check_launch(const char* what);
// 5^rot mod 2N: the Galois element of a left rotation by rot slots.
uint64_t galois_of(int rot_slots, int N);

}  // namespace aoa
