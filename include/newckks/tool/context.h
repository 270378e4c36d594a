// Context: one HIP stream per context, a QB-sized scratch buffer and the
// sampler (aoa-ntt/include/newckks/tool/context.h:12-49).
#ifndef NEWCKKS_CONTEXT_H
#define NEWCKKS_CONTEXT_H

#include <hip/hip_runtime.h>
#include <stdint.h>

class Sampler;

class Context {
 private:
  hipStream_t stream;
  Sampler* s;

 public:
  uint64_t* d_tmp;  // QB-sized scratch

  Context();
  virtual ~Context();
  hipStream_t get_stream() { return stream; }
  Sampler* get_sampler() { return s; }
};

#endif
