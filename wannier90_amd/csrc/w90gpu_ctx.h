// The library context shared by the translation units of libw90gpu (the
// spread minimisation in w90gpu.hip, the disentanglement in w90dis.hip).
#pragma once

#include <string>

struct w90gpu_ctx {
  int device = 0;
  std::string err;
  char* scratch = nullptr;  // this device's copy of `scratch` (w90gpu_omega_noalloc)
};

// Sets the message w90gpu_last_error returns for ctx (for ctx == nullptr,
// the calling thread's message of calls without a context).
extern "C" void w90gpu_set_error(w90gpu_ctx* ctx, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
