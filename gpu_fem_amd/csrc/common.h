// Internal helpers shared by the gpufem_amd translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/gpufem_amd.h"

struct gf_context_s {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  int num_cus = 0;
};

namespace gf {

inline int set_error(gf_context ctx, int code, const char* fmt, ...) {
  if (ctx) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->last_error = buf;
  }
  return code;
}

inline int check_hip(gf_context ctx, hipError_t e, const char* what) {
  if (e == hipSuccess) return GF_OK;
  return set_error(ctx, e == hipErrorOutOfMemory ? GF_ERR_NOMEM : GF_ERR_DEVICE,
                   "%s: %s", what, hipGetErrorString(e));
}

// Check the launch status of the last kernel on this thread.
inline int check_launch(gf_context ctx, const char* what) {
  return check_hip(ctx, hipGetLastError(), what);
}

inline unsigned div_up(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace gf

#define GF_TRY(expr)             \
  do {                           \
    int _st = (expr);            \
    if (_st != GF_OK) return _st; \
  } while (0)

#define GF_HIP(ctx, call) GF_TRY(gf::check_hip((ctx), (call), #call))
