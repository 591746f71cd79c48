// Internal helpers shared by the gpufem_amd translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>
#include <unordered_map>

#include "../../include/gpufem_amd.h"

struct gf_context_s {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  int num_cus = 0;
  // pinned host slots for device-to-host scalars (GMRES estimates); solves on
  // one context are sequential, so every solve may use them
  double* pinned = nullptr;
  // Caching contexts (serving workers) keep freed device blocks for reuse:
  // hipFree waits for the whole device, which would serialise the streams of
  // concurrent requests.  Blocks are reused only on this context's stream, so
  // stream order protects them.
  bool caching = false;
  std::multimap<size_t, void*> free_blocks;
  std::unordered_map<void*, size_t> block_size;
  size_t cached_bytes = 0;
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

// Creates a context; worker contexts (serving) get a non-blocking stream and
// cache freed device blocks.
int context_create(int device, bool worker, gf_context* out);

// Device allocation through the context (plain hipMalloc / hipFree unless the
// context caches).  Cached sizes are rounded up to a quarter of their power
// of two, and a cached block is reused for a request up to 1/3 smaller.
constexpr size_t kCacheLimit = size_t(8) << 30;  // bytes kept per context
inline size_t cache_class(size_t bytes) {
  if (bytes <= 4096) return 4096;
  size_t p = 1;
  while (p * 2 <= bytes) p *= 2;
  const size_t q = p / 4;
  return (bytes + q - 1) / q * q;
}
inline int dmalloc(gf_context ctx, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 1;
  if (!ctx || !ctx->caching)
    return check_hip(ctx, hipMalloc(p, bytes), "hipMalloc");
  const size_t sz = cache_class(bytes);
  auto it = ctx->free_blocks.lower_bound(sz);
  if (it != ctx->free_blocks.end() && it->first <= sz + sz / 2) {
    *p = it->second;
    ctx->cached_bytes -= it->first;
    ctx->block_size[*p] = it->first;
    ctx->free_blocks.erase(it);
    return GF_OK;
  }
  hipError_t e = hipMalloc(p, sz);
  if (e == hipErrorOutOfMemory && !ctx->free_blocks.empty()) {
    // give the cache back and try once more
    hipGetLastError();
    for (auto& kv : ctx->free_blocks) hipFree(kv.second);
    ctx->free_blocks.clear();
    ctx->cached_bytes = 0;
    e = hipMalloc(p, sz);
  }
  if (e != hipSuccess) return check_hip(ctx, e, "hipMalloc");
  ctx->block_size[*p] = sz;
  return GF_OK;
}
inline void dfree(gf_context ctx, void* p) {
  if (!p) return;
  if (!ctx || !ctx->caching) {
    hipFree(p);
    return;
  }
  auto it = ctx->block_size.find(p);
  if (it == ctx->block_size.end()) {
    hipFree(p);
    return;
  }
  const size_t sz = it->second;
  ctx->block_size.erase(it);
  if (ctx->cached_bytes + sz > kCacheLimit) {
    hipFree(p);
    return;
  }
  ctx->free_blocks.emplace(sz, p);
  ctx->cached_bytes += sz;
}
// frees every cached block (context teardown)
inline void release_cache(gf_context ctx) {
  for (auto& kv : ctx->free_blocks) hipFree(kv.second);
  ctx->free_blocks.clear();
  ctx->cached_bytes = 0;
}

}  // namespace gf

#define GF_TRY(expr)             \
  do {                           \
    int _st = (expr);            \
    if (_st != GF_OK) return _st; \
  } while (0)

#define GF_HIP(ctx, call) GF_TRY(gf::check_hip((ctx), (call), #call))
