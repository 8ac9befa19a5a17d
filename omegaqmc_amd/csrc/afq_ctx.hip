// Context lifecycle, error reporting and scratch management of the C ABI.
#include <stdarg.h>
#include <string.h>

#include "afq_common.h"

namespace afq {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  snprintf(g_err, sizeof(g_err), "HIP error %d (%s) in %s", (int)e, hipGetErrorString(e), what);
  return e == hipErrorOutOfMemory ? AFQ_ENOMEM : AFQ_EHIP;
}

void* scratch(afq_ctx* ctx, size_t bytes) {
  bytes = align256(bytes);
  if (ctx->scratch_bytes >= bytes) return ctx->scratch;
  // Growing the scratch must not free memory a queued kernel still reads.
  if (ctx->scratch) {
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) return nullptr;
    (void)hipFree(ctx->scratch);
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
  }
  if (hipMalloc(&ctx->scratch, bytes) != hipSuccess) {
    ctx->scratch = nullptr;
    set_error("hipMalloc of %zu scratch bytes failed", bytes);
    return nullptr;
  }
  ctx->scratch_bytes = bytes;
  return ctx->scratch;
}

size_t max_lds_bytes() {
  static size_t cached = 0;
  if (cached) return cached;
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 65536;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
    v = 65536;
  cached = (size_t)v;
  return cached;
}

int num_cus() {
  static int cached = 0;
  if (cached) return cached;
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
    v = 256;
  cached = v;
  return cached;
}

}  // namespace afq

extern "C" {

int afq_version(void) { return 10000; }

const char* afq_last_error(void) { return afq::g_err; }

int afq_ctx_create(afq_ctx** out, int device, void* stream) {
  AFQ_REQUIRE(out != nullptr, "afq_ctx_create: out is NULL");
  int ndev = 0;
  AFQ_CHECK_HIP(hipGetDeviceCount(&ndev));
  AFQ_REQUIRE(device >= 0 && device < ndev, "afq_ctx_create: device %d of %d", device, ndev);
  AFQ_CHECK_HIP(hipSetDevice(device));
  double* comb_static = nullptr;
  int rc = afq::comb_static_scratch(device, &comb_static);
  if (rc) return rc;
  afq_ctx* c = new afq_ctx();
  c->device = device;
  c->stream = (hipStream_t)stream;
  c->comb_static = comb_static;
  *out = c;
  return AFQ_OK;
}

int afq_ctx_destroy(afq_ctx* ctx) {
  if (!ctx) return AFQ_OK;
  if (ctx->scratch) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->scratch);
  }
  delete ctx;
  return AFQ_OK;
}

int afq_ctx_set_stream(afq_ctx* ctx, void* stream) {
  AFQ_REQUIRE(ctx != nullptr, "afq_ctx_set_stream: ctx is NULL");
  ctx->stream = (hipStream_t)stream;
  return AFQ_OK;
}

int afq_ctx_synchronize(afq_ctx* ctx) {
  AFQ_REQUIRE(ctx != nullptr, "afq_ctx_synchronize: ctx is NULL");
  AFQ_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return AFQ_OK;
}

int afq_ctx_reserve(afq_ctx* ctx, size_t bytes) {
  AFQ_REQUIRE(ctx != nullptr, "afq_ctx_reserve: ctx is NULL");
  if (!afq::scratch(ctx, bytes)) return AFQ_ENOMEM;
  return AFQ_OK;
}

}  // extern "C"
