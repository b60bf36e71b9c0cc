// fot_runtime.hip -- status codes, device memory, streams, timers and the
// out -> in copy of libfot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fot_internal.hpp"
#include "fot_hip.hpp"

namespace {
thread_local std::string g_message;
}

namespace fot {
fot_status fail(fot_status status, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_message = buf;
  return status;
}
} // namespace fot

using fot::fail;


extern "C" const char *fot_status_string(fot_status s) {
  switch (s) {
  case FOT_OK: return "ok";
  case FOT_ERR_INVALID_ARGUMENT: return "invalid argument";
  case FOT_ERR_HIP: return "HIP runtime error";
  case FOT_ERR_OUT_OF_MEMORY: return "out of memory";
  case FOT_ERR_COLOUR_OVERFLOW: return "colour overflow";
  case FOT_ERR_UNSUPPORTED: return "unsupported";
  case FOT_ERR_INVALID_INPUT: return "invalid input";
  case FOT_ERR_NO_DEVICE: return "no device";
  case FOT_ERR_TIMEOUT: return "device wait timed out";
  }
  return "unknown status";
}

extern "C" const char *fot_last_error_message(void) { return g_message.c_str(); }
extern "C" int fot_version(void) { return 1; }

static fot_status fot_device_count_impl(int *count) {
  if (!count) return fail(FOT_ERR_INVALID_ARGUMENT, "null count");
  *count = 0;
  const hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return fail(FOT_ERR_NO_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
  }
  return FOT_OK;
}
extern "C" fot_status fot_device_count(int *count) {
  return fot::guarded("fot_device_count", [&] { return fot_device_count_impl(count); });
}

static fot_status fot_set_device_impl(int device) {
  FOT_HIP(hipSetDevice(device));
  return FOT_OK;
}
extern "C" fot_status fot_set_device(int device) {
  return fot::guarded("fot_set_device", [&] { return fot_set_device_impl(device); });
}

static fot_status fot_malloc_impl(void **p, size_t bytes) {
  if (!p) return fail(FOT_ERR_INVALID_ARGUMENT, "null pointer");
  *p = nullptr;
  if (bytes == 0) return FOT_OK;
  FOT_HIP(hipMalloc(p, bytes));
  return FOT_OK;
}
extern "C" fot_status fot_malloc(void **p, size_t bytes) {
  return fot::guarded("fot_malloc", [&] { return fot_malloc_impl(p, bytes); });
}

static fot_status fot_free_impl(void *p) {
  if (p) FOT_HIP(hipFree(p));
  return FOT_OK;
}
extern "C" fot_status fot_free(void *p) {
  return fot::guarded("fot_free", [&] { return fot_free_impl(p); });
}

static fot_status fot_memcpy_htod_impl(void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return FOT_OK;
  if (!dst || !src) return fail(FOT_ERR_INVALID_ARGUMENT, "null pointer");
  FOT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return FOT_OK;
}
extern "C" fot_status fot_memcpy_htod(void *dst, const void *src, size_t bytes) {
  return fot::guarded("fot_memcpy_htod", [&] { return fot_memcpy_htod_impl(dst, src, bytes); });
}

static fot_status fot_memcpy_dtoh_impl(void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return FOT_OK;
  if (!dst || !src) return fail(FOT_ERR_INVALID_ARGUMENT, "null pointer");
  FOT_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return FOT_OK;
}
extern "C" fot_status fot_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  return fot::guarded("fot_memcpy_dtoh", [&] { return fot_memcpy_dtoh_impl(dst, src, bytes); });
}

static fot_status fot_memcpy_dtod_impl(void *dst, const void *src, size_t bytes,
                                      void *stream) {
  if (bytes == 0) return FOT_OK;
  if (!dst || !src) return fail(FOT_ERR_INVALID_ARGUMENT, "null pointer");
  FOT_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice,
                         static_cast<hipStream_t>(stream)));
  return FOT_OK;
}
extern "C" fot_status fot_memcpy_dtod(void *dst, const void *src, size_t bytes,
                                      void *stream) {
  return fot::guarded("fot_memcpy_dtod", [&] {
    return fot_memcpy_dtod_impl(dst, src, bytes, stream);
  });
}

static fot_status fot_device_synchronize_impl(void) {
  FOT_HIP(hipDeviceSynchronize());
  return FOT_OK;
}
extern "C" fot_status fot_device_synchronize(void) {
  return fot::guarded("fot_device_synchronize", [&] { return fot_device_synchronize_impl(); });
}

static fot_status fot_stream_create_impl(void **stream) {
  if (!stream) return fail(FOT_ERR_INVALID_ARGUMENT, "null stream");
  hipStream_t s;
  // default priority: streams of other priorities are not used anywhere
  FOT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return FOT_OK;
}
extern "C" fot_status fot_stream_create(void **stream) {
  return fot::guarded("fot_stream_create", [&] { return fot_stream_create_impl(stream); });
}

static fot_status fot_stream_destroy_impl(void *stream) {
  if (stream) FOT_HIP(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return FOT_OK;
}
extern "C" fot_status fot_stream_destroy(void *stream) {
  return fot::guarded("fot_stream_destroy", [&] { return fot_stream_destroy_impl(stream); });
}

static fot_status fot_stream_synchronize_impl(void *stream) {
  FOT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return FOT_OK;
}
extern "C" fot_status fot_stream_synchronize(void *stream) {
  return fot::guarded("fot_stream_synchronize", [&] {
    return fot_stream_synchronize_impl(stream);
  });
}

// copyKernel (data_t.cu:76-93), streaming form: each workgroup moves tiles
// of kCopyUnroll x 256 16-byte vectors, all loads of a tile issued before
// its stores (four 16-byte loads in flight per lane), with nontemporal hints
// so a copy larger than the caches does not evict what the loop keeps
// there.  Lanes of one load instruction touch consecutive vectors.  The
// scalar tail (bytes % 16) is copied by workgroup 0.
constexpr unsigned kCopyThreads = 256;
constexpr unsigned kCopyUnroll = 4;
constexpr size_t kCopyTile = size_t(kCopyThreads) * kCopyUnroll;
// the nontemporal builtins take native vector types, not HIP's uint4 struct
typedef unsigned int copy_vec __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kCopyThreads) void fot_copy_kernel(
    const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16,
    const unsigned char *__restrict__ tail_src, unsigned char *__restrict__ tail_dst,
    unsigned tail) {
  const size_t full = n16 / kCopyTile;
  for (size_t t = blockIdx.x; t < full; t += gridDim.x) {
    const size_t base = t * kCopyTile + threadIdx.x;
    const copy_vec *s = reinterpret_cast<const copy_vec *>(src) + base;
    copy_vec *d = reinterpret_cast<copy_vec *>(dst) + base;
    copy_vec v[kCopyUnroll];
#pragma unroll
    for (unsigned k = 0; k < kCopyUnroll; ++k) v[k] = __builtin_nontemporal_load(s + k * kCopyThreads);
#pragma unroll
    for (unsigned k = 0; k < kCopyUnroll; ++k) __builtin_nontemporal_store(v[k], d + k * kCopyThreads);
  }
  // the vectors after the last whole tile: fewer than kCopyTile
  if (blockIdx.x == gridDim.x - 1) {
    for (size_t i = full * kCopyTile + threadIdx.x; i < n16; i += kCopyThreads)
      dst[i] = src[i];
  }
  if (blockIdx.x == 0 && threadIdx.x < tail) tail_dst[threadIdx.x] = tail_src[threadIdx.x];
}

// Workgroups of one copy launch: enough for every tile up to 8 per CU of
// the 256 CUs at once, then the tiles are taken grid-stride.
static unsigned copy_grid(size_t n16) {
  const size_t tiles = n16 / kCopyTile;
  return unsigned(std::max<size_t>(1, std::min<size_t>(tiles, 2048)));
}

static fot_status fot_copy_impl(void *dst, const void *src, size_t bytes, void *stream) {
  if (bytes == 0) return FOT_OK;
  if (!dst || !src) return fail(FOT_ERR_INVALID_ARGUMENT, "null pointer");
  if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15u)
    return fot_memcpy_dtod(dst, src, bytes, stream);
  const size_t n16 = bytes / 16;
  const unsigned tail = unsigned(bytes % 16);
  hipLaunchKernelGGL(fot_copy_kernel, dim3(copy_grid(n16)), dim3(kCopyThreads), 0,
                     static_cast<hipStream_t>(stream), static_cast<const uint4 *>(src),
                     static_cast<uint4 *>(dst), n16,
                     static_cast<const unsigned char *>(src) + n16 * 16,
                     static_cast<unsigned char *>(dst) + n16 * 16, tail);
  FOT_HIP(hipGetLastError());
  return FOT_OK;
}
extern "C" fot_status fot_copy(void *dst, const void *src, size_t bytes, void *stream) {
  return fot::guarded("fot_copy", [&] { return fot_copy_impl(dst, src, bytes, stream); });
}

struct fot_timer {
  hipEvent_t start = nullptr, stop = nullptr;
};

static fot_status fot_timer_create_impl(fot_timer **t) {
  if (!t) return fail(FOT_ERR_INVALID_ARGUMENT, "null timer");
  fot_timer *p = new fot_timer;
  hipError_t e = hipEventCreate(&p->start);
  if (e == hipSuccess) e = hipEventCreate(&p->stop);
  if (e != hipSuccess) {
    if (p->start) (void)hipEventDestroy(p->start);
    delete p;
    return fail(FOT_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(e));
  }
  *t = p;
  return FOT_OK;
}
extern "C" fot_status fot_timer_create(fot_timer **t) {
  return fot::guarded("fot_timer_create", [&] { return fot_timer_create_impl(t); });
}

static fot_status fot_timer_destroy_impl(fot_timer *t) {
  if (!t) return FOT_OK;
  (void)hipEventDestroy(t->start);
  (void)hipEventDestroy(t->stop);
  delete t;
  return FOT_OK;
}
extern "C" fot_status fot_timer_destroy(fot_timer *t) {
  return fot::guarded("fot_timer_destroy", [&] { return fot_timer_destroy_impl(t); });
}

static fot_status fot_timer_start_impl(fot_timer *t, void *stream) {
  if (!t) return fail(FOT_ERR_INVALID_ARGUMENT, "null timer");
  FOT_HIP(hipEventRecord(t->start, static_cast<hipStream_t>(stream)));
  return FOT_OK;
}
extern "C" fot_status fot_timer_start(fot_timer *t, void *stream) {
  return fot::guarded("fot_timer_start", [&] { return fot_timer_start_impl(t, stream); });
}

static fot_status fot_timer_stop_impl(fot_timer *t, void *stream) {
  if (!t) return fail(FOT_ERR_INVALID_ARGUMENT, "null timer");
  FOT_HIP(hipEventRecord(t->stop, static_cast<hipStream_t>(stream)));
  return FOT_OK;
}
extern "C" fot_status fot_timer_stop(fot_timer *t, void *stream) {
  return fot::guarded("fot_timer_stop", [&] { return fot_timer_stop_impl(t, stream); });
}

static fot_status fot_timer_elapsed_ms_impl(fot_timer *t, float *ms) {
  if (!t || !ms) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  FOT_HIP(hipEventSynchronize(t->stop));
  FOT_HIP(hipEventElapsedTime(ms, t->start, t->stop));
  return FOT_OK;
}
extern "C" fot_status fot_timer_elapsed_ms(fot_timer *t, float *ms) {
  return fot::guarded("fot_timer_elapsed_ms", [&] { return fot_timer_elapsed_ms_impl(t, ms); });
}

// ---------------------------------------------------------------------
// Context: a device, one stream, and the buffers allocated through it.
struct fot_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<void *> bufs;
};

static fot_status fot_ctx_create_impl(fot_ctx **ctx, int device) {
  if (!ctx) return fail(FOT_ERR_INVALID_ARGUMENT, "null context");
  *ctx = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(FOT_ERR_NO_DEVICE, "no HIP device");
  if (device < 0 || device >= n)
    return fail(FOT_ERR_INVALID_ARGUMENT, "device %d of %d", device, n);
  FOT_HIP(hipSetDevice(device));
  fot_ctx *c = new (std::nothrow) fot_ctx;
  if (!c) return fail(FOT_ERR_OUT_OF_MEMORY, "fot_ctx_create");
  c->device = device;
  // default priority: streams of other priorities are not used anywhere
  const hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(FOT_ERR_HIP, "hipStreamCreateWithFlags: %s", hipGetErrorString(e));
  }
  *ctx = c;
  return FOT_OK;
}
extern "C" fot_status fot_ctx_create(fot_ctx **ctx, int device) {
  return fot::guarded("fot_ctx_create", [&] { return fot_ctx_create_impl(ctx, device); });
}

static fot_status fot_ctx_destroy_impl(fot_ctx *ctx) {
  if (!ctx) return FOT_OK;
  (void)hipSetDevice(ctx->device);
  const hipError_t e = hipStreamSynchronize(ctx->stream);
  for (void *p : ctx->bufs) (void)hipFree(p);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  if (e != hipSuccess) return fail(FOT_ERR_HIP, "context stream: %s", hipGetErrorString(e));
  return FOT_OK;
}
extern "C" fot_status fot_ctx_destroy(fot_ctx *ctx) {
  return fot::guarded("fot_ctx_destroy", [&] { return fot_ctx_destroy_impl(ctx); });
}

static fot_status fot_ctx_stream_impl(const fot_ctx *ctx, void **stream) {
  if (!ctx || !stream) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  *stream = ctx->stream;
  return FOT_OK;
}
extern "C" fot_status fot_ctx_stream(const fot_ctx *ctx, void **stream) {
  return fot::guarded("fot_ctx_stream", [&] { return fot_ctx_stream_impl(ctx, stream); });
}

static fot_status fot_ctx_synchronize_impl(fot_ctx *ctx) {
  if (!ctx) return fail(FOT_ERR_INVALID_ARGUMENT, "null context");
  FOT_HIP(hipStreamSynchronize(ctx->stream));
  return FOT_OK;
}
extern "C" fot_status fot_ctx_synchronize(fot_ctx *ctx) {
  return fot::guarded("fot_ctx_synchronize", [&] { return fot_ctx_synchronize_impl(ctx); });
}

static fot_status fot_buf_alloc_impl(fot_ctx *ctx, size_t bytes, void **device_ptr) {
  if (!ctx || !device_ptr) return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  *device_ptr = nullptr;
  FOT_HIP(hipSetDevice(ctx->device));
  void *p = nullptr;
  if (bytes) {
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(FOT_ERR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", bytes,
                                     hipGetErrorString(e));
    ctx->bufs.push_back(p);
  }
  *device_ptr = p;
  return FOT_OK;
}
extern "C" fot_status fot_buf_alloc(fot_ctx *ctx, size_t bytes, void **device_ptr) {
  return fot::guarded("fot_buf_alloc", [&] { return fot_buf_alloc_impl(ctx, bytes, device_ptr); });
}

static fot_status fot_buf_upload_impl(fot_ctx *ctx, void *device_dst, const void *host_src,
                                     size_t bytes) {
  if (!ctx || (bytes && (!device_dst || !host_src)))
    return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  if (!bytes) return FOT_OK;
  FOT_HIP(hipMemcpyAsync(device_dst, host_src, bytes, hipMemcpyHostToDevice, ctx->stream));
  FOT_HIP(hipStreamSynchronize(ctx->stream));
  return FOT_OK;
}
extern "C" fot_status fot_buf_upload(fot_ctx *ctx, void *device_dst, const void *host_src,
                                     size_t bytes) {
  return fot::guarded("fot_buf_upload", [&] {
    return fot_buf_upload_impl(ctx, device_dst, host_src, bytes);
  });
}

static fot_status fot_buf_download_impl(fot_ctx *ctx, void *host_dst, const void *device_src,
                                       size_t bytes) {
  if (!ctx || (bytes && (!host_dst || !device_src)))
    return fail(FOT_ERR_INVALID_ARGUMENT, "null argument");
  if (!bytes) return FOT_OK;
  FOT_HIP(hipMemcpyAsync(host_dst, device_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  FOT_HIP(hipStreamSynchronize(ctx->stream));
  return FOT_OK;
}
extern "C" fot_status fot_buf_download(fot_ctx *ctx, void *host_dst, const void *device_src,
                                       size_t bytes) {
  return fot::guarded("fot_buf_download", [&] {
    return fot_buf_download_impl(ctx, host_dst, device_src, bytes);
  });
}

static fot_status fot_buf_free_impl(fot_ctx *ctx, void *device_ptr) {
  if (!ctx) return fail(FOT_ERR_INVALID_ARGUMENT, "null context");
  if (!device_ptr) return FOT_OK;
  auto it = std::find(ctx->bufs.begin(), ctx->bufs.end(), device_ptr);
  if (it == ctx->bufs.end())
    return fail(FOT_ERR_INVALID_ARGUMENT, "buffer %p is not this context's", device_ptr);
  ctx->bufs.erase(it);
  // the stream may still use the buffer
  FOT_HIP(hipStreamSynchronize(ctx->stream));
  FOT_HIP(hipFree(device_ptr));
  return FOT_OK;
}
extern "C" fot_status fot_buf_free(fot_ctx *ctx, void *device_ptr) {
  return fot::guarded("fot_buf_free", [&] { return fot_buf_free_impl(ctx, device_ptr); });
}
