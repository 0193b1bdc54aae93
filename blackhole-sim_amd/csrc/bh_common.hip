// Module identity, errors, device contexts and memory helpers of the bhgrrt C ABI.
//
// bh_core_version / bh_detect_arch replace core_version() and detect_arch() of
// the reference's PyO3 module (native/core/src/lib.rs:9-35).

#include "bh_internal.h"

#include <sys/utsname.h>

#include <cstring>

namespace bh {

static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

int ctx_check(bh_ctx *ctx) {
    if (!ctx) return set_error(BH_ERR_STATE, "context is NULL");
    BH_HIP(hipSetDevice(ctx->device));
    return BH_OK;
}

void begin_timing(bh_ctx *ctx) {
    ctx->timed_launches = 0;
    (void)hipEventRecord(ctx->ev_start, ctx->stream);
}

void end_timing(bh_ctx *ctx, int launches) {
    (void)hipEventRecord(ctx->ev_stop, ctx->stream);
    ctx->timed_launches = launches;
}

__global__ void fill_u32_kernel(unsigned int *dst, unsigned int v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

int fill_u32(bh_ctx *ctx, unsigned int *dst, unsigned int v, int n) {
    if (n <= 0) return BH_OK;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, dst, v, n);
    BH_HIP(hipGetLastError());
    return BH_OK;
}

static std::mutex g_default_mutex;
static bh_ctx *g_default = nullptr;
static void *g_staging = nullptr;
static size_t g_staging_bytes = 0;

std::mutex &default_ctx_mutex() { return g_default_mutex; }

bh_ctx *default_ctx(int *status) {
    if (!g_default) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) dev = 0;
        bh_ctx *c = nullptr;
        int st = bh_ctx_create(dev, &c);
        if (st != BH_OK) {
            *status = st;
            return nullptr;
        }
        g_default = c;
    }
    *status = ctx_check(g_default);
    return *status == BH_OK ? g_default : nullptr;
}

// Pinned, coherent host memory mapped into the device's address space: kernels of small host-buffer calls
// read their inputs from it and write their outputs to it directly, so such a call is one launch and one
// wait, with no copy-engine transfer.  Grow-only, like the device staging buffer.
static void *g_host_staging = nullptr;
static void *g_host_staging_dev = nullptr;
static size_t g_host_staging_bytes = 0;

int host_staging(size_t bytes, void **host, void **dev) {
    if (bytes > g_host_staging_bytes) {
        if (g_host_staging) (void)hipHostFree(g_host_staging);
        g_host_staging = g_host_staging_dev = nullptr;
        g_host_staging_bytes = 0;
        size_t want = bytes < (size_t(1) << 18) ? (size_t(1) << 18) : bytes;
        if (hipHostMalloc(&g_host_staging, want, hipHostMallocCoherent | hipHostMallocPortable | hipHostMallocMapped) !=
            hipSuccess) {
            g_host_staging = nullptr;
            return set_error(BH_ERR_MEMORY, "pinned host staging allocation of %zu bytes failed", want);
        }
        if (hipHostGetDevicePointer(&g_host_staging_dev, g_host_staging, 0) != hipSuccess) {
            (void)hipHostFree(g_host_staging);
            g_host_staging = g_host_staging_dev = nullptr;
            return set_error(BH_ERR_DEVICE, "pinned host staging is not mapped for the device");
        }
        g_host_staging_bytes = want;
    }
    *host = g_host_staging;
    *dev = g_host_staging_dev;
    return BH_OK;
}

int staging(size_t bytes, void **ptr) {
    if (bytes > g_staging_bytes) {
        if (g_staging) (void)hipFree(g_staging);
        g_staging = nullptr;
        g_staging_bytes = 0;
        size_t want = bytes < (size_t(1) << 20) ? (size_t(1) << 20) : bytes;
        if (hipMalloc(&g_staging, want) != hipSuccess)
            return set_error(BH_ERR_MEMORY, "device staging allocation of %zu bytes failed", want);
        g_staging_bytes = want;
    }
    *ptr = g_staging;
    return BH_OK;
}

}  // namespace bh

using namespace bh;

extern "C" {

const char *bh_core_version(void) { return "0.9.0"; }

const char *bh_last_error(void) { return g_last_error.c_str(); }

int bh_detect_arch(char *buf, size_t buflen) {
    // normalized_arch of lib.rs:9-17
    struct utsname u;
    const char *arch = "unknown";
    if (uname(&u) == 0) {
        if (!strcmp(u.machine, "x86_64") || !strcmp(u.machine, "amd64")) arch = "x86_64";
        else if (!strcmp(u.machine, "aarch64") || !strcmp(u.machine, "arm64")) arch = "arm64";
        else if (!strcmp(u.machine, "i386") || !strcmp(u.machine, "i686") || !strcmp(u.machine, "x86")) arch = "x86";
        else if (!strncmp(u.machine, "arm", 3)) arch = "arm";
    }
    if (!buf || buflen == 0) return set_error(BH_ERR_VALUE, "buffer is empty");
    int n = snprintf(buf, buflen, "arch=%s;os=linux;family=unix;pointer_width=%d", arch, int(sizeof(void *) * 8));
    if (n < 0 || size_t(n) >= buflen) return set_error(BH_ERR_VALUE, "buffer too small");
    return BH_OK;
}

int bh_device_count(int *count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return BH_OK;
}

int bh_device_name(int device, char *buf, size_t buflen, int *compute_units) {
    hipDeviceProp_t p;
    BH_HIP(hipGetDeviceProperties(&p, device));
    if (buf && buflen) snprintf(buf, buflen, "%s", p.gcnArchName);
    if (compute_units) *compute_units = p.multiProcessorCount;
    return BH_OK;
}

int bh_ctx_create(int device, bh_ctx **out) {
    if (!out) return set_error(BH_ERR_VALUE, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return set_error(BH_ERR_DEVICE, "no HIP device is available; the bhgrrt kernels need an MI355X (gfx950) GPU");
    if (device < 0 || device >= n) return set_error(BH_ERR_VALUE, "device %d out of range (%d devices)", device, n);
    BH_HIP(hipSetDevice(device));
    hipDeviceProp_t p;
    BH_HIP(hipGetDeviceProperties(&p, device));
    if (strncmp(p.gcnArchName, "gfx950", 6) != 0)
        return set_error(BH_ERR_DEVICE, "device %d is %s; this library is built for gfx950 only", device, p.gcnArchName);
    bh_ctx *c = new bh_ctx();
    c->device = device;
    c->cus = p.multiProcessorCount;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev_start);
    if (e == hipSuccess) e = hipEventCreate(&c->ev_stop);
    if (e == hipSuccess) e = hipMalloc(&c->d_counters, kCounterWords * sizeof(unsigned int));
    if (e != hipSuccess) {
        bh_ctx_destroy(c);
        return set_error(BH_ERR_DEVICE, "context setup failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return BH_OK;
}

int bh_ctx_destroy(bh_ctx *ctx) {
    if (!ctx) return BH_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &b : ctx->bricks) {
        if (b.coeffs) (void)hipFree(b.coeffs);
        if (b.grids) (void)hipFree(b.grids);
    }
    if (ctx->d_counters) (void)hipFree(ctx->d_counters);
    if (ctx->d_order) (void)hipFree(ctx->d_order);
    if (ctx->d_synchro) (void)hipFree(ctx->d_synchro);
    if (ctx->d_sampler_work) (void)hipFree(ctx->d_sampler_work);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BH_OK;
}

int bh_ctx_synchronize(bh_ctx *ctx) {
    int st = ctx_check(ctx);
    if (st) return st;
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return BH_OK;
}

int bh_ctx_last_kernel_ms(bh_ctx *ctx, double *ms) {
    int st = ctx_check(ctx);
    if (st) return st;
    if (ctx->timed_launches <= 0) return set_error(BH_ERR_STATE, "no timed launch recorded");
    BH_HIP(hipEventSynchronize(ctx->ev_stop));
    float t = 0.f;
    BH_HIP(hipEventElapsedTime(&t, ctx->ev_start, ctx->ev_stop));
    *ms = double(t) / ctx->timed_launches;
    return BH_OK;
}

int bh_device_alloc(bh_ctx *ctx, size_t bytes, void **ptr) {
    int st = ctx_check(ctx);
    if (st) return st;
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess)
        return set_error(BH_ERR_MEMORY, "hipMalloc of %zu bytes failed", bytes);
    return BH_OK;
}

int bh_device_free(bh_ctx *ctx, void *ptr) {
    int st = ctx_check(ctx);
    if (st) return st;
    BH_HIP(hipStreamSynchronize(ctx->stream));
    BH_HIP(hipFree(ptr));
    return BH_OK;
}

int bh_copy_to_device(bh_ctx *ctx, void *dst, const void *src, size_t bytes) {
    int st = ctx_check(ctx);
    if (st) return st;
    BH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return BH_OK;
}

int bh_copy_to_host(bh_ctx *ctx, void *dst, const void *src, size_t bytes) {
    int st = ctx_check(ctx);
    if (st) return st;
    BH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return BH_OK;
}

int bh_copy_device(bh_ctx *ctx, void *dst, const void *src, size_t bytes) {
    int st = ctx_check(ctx);
    if (st) return st;
    BH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return BH_OK;
}

}  // extern "C"
