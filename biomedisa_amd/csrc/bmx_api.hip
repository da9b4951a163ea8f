// bmx_api.hip -- context, device memory and error reporting of libbmx_hip.
#include "bmx_internal.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>

#include <mutex>

// Scratch of bmx_walk_noalloc: one static buffer per GPU, held until the
// process ends.  In host code `scratch` only names the symbol; kernels get its
// device address, looked up once per device in bmx_open.
static __device__ __align__(256) unsigned char scratch[BMX_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {nullptr};
std::mutex g_scratch_mu;
}  // namespace

void *bmx_noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

// With `device` current: its address of the scratch buffer, once.
static int resolve_scratch(int device)
{
    BMX_CHECK(device < kMaxDevices, "bmx_open: device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return BMX_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
        return bmx_set_error(BMX_E_DEVICE, "bmx_open: hipGetSymbolAddress(scratch) failed: %s",
                             hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
    return BMX_OK;
}

static thread_local char g_err[512] = "";

int bmx_set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int bmx_set_device(bmx_ctx *ctx)
{
    BMX_CHECK(ctx != nullptr, "null context");
    BMX_HIP(hipSetDevice(ctx->device));
    return BMX_OK;
}

int bmx_aux_streams(bmx_ctx *ctx)
{
    if (ctx->aux[0]) return BMX_OK;
    hipStream_t st[2] = {};
    hipEvent_t ev[3] = {};
    hipError_t e = hipSuccess;
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipStreamCreateWithFlags(&st[k], hipStreamNonBlocking);
    for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipEventCreateWithFlags(&ev[k], hipEventDisableTiming);
    if (e != hipSuccess) {
        for (int k = 0; k < 2; ++k) if (st[k]) (void)hipStreamDestroy(st[k]);
        for (int k = 0; k < 3; ++k) if (ev[k]) (void)hipEventDestroy(ev[k]);
        return bmx_set_error(BMX_E_DEVICE, "bmx: stream creation failed: %s", hipGetErrorString(e));
    }
    for (int k = 0; k < 2; ++k) ctx->aux[k] = st[k];
    for (int k = 0; k < 3; ++k) ctx->aux_ev[k] = ev[k];
    return BMX_OK;
}

extern "C" {

const char *bmx_last_error(void) { return g_err; }

int bmx_version(void) { return 1; }

int bmx_open(int device, bmx_ctx **out)
{
    BMX_CHECK(out != nullptr, "bmx_open: out is null");
    *out = nullptr;
    int n = 0;
    BMX_HIP(hipGetDeviceCount(&n));
    BMX_CHECK(device >= 0 && device < n, "bmx_open: device %d not in [0,%d)", device, n);
    BMX_HIP(hipSetDevice(device));
    const int rs = resolve_scratch(device);
    if (rs) return rs;
    bmx_ctx *c = (bmx_ctx *)calloc(1, sizeof(bmx_ctx));
    if (!c) return bmx_set_error(BMX_E_OOM, "bmx_open: host allocation failed");
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e == hipSuccess) e = hipMalloc((void **)&c->d_counter, 16 * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&c->d_small, kSmallBytes);
    if (e != hipSuccess) {
        free(c);
        return bmx_set_error(BMX_E_DEVICE, "bmx_open: %s", hipGetErrorString(e));
    }
    *out = c;
    return BMX_OK;
}

int bmx_close(bmx_ctx *ctx)
{
    if (!ctx) return BMX_OK;
    BMX_HIP(hipSetDevice(ctx->device));
    BMX_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->aux[0]) {
        for (int k = 0; k < 2; ++k) {
            BMX_HIP(hipStreamSynchronize(ctx->aux[k]));
            BMX_HIP(hipStreamDestroy(ctx->aux[k]));
        }
        for (int k = 0; k < 3; ++k) BMX_HIP(hipEventDestroy(ctx->aux_ev[k]));
    }
    if (ctx->d_block_counts) BMX_HIP(hipFree(ctx->d_block_counts));
    if (ctx->d_rec) BMX_HIP(hipFree(ctx->d_rec));
    if (ctx->d_state) BMX_HIP(hipFree(ctx->d_state));
    if (ctx->d_keys) BMX_HIP(hipFree(ctx->d_keys));
    if (ctx->d_buckets) BMX_HIP(hipFree(ctx->d_buckets));
    if (ctx->d_scan) BMX_HIP(hipFree(ctx->d_scan));
    if (ctx->d_vols) BMX_HIP(hipFree(ctx->d_vols));
    if (ctx->d_wseeds) BMX_HIP(hipFree(ctx->d_wseeds));
    if (ctx->d_netpart) BMX_HIP(hipFree(ctx->d_netpart));
    if (ctx->seg_ev) {
        for (int k = 0; k < 2 * 1024; ++k) BMX_HIP(hipEventDestroy(ctx->seg_ev[k]));
        free(ctx->seg_ev);
    }
    BMX_HIP(hipFree(ctx->d_counter));
    BMX_HIP(hipFree(ctx->d_small));
    BMX_HIP(hipEventDestroy(ctx->ev0));
    BMX_HIP(hipEventDestroy(ctx->ev1));
    BMX_HIP(hipStreamDestroy(ctx->stream));
    free(ctx);
    return BMX_OK;
}

int bmx_sync(bmx_ctx *ctx)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_HIP(hipStreamSynchronize(ctx->stream));
    return BMX_OK;
}

int bmx_malloc(bmx_ctx *ctx, size_t bytes, void **dptr)
{
    BMX_CHECK(dptr != nullptr, "bmx_malloc: dptr is null");
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    *dptr = nullptr;
    if (bytes == 0) bytes = 1;
    hipError_t e = hipMalloc(dptr, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return bmx_set_error(e == hipErrorOutOfMemory ? BMX_E_OOM : BMX_E_DEVICE,
                             "bmx_malloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    return BMX_OK;
}

int bmx_free(bmx_ctx *ctx, void *dptr)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    if (!dptr) return BMX_OK;
    BMX_HIP(hipStreamSynchronize(ctx->stream));
    BMX_HIP(hipFree(dptr));
    return BMX_OK;
}

int bmx_upload(bmx_ctx *ctx, void *dst_dev, const void *src_host, size_t bytes)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    if (bytes == 0) return BMX_OK;
    BMX_CHECK(dst_dev && src_host, "bmx_upload: null pointer");
    BMX_HIP(hipMemcpyAsync(dst_dev, src_host, bytes, hipMemcpyHostToDevice, ctx->stream));
    BMX_HIP(hipStreamSynchronize(ctx->stream));
    return BMX_OK;
}

int bmx_download(bmx_ctx *ctx, void *dst_host, const void *src_dev, size_t bytes)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    if (bytes == 0) return BMX_OK;
    BMX_CHECK(dst_host && src_dev, "bmx_download: null pointer");
    BMX_HIP(hipMemcpyAsync(dst_host, src_dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    BMX_HIP(hipStreamSynchronize(ctx->stream));
    return BMX_OK;
}

int bmx_copy(bmx_ctx *ctx, void *dst_dev, const void *src_dev, size_t bytes)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    if (bytes == 0) return BMX_OK;
    BMX_CHECK(dst_dev && src_dev, "bmx_copy: null pointer");
    BMX_HIP(hipMemcpyAsync(dst_dev, src_dev, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return BMX_OK;
}

int bmx_zero(bmx_ctx *ctx, void *dst_dev, size_t bytes)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    if (bytes == 0) return BMX_OK;
    BMX_CHECK(dst_dev, "bmx_zero: null pointer");
    BMX_HIP(hipMemsetAsync(dst_dev, 0, bytes, ctx->stream));
    return BMX_OK;
}

float bmx_last_walk_ms(bmx_ctx *ctx) { return ctx ? ctx->last_walk_ms : -1.0f; }

}  // extern "C"

// VALU calibration: `threads` threads (256 per block) each run `iters` x 16
// steps of 8 independent fused multiply-add chains, i.e. threads*iters*128/64
// wave-level v_fma_f32 instructions with every lane active, so that a VALU
// instruction counter read over this kernel gives the rate of a full-rate op.
__global__ __launch_bounds__(256) void k_valu_cal(float *__restrict__ out, int iters, float b, float c)
{
    float a0 = (float)threadIdx.x, a1 = a0 + 1.f, a2 = a0 + 2.f, a3 = a0 + 3.f;
    float a4 = a0 + 4.f, a5 = a0 + 5.f, a6 = a0 + 6.f, a7 = a0 + 7.f;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            a0 = __builtin_fmaf(a0, b, c); a1 = __builtin_fmaf(a1, b, c);
            a2 = __builtin_fmaf(a2, b, c); a3 = __builtin_fmaf(a3, b, c);
            a4 = __builtin_fmaf(a4, b, c); a5 = __builtin_fmaf(a5, b, c);
            a6 = __builtin_fmaf(a6, b, c); a7 = __builtin_fmaf(a7, b, c);
        }
    }
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7;
}

extern "C" int bmx_valu_calibrate(bmx_ctx *ctx, float *out, int64_t threads, int iters, float *ms)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(out && ms, "bmx_valu_calibrate: null pointer");
    BMX_CHECK(threads >= 256 && threads % 256 == 0 && threads / 256 < ((int64_t)1 << 31) && iters >= 1,
              "bmx_valu_calibrate: threads must be a positive multiple of 256, iters >= 1");
    BMX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    k_valu_cal<<<(unsigned)(threads / 256), 256, 0, ctx->stream>>>(out, iters, 0.999f, 0.001f);
    BMX_LAUNCH_CHECK();
    BMX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    BMX_HIP(hipEventSynchronize(ctx->ev1));
    BMX_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
    return BMX_OK;
}
