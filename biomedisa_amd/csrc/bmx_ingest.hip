// bmx_ingest.hip -- the host pre- and post-processing of smart_interpolation
// moved onto the device: label set and label boxes, image normalisation, the
// 1-voxel pad and the crop to the label box, the label-value -> label-index
// table, result placement and the Dice sanity check.
//
// The raw arrays the caller passed are uploaded once; every volume the walk
// needs is written from them on the device, and the results come back as
// whole (Z,Y,X) volumes ready to hand out.
#include "bmx_internal.h"

#include <limits.h>
#include <string.h>

namespace {

constexpr int kClasses = BMX_LABEL_CLASSES;  // 0..255, negative, > 255
constexpr int kNeg = 256, kBig = 257;

__device__ __forceinline__ void wave_sum_add(unsigned long long *dst, uint64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

template <typename T> __device__ __forceinline__ int label_class(T v)
{
    if ((long long)v < 0) return kNeg;
    if ((long long)v > 255) return kBig;
    return (int)v;
}

// Per-block counts and boxes of the label classes in LDS, merged into the
// global arrays at the end.  Only non-zero voxels do work.
template <typename T>
__device__ __forceinline__ void stats_add(T v, uint32_t i, uint32_t X, uint32_t sz, uint32_t *cnt, int *box)
{
    if (v == (T)0) return;
    const int c = label_class(v);
    const uint32_t z = i / sz, r = i - z * sz, y = r / X, x = r - y * X;
    atomicAdd(&cnt[c], 1u);
    atomicMin(&box[6 * c + 0], (int)z);
    atomicMax(&box[6 * c + 1], (int)z);
    atomicMin(&box[6 * c + 2], (int)y);
    atomicMax(&box[6 * c + 3], (int)y);
    atomicMin(&box[6 * c + 4], (int)x);
    atomicMax(&box[6 * c + 5], (int)x);
}

__global__ __launch_bounds__(256) void k_stats_init(unsigned long long *gcnt, int *gbox)
{
    for (int c = threadIdx.x; c < kClasses; c += blockDim.x) {
        gcnt[c] = 0ull;
        gbox[6 * c + 0] = INT_MAX; gbox[6 * c + 1] = -1;
        gbox[6 * c + 2] = INT_MAX; gbox[6 * c + 3] = -1;
        gbox[6 * c + 4] = INT_MAX; gbox[6 * c + 5] = -1;
    }
}

// 16 bytes of labels per thread and iteration (V/VEC vectors), then the tail.
template <typename T>
__global__ __launch_bounds__(256) void k_label_stats(const T *__restrict__ lab, uint32_t V, uint32_t X, uint32_t sz,
                                                     unsigned long long *gcnt, int *gbox)
{
    constexpr int VEC = 16 / sizeof(T);
    __shared__ uint32_t cnt[kClasses];
    __shared__ int box[6 * kClasses];
    for (int c = threadIdx.x; c < kClasses; c += blockDim.x) {
        cnt[c] = 0u;
        box[6 * c + 0] = INT_MAX; box[6 * c + 1] = -1;
        box[6 * c + 2] = INT_MAX; box[6 * c + 3] = -1;
        box[6 * c + 4] = INT_MAX; box[6 * c + 5] = -1;
    }
    __syncthreads();
    const uint32_t nvec = V / VEC;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nvec; j += stride) {
        union { uint4 q; T v[VEC]; } u;
        u.q = ((const uint4 *)lab)[j];
        bool any = false;
#pragma unroll
        for (int k = 0; k < VEC; ++k) any |= u.v[k] != (T)0;
        if (any) {
#pragma unroll
            for (int k = 0; k < VEC; ++k) stats_add<T>(u.v[k], j * VEC + k, X, sz, cnt, box);
        }
    }
    for (uint32_t i = nvec * VEC + blockIdx.x * blockDim.x + threadIdx.x; i < V; i += stride)
        stats_add<T>(lab[i], i, X, sz, cnt, box);
    __syncthreads();
    for (int c = threadIdx.x; c < kClasses; c += blockDim.x) {
        if (!cnt[c]) continue;
        atomicAdd(&gcnt[c], (unsigned long long)cnt[c]);
        atomicMin(&gbox[6 * c + 0], box[6 * c + 0]);
        atomicMax(&gbox[6 * c + 1], box[6 * c + 1]);
        atomicMin(&gbox[6 * c + 2], box[6 * c + 2]);
        atomicMax(&gbox[6 * c + 3], box[6 * c + 3]);
        atomicMin(&gbox[6 * c + 4], box[6 * c + 4]);
        atomicMax(&gbox[6 * c + 5], box[6 * c + 5]);
    }
}

// Smallest and largest value (as double, exact for every supported type).
template <typename T>
__global__ __launch_bounds__(256) void k_minmax(const T *__restrict__ a, uint32_t n, double *out)
{
    __shared__ double smin[4], smax[4];
    double lo = INFINITY, hi = -INFINITY;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double v = (double)a[i];
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off, 64));
        hi = fmax(hi, __shfl_xor(hi, off, 64));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        smin[w] = lo;
        smax[w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            lo = fmin(lo, smin[k]);
            hi = fmax(hi, smax[k]);
        }
        // doubles ordered as integers: flip the magnitude bits of negatives
        const auto key = [](double d) {
            const long long b = __double_as_longlong(d);
            return b < 0 ? (unsigned long long)(~b) : (unsigned long long)b | 0x8000000000000000ull;
        };
        atomicMin((unsigned long long *)&out[0], key(lo));
        atomicMax((unsigned long long *)&out[1], key(hi));
    }
}

__global__ void k_minmax_init(double *out)
{
    if (threadIdx.x == 0) {
        ((unsigned long long *)out)[0] = ~0ull;
        ((unsigned long long *)out)[1] = 0ull;
    }
}

double minmax_key_to_double(unsigned long long k)
{
    const unsigned long long b = (k & 0x8000000000000000ull) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    double d;
    memcpy(&d, &b, sizeof(d));
    return d;
}

// Crop geometry: output voxel (zc, yc, xc) is voxel (z0+zc-1, y0+yc-1, x0+xc-1)
// of the unpadded input, or a pad voxel (0) when that lies outside it.
struct Crop {
    int Z, Y, X;        // input (unpadded)
    int z0, y0, x0;     // box origin in padded coordinates
    int Zc, Yc, Xc;     // output
};

__device__ __forceinline__ bool crop_src(const Crop &c, int zc, int yc, int &z, int &y)
{
    z = c.z0 + zc - 1;
    y = c.y0 + yc - 1;
    return z >= 0 && z < c.Z && y >= 0 && y < c.Y;
}

// Image modes: 0 copy (uint8 or float32 as it is), 1 int8 -> uint8 (+128),
// 2 any type -> float32 ((v - vmin) / vrange * 255, the IEEE operations of
// interpolation.py:140-144 in that order).
template <typename T, typename O, int MODE>
__device__ __forceinline__ O img_value(T v, float vmin, float vrange)
{
    if (MODE == 0) return (O)v;
    if (MODE == 1) return (O)(uint8_t)((uint8_t)v ^ 0x80u);
    float d = (float)v;
    d = d - vmin;
    d = d / vrange;
    d = d * 255.0f;
    return (O)d;
}

template <typename T, typename O, int MODE>
__global__ __launch_bounds__(256) void k_crop_image(const T *__restrict__ in, Crop c, float vmin, float vrange,
                                                    O *__restrict__ out)
{
    const uint32_t rows = (uint32_t)c.Zc * (uint32_t)c.Yc;
    for (uint32_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const int zc = (int)(row / (uint32_t)c.Yc), yc = (int)(row - (uint32_t)zc * (uint32_t)c.Yc);
        int z, y;
        const bool inrow = crop_src(c, zc, yc, z, y);
        O *o = out + (size_t)row * c.Xc;
        const T *src = in + (inrow ? ((size_t)z * c.Y + y) * c.X : 0);
        for (int xc = threadIdx.x; xc < c.Xc; xc += blockDim.x) {
            const int x = c.x0 + xc - 1;
            o[xc] = (inrow && x >= 0 && x < c.X) ? img_value<T, O, MODE>(src[x], vmin, vrange) : (O)0;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_crop_labels(const T *__restrict__ in, Crop c, const uint8_t *__restrict__ lut,
                                                     uint8_t *__restrict__ out)
{
    __shared__ uint8_t slut[kClasses];
    for (int k = threadIdx.x; k < kClasses; k += blockDim.x) slut[k] = lut[k];
    __syncthreads();
    const uint32_t rows = (uint32_t)c.Zc * (uint32_t)c.Yc;
    for (uint32_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const int zc = (int)(row / (uint32_t)c.Yc), yc = (int)(row - (uint32_t)zc * (uint32_t)c.Yc);
        int z, y;
        const bool inrow = crop_src(c, zc, yc, z, y);
        uint8_t *o = out + (size_t)row * c.Xc;
        const T *src = in + (inrow ? ((size_t)z * c.Y + y) * c.X : 0);
        for (int xc = threadIdx.x; xc < c.Xc; xc += blockDim.x) {
            const int x = c.x0 + xc - 1;
            o[xc] = (inrow && x >= 0 && x < c.X) ? slut[label_class(src[x])] : slut[0];
        }
    }
}

// out (Z,Y,X) = crop voxel where the box covers the (padded) voxel, else 0.
__global__ __launch_bounds__(256) void k_place(const uint8_t *__restrict__ crop, Crop c, uint8_t *__restrict__ out)
{
    const uint32_t rows = (uint32_t)c.Z * (uint32_t)c.Y;
    for (uint32_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const int z = (int)(row / (uint32_t)c.Y), y = (int)(row - (uint32_t)z * (uint32_t)c.Y);
        const int zc = z + 1 - c.z0, yc = y + 1 - c.y0;
        const bool inrow = zc >= 0 && zc < c.Zc && yc >= 0 && yc < c.Yc;
        uint8_t *o = out + (size_t)row * c.X;
        const uint8_t *src = crop + (inrow ? ((size_t)zc * c.Yc + yc) * c.Xc : 0);
        for (int x = threadIdx.x; x < c.X; x += blockDim.x) {
            const int xc = x + 1 - c.x0;
            o[x] = (inrow && xc >= 0 && xc < c.Xc) ? src[xc] : (uint8_t)0;
        }
    }
}

// Dice counts of the sanity check: a = matches, b = labelled voxels, c =
// labelled voxels with a non-zero result.
__global__ __launch_bounds__(256) void k_dice_counts(const uint8_t *__restrict__ idx, const uint8_t *__restrict__ res,
                                                     uint32_t n, const uint8_t *__restrict__ values,
                                                     const uint8_t *__restrict__ counted, int nlabels,
                                                     unsigned long long *out)
{
    __shared__ uint8_t sval[256], scnt[256];
    sval[threadIdx.x] = values[threadIdx.x];
    scnt[threadIdx.x] = counted[threadIdx.x];
    __syncthreads();
    uint64_t a = 0, b = 0, cc = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t k = idx[i];
        if (k == 0u || k >= (uint32_t)nlabels) continue;   // background, removed or ignored
        const uint32_t v = sval[k], r = res[i];
        b += 1;
        cc += r != 0u;
        a += (r == v) & (scnt[v] != 0u);
    }
    wave_sum_add(out + 0, a);
    wave_sum_add(out + 1, b);
    wave_sum_add(out + 2, cc);
}

Crop make_crop(int64_t Z, int64_t Y, int64_t X, const int64_t *box)
{
    Crop c;
    c.Z = (int)Z; c.Y = (int)Y; c.X = (int)X;
    c.z0 = (int)box[0]; c.y0 = (int)box[2]; c.x0 = (int)box[4];
    c.Zc = (int)(box[1] - box[0]); c.Yc = (int)(box[3] - box[2]); c.Xc = (int)(box[5] - box[4]);
    return c;
}

int check_box(int64_t Z, int64_t Y, int64_t X, const int64_t *box, const char *fn)
{
    BMX_CHECK(box != nullptr, "%s: null box", fn);
    BMX_CHECK(Z >= 1 && Y >= 1 && X >= 1 && Z * Y * X < ((int64_t)1 << 32), "%s: bad dims", fn);
    const int64_t P[3] = {Z + 2, Y + 2, X + 2};
    for (int a = 0; a < 3; ++a)
        BMX_CHECK(box[2 * a] >= 0 && box[2 * a] < box[2 * a + 1] && box[2 * a + 1] <= P[a],
                  "%s: box [%lld, %lld) outside the padded axis %d of %lld", fn, (long long)box[2 * a],
                  (long long)box[2 * a + 1], a, (long long)P[a]);
    BMX_CHECK((box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4]) < ((int64_t)1 << 32), "%s: box too large", fn);
    return BMX_OK;
}

unsigned row_grid(int64_t rows)
{
    return (unsigned)(rows < 1 ? 1 : (rows < 65536 ? rows : 65536));
}

size_t raw_size(int t)
{
    switch (t) {
    case BMX_RAW_U8: case BMX_RAW_I8: return 1;
    case BMX_RAW_U16: case BMX_RAW_I16: return 2;
    case BMX_RAW_U32: case BMX_RAW_I32: case BMX_RAW_F32: return 4;
    default: return 0;
    }
}

}  // namespace

extern "C" int bmx_label_stats(bmx_ctx *ctx, const void *lab, int lab_dtype, int64_t Z, int64_t Y, int64_t X,
                               uint64_t *counts, int32_t *boxes)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(lab && counts && boxes, "bmx_label_stats: null pointer");
    BMX_CHECK(lab_dtype != BMX_RAW_F32 && raw_size(lab_dtype) > 0, "bmx_label_stats: bad label dtype %d", lab_dtype);
    BMX_CHECK(Z >= 1 && Y >= 1 && X >= 1 && Z * Y * X < ((int64_t)1 << 32), "bmx_label_stats: bad dims");
    const uint32_t V = (uint32_t)(Z * Y * X), sz = (uint32_t)(Y * X);
    unsigned long long *gcnt = (unsigned long long *)ctx->d_small;
    int *gbox = (int *)(gcnt + kClasses);
    k_stats_init<<<1, 256, 0, ctx->stream>>>(gcnt, gbox);
    BMX_LAUNCH_CHECK();
    const unsigned grid = bmx_stride_blocks((int64_t)V / (16 / (int64_t)raw_size(lab_dtype)) + 1, 256);
    const unsigned g = grid < 2048 ? grid : 2048;
    switch (lab_dtype) {
    case BMX_RAW_U8: k_label_stats<uint8_t><<<g, 256, 0, ctx->stream>>>((const uint8_t *)lab, V, (uint32_t)X, sz, gcnt, gbox); break;
    case BMX_RAW_I8: k_label_stats<int8_t><<<g, 256, 0, ctx->stream>>>((const int8_t *)lab, V, (uint32_t)X, sz, gcnt, gbox); break;
    case BMX_RAW_U16: k_label_stats<uint16_t><<<g, 256, 0, ctx->stream>>>((const uint16_t *)lab, V, (uint32_t)X, sz, gcnt, gbox); break;
    case BMX_RAW_I16: k_label_stats<int16_t><<<g, 256, 0, ctx->stream>>>((const int16_t *)lab, V, (uint32_t)X, sz, gcnt, gbox); break;
    case BMX_RAW_U32: k_label_stats<uint32_t><<<g, 256, 0, ctx->stream>>>((const uint32_t *)lab, V, (uint32_t)X, sz, gcnt, gbox); break;
    default: k_label_stats<int32_t><<<g, 256, 0, ctx->stream>>>((const int32_t *)lab, V, (uint32_t)X, sz, gcnt, gbox); break;
    }
    BMX_LAUNCH_CHECK();
    BMX_HIP(hipMemcpyAsync(counts, gcnt, kClasses * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
    BMX_HIP(hipMemcpyAsync(boxes, gbox, 6 * kClasses * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    BMX_HIP(hipStreamSynchronize(ctx->stream));
    return BMX_OK;
}

extern "C" int bmx_image_range(bmx_ctx *ctx, const void *img, int img_raw_dtype, int64_t n, double *vmin,
                               double *vmax)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(img && vmin && vmax, "bmx_image_range: null pointer");
    BMX_CHECK(raw_size(img_raw_dtype) > 0, "bmx_image_range: bad dtype %d", img_raw_dtype);
    BMX_CHECK(n >= 1 && n < ((int64_t)1 << 32), "bmx_image_range: bad size");
    double *d = (double *)ctx->d_small;
    k_minmax_init<<<1, 64, 0, ctx->stream>>>(d);
    BMX_LAUNCH_CHECK();
    const unsigned grid = bmx_stride_blocks(n, 256);
    const unsigned g = grid < 2048 ? grid : 2048;
    const uint32_t N = (uint32_t)n;
    switch (img_raw_dtype) {
    case BMX_RAW_U8: k_minmax<uint8_t><<<g, 256, 0, ctx->stream>>>((const uint8_t *)img, N, d); break;
    case BMX_RAW_I8: k_minmax<int8_t><<<g, 256, 0, ctx->stream>>>((const int8_t *)img, N, d); break;
    case BMX_RAW_U16: k_minmax<uint16_t><<<g, 256, 0, ctx->stream>>>((const uint16_t *)img, N, d); break;
    case BMX_RAW_I16: k_minmax<int16_t><<<g, 256, 0, ctx->stream>>>((const int16_t *)img, N, d); break;
    case BMX_RAW_U32: k_minmax<uint32_t><<<g, 256, 0, ctx->stream>>>((const uint32_t *)img, N, d); break;
    case BMX_RAW_I32: k_minmax<int32_t><<<g, 256, 0, ctx->stream>>>((const int32_t *)img, N, d); break;
    default: k_minmax<float><<<g, 256, 0, ctx->stream>>>((const float *)img, N, d); break;
    }
    BMX_LAUNCH_CHECK();
    unsigned long long k[2];
    BMX_HIP(hipMemcpyAsync(k, d, sizeof(k), hipMemcpyDeviceToHost, ctx->stream));
    BMX_HIP(hipStreamSynchronize(ctx->stream));
    *vmin = minmax_key_to_double(k[0]);
    *vmax = minmax_key_to_double(k[1]);
    return BMX_OK;
}

extern "C" int bmx_crop_image(bmx_ctx *ctx, const void *img, int img_raw_dtype, int normalize, int64_t Z, int64_t Y,
                              int64_t X, const int64_t *box, float vmin, float vrange, void *out)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(img && out, "bmx_crop_image: null pointer");
    rc = check_box(Z, Y, X, box, "bmx_crop_image");
    if (rc) return rc;
    BMX_CHECK(raw_size(img_raw_dtype) > 0, "bmx_crop_image: bad dtype %d", img_raw_dtype);
    const Crop c = make_crop(Z, Y, X, box);
    const unsigned g = row_grid((int64_t)c.Zc * c.Yc);
    if (!normalize) {
        BMX_CHECK(img_raw_dtype == BMX_RAW_U8 || img_raw_dtype == BMX_RAW_I8 || img_raw_dtype == BMX_RAW_F32,
                  "bmx_crop_image: without normalisation the image must be uint8, int8 or float32");
        if (img_raw_dtype == BMX_RAW_U8)
            k_crop_image<uint8_t, uint8_t, 0><<<g, 256, 0, ctx->stream>>>((const uint8_t *)img, c, 0.f, 1.f,
                                                                           (uint8_t *)out);
        else if (img_raw_dtype == BMX_RAW_I8)
            k_crop_image<int8_t, uint8_t, 1><<<g, 256, 0, ctx->stream>>>((const int8_t *)img, c, 0.f, 1.f,
                                                                          (uint8_t *)out);
        else
            k_crop_image<float, float, 0><<<g, 256, 0, ctx->stream>>>((const float *)img, c, 0.f, 1.f, (float *)out);
    } else {
        float *o = (float *)out;
        switch (img_raw_dtype) {
        case BMX_RAW_U8: k_crop_image<uint8_t, float, 2><<<g, 256, 0, ctx->stream>>>((const uint8_t *)img, c, vmin, vrange, o); break;
        case BMX_RAW_I8: k_crop_image<int8_t, float, 2><<<g, 256, 0, ctx->stream>>>((const int8_t *)img, c, vmin, vrange, o); break;
        case BMX_RAW_U16: k_crop_image<uint16_t, float, 2><<<g, 256, 0, ctx->stream>>>((const uint16_t *)img, c, vmin, vrange, o); break;
        case BMX_RAW_I16: k_crop_image<int16_t, float, 2><<<g, 256, 0, ctx->stream>>>((const int16_t *)img, c, vmin, vrange, o); break;
        case BMX_RAW_U32: k_crop_image<uint32_t, float, 2><<<g, 256, 0, ctx->stream>>>((const uint32_t *)img, c, vmin, vrange, o); break;
        case BMX_RAW_I32: k_crop_image<int32_t, float, 2><<<g, 256, 0, ctx->stream>>>((const int32_t *)img, c, vmin, vrange, o); break;
        default: k_crop_image<float, float, 2><<<g, 256, 0, ctx->stream>>>((const float *)img, c, vmin, vrange, o); break;
        }
    }
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

extern "C" int bmx_crop_labels(bmx_ctx *ctx, const void *lab, int lab_dtype, int64_t Z, int64_t Y, int64_t X,
                               const int64_t *box, const uint8_t *lut, uint8_t *out)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(lab && lut && out, "bmx_crop_labels: null pointer");
    BMX_CHECK(lab_dtype != BMX_RAW_F32 && raw_size(lab_dtype) > 0, "bmx_crop_labels: bad label dtype %d", lab_dtype);
    rc = check_box(Z, Y, X, box, "bmx_crop_labels");
    if (rc) return rc;
    const Crop c = make_crop(Z, Y, X, box);
    uint8_t *dlut = (uint8_t *)ctx->d_small + kSmallLutOffset;
    BMX_HIP(hipMemcpyAsync(dlut, lut, kClasses, hipMemcpyHostToDevice, ctx->stream));
    const unsigned g = row_grid((int64_t)c.Zc * c.Yc);
    switch (lab_dtype) {
    case BMX_RAW_U8: k_crop_labels<uint8_t><<<g, 256, 0, ctx->stream>>>((const uint8_t *)lab, c, dlut, out); break;
    case BMX_RAW_I8: k_crop_labels<int8_t><<<g, 256, 0, ctx->stream>>>((const int8_t *)lab, c, dlut, out); break;
    case BMX_RAW_U16: k_crop_labels<uint16_t><<<g, 256, 0, ctx->stream>>>((const uint16_t *)lab, c, dlut, out); break;
    case BMX_RAW_I16: k_crop_labels<int16_t><<<g, 256, 0, ctx->stream>>>((const int16_t *)lab, c, dlut, out); break;
    case BMX_RAW_U32: k_crop_labels<uint32_t><<<g, 256, 0, ctx->stream>>>((const uint32_t *)lab, c, dlut, out); break;
    default: k_crop_labels<int32_t><<<g, 256, 0, ctx->stream>>>((const int32_t *)lab, c, dlut, out); break;
    }
    BMX_LAUNCH_CHECK();
    // the table is read by the kernel before the next call can overwrite it
    BMX_HIP(hipStreamSynchronize(ctx->stream));
    return BMX_OK;
}

extern "C" int bmx_place(bmx_ctx *ctx, const uint8_t *crop, int64_t Z, int64_t Y, int64_t X, const int64_t *box,
                         uint8_t *out)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(crop && out && (const void *)crop != (const void *)out, "bmx_place: null or aliased pointers");
    rc = check_box(Z, Y, X, box, "bmx_place");
    if (rc) return rc;
    const Crop c = make_crop(Z, Y, X, box);
    k_place<<<row_grid(Z * Y), 256, 0, ctx->stream>>>(crop, c, out);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

extern "C" int bmx_dice_counts(bmx_ctx *ctx, const uint8_t *labidx, const uint8_t *result, int64_t n, int nlabels,
                               const uint8_t *values, const uint8_t *counted, uint64_t *counts)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(labidx && result && values && counted && counts, "bmx_dice_counts: null pointer");
    BMX_CHECK(n >= 0 && n < ((int64_t)1 << 32), "bmx_dice_counts: bad size");
    BMX_CHECK(nlabels >= 1 && nlabels <= 256, "bmx_dice_counts: bad label count");
    unsigned long long *d = (unsigned long long *)ctx->d_small;
    uint8_t *dv = (uint8_t *)ctx->d_small + kSmallLutOffset;
    BMX_HIP(hipMemcpyAsync(dv, values, (size_t)nlabels, hipMemcpyHostToDevice, ctx->stream));
    if (nlabels < 256) BMX_HIP(hipMemsetAsync(dv + nlabels, 0, (size_t)(256 - nlabels), ctx->stream));
    BMX_HIP(hipMemcpyAsync(dv + 256, counted, 256, hipMemcpyHostToDevice, ctx->stream));
    BMX_HIP(hipMemsetAsync(d, 0, 3 * sizeof(unsigned long long), ctx->stream));
    if (n > 0) {
        const unsigned grid = bmx_stride_blocks(n, 256);
        k_dice_counts<<<grid < 2048 ? grid : 2048, 256, 0, ctx->stream>>>(labidx, result, (uint32_t)n, dv, dv + 256,
                                                                           nlabels, d);
        BMX_LAUNCH_CHECK();
    }
    unsigned long long h[3];
    BMX_HIP(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    BMX_HIP(hipStreamSynchronize(ctx->stream));
    counts[0] = h[0];
    counts[1] = h[1];
    counts[2] = h[2];
    return BMX_OK;
}
