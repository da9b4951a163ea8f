// bmx_walk.hip -- the random walker.
//
// Transition law of the reference kernel randomWalk
// (biomedisa/features/random_walk/pyopencl_small.py:192-314): at voxel i the
// walker started at seed s (intensity B, variance var) weighs the six faces
// W_d = exp(-(B - img[i+delta_d])^2 / (2 var)) in the order +z,-z,+y,-y,+x,-x,
// prefix-sums them, draws rand = W5*u and takes the first d whose prefix
// exceeds rand (rand == 0 -> +z).  A move that would leave the interior is
// rejected but still counts as a step; an accepted move adds one hit to the
// seed label's counter at the new voxel.  Each walk restarts at the seed.
//
// Two random streams:
//   Philox (default): one thread per (seed, walk); counter (pos, 0, walk,
//     step/4) under a 64-bit key, word step%4, u = (word >> 8) * 2^-24.
//     Walks of a seed are independent, so the grid has nseeds*nbrw threads.
//   Reference: one thread per seed, the float MRG recurrence seeded with the
//     voxel index and carried across all walks (:213-223, :256-282).
//
// Memory layout for the Philox walker: the image is re-packed once per call
// into one record per voxel holding the six neighbour intensities and a
// 6-bit mask of which neighbours are interior, so a step is one load (8 bytes
// for uint8 images, 32 for float32) plus one hit atomic, and the walker needs
// no coordinates.  For uint8 images the six weights come from a per-seed
// table exp(-t^2 * div1), t = |B - A| in 0..255, built in LDS with the same
// bmx_weight arithmetic, so table lookups return exactly the computed value.
// Blocks are remapped so that each XCD walks a contiguous range of seeds.
#include "bmx_walk_common.h"

namespace {
constexpr int kLutMinWalks = 8;   // table path when a block holds <= 32 seeds

struct RecF32 {
    float a[6];
    uint32_t mask;
    uint32_t pad;
};

// Interior bits of the six neighbours of interior voxel (z,y,x).
__device__ __forceinline__ uint32_t interior_mask(const WalkGeom &g, int z, int y, int x)
{
    return (uint32_t)(z + 1 < g.Z - 1) | ((uint32_t)(z - 1 > 0) << 1) |
           ((uint32_t)(y + 1 < g.Y - 1) << 2) | ((uint32_t)(y - 1 > 0) << 3) |
           ((uint32_t)(x + 1 < g.X - 1) << 4) | ((uint32_t)(x - 1 > 0) << 5);
}

__global__ __launch_bounds__(256) void k_pack_u8(const uint8_t *__restrict__ img, WalkGeom g,
                                                 uint64_t *__restrict__ rec)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.V) return;
    const int z = (int)(v / g.sz);
    const uint32_t r = (uint32_t)(v - (int64_t)z * g.sz);
    const int y = (int)(r / g.sy), x = (int)(r - (uint32_t)y * g.sy);
    uint64_t out = 0;
    if (z > 0 && z < g.Z - 1 && y > 0 && y < g.Y - 1 && x > 0 && x < g.X - 1) {
        out = (uint64_t)img[v + g.sz] | ((uint64_t)img[v - g.sz] << 8) |
              ((uint64_t)img[v + g.sy] << 16) | ((uint64_t)img[v - g.sy] << 24) |
              ((uint64_t)img[v + 1] << 32) | ((uint64_t)img[v - 1] << 40) |
              ((uint64_t)interior_mask(g, z, y, x) << 48);
    }
    rec[v] = out;
}

__global__ __launch_bounds__(256) void k_pack_f32(const float *__restrict__ img, WalkGeom g,
                                                  RecF32 *__restrict__ rec)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.V) return;
    const int z = (int)(v / g.sz);
    const uint32_t r = (uint32_t)(v - (int64_t)z * g.sz);
    const int y = (int)(r / g.sy), x = (int)(r - (uint32_t)y * g.sy);
    RecF32 out = {{0, 0, 0, 0, 0, 0}, 0u, 0u};
    if (z > 0 && z < g.Z - 1 && y > 0 && y < g.Y - 1 && x > 0 && x < g.X - 1) {
        out.a[0] = img[v + g.sz];
        out.a[1] = img[v - g.sz];
        out.a[2] = img[v + g.sy];
        out.a[3] = img[v - g.sy];
        out.a[4] = img[v + 1];
        out.a[5] = img[v - 1];
        out.mask = interior_mask(g, z, y, x);
    }
    rec[v] = out;
}

// Philox walker, uint8 image, weights from per-seed LDS tables.
// Block = spb seeds x nbrw walks (spb * nbrw <= 256).
__global__ __launch_bounds__(kWalkThreads) void k_walk_philox_u8_lut(
    const uint64_t *__restrict__ rec, const uint8_t *__restrict__ img, WalkGeom g,
    const bmx_seed *__restrict__ seeds, int64_t nseeds, int nbrw, int sorw, int spb, uint32_t k0,
    uint32_t k1, uint32_t *__restrict__ hits, unsigned long long *accepted)
{
    extern __shared__ float lut[];   // [spb][256]
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t s0 = (int64_t)lb * spb;
    for (int k = threadIdx.x; k < spb * 256; k += kWalkThreads) {
        const int slot = k >> 8, t = k & 255;
        const int64_t s = s0 + slot;
        float w = 0.0f;
        if (s < nseeds) {
            const bmx_seed sd = seeds[s];
            const float div1 = 1.0f / (2.0f * sd.var);
            const float ft = (float)t;
            w = bmx_expf(((-ft) * ft) * div1);
        }
        lut[k] = w;
    }
    __syncthreads();
    const int slot = threadIdx.x / nbrw;
    const int w = threadIdx.x - slot * nbrw;
    const int64_t s = s0 + slot;
    uint64_t acc = 0;
    if (slot < spb && s < nseeds) {
        const bmx_seed sd = seeds[s];
        const uint32_t p0 = sd.pos;
        const int Bi = img[p0];
        const float *__restrict__ tab = lut + slot * 256;
        uint32_t *__restrict__ h = hits + (int64_t)sd.label * g.V;
        uint32_t i = p0;
        for (int sb = 0; sb < sorw; sb += 4) {
            const bmx_u4 r = bmx_philox(p0, 0u, (uint32_t)w, (uint32_t)(sb >> 2), k0, k1);
            const uint32_t rr[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (sb + j < sorw) {
                    const uint64_t q = rec[i];
                    const uint32_t lo = (uint32_t)q, hi = (uint32_t)(q >> 32);
                    const float w0 = tab[abs(Bi - (int)(lo & 255u))];
                    const float w1 = tab[abs(Bi - (int)((lo >> 8) & 255u))];
                    const float w2 = tab[abs(Bi - (int)((lo >> 16) & 255u))];
                    const float w3 = tab[abs(Bi - (int)(lo >> 24))];
                    const float w4 = tab[abs(Bi - (int)(hi & 255u))];
                    const float w5 = tab[abs(Bi - (int)((hi >> 8) & 255u))];
                    const int d = choose(w0, w1, w2, w3, w4, w5, u01(rr[j]));
                    if ((hi >> (16 + d)) & 1u) {
                        i += step_delta(d, g.sz, g.sy);
                        atomicAdd(&h[i], 1u);
                        ++acc;
                    }
                }
            }
        }
    }
    wave_add_u64(accepted, acc);
}

// Philox walker with computed weights: uint8 (8-byte records) or float32
// (32-byte records).  One thread per (seed, walk), seed-major.
template <typename T>
__global__ __launch_bounds__(kWalkThreads) void k_walk_philox_exp(
    const void *__restrict__ recv, const T *__restrict__ img, WalkGeom g,
    const bmx_seed *__restrict__ seeds, int64_t nseeds, int nbrw, int sorw, uint32_t k0, uint32_t k1,
    uint32_t *__restrict__ hits, unsigned long long *accepted)
{
    const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
    const int64_t tid = (int64_t)lb * kWalkThreads + threadIdx.x;
    const int64_t s = tid / nbrw;
    uint64_t acc = 0;
    if (s < nseeds) {
        const int w = (int)(tid - s * nbrw);
        const bmx_seed sd = seeds[s];
        const uint32_t p0 = sd.pos;
        const float B = bmx_ld(img, (int64_t)p0);
        const float div1 = 1.0f / (2.0f * sd.var);
        uint32_t *__restrict__ h = hits + (int64_t)sd.label * g.V;
        uint32_t i = p0;
        for (int sb = 0; sb < sorw; sb += 4) {
            const bmx_u4 r = bmx_philox(p0, 0u, (uint32_t)w, (uint32_t)(sb >> 2), k0, k1);
            const uint32_t rr[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (sb + j < sorw) {
                    float a[6];
                    uint32_t mask;
                    if constexpr (sizeof(T) == 1) {
                        const uint64_t q = ((const uint64_t *)recv)[i];
#pragma unroll
                        for (int k = 0; k < 6; ++k) a[k] = (float)(uint32_t)((q >> (8 * k)) & 255u);
                        mask = (uint32_t)(q >> 48);
                    } else {
                        const float4 *p = (const float4 *)((const RecF32 *)recv + i);
                        const float4 q0 = p[0], q1 = p[1];
                        a[0] = q0.x; a[1] = q0.y; a[2] = q0.z; a[3] = q0.w;
                        a[4] = q1.x; a[5] = q1.y;
                        mask = __float_as_uint(q1.z);
                    }
                    const int d = choose(bmx_weight(B, a[0], div1), bmx_weight(B, a[1], div1),
                                         bmx_weight(B, a[2], div1), bmx_weight(B, a[3], div1),
                                         bmx_weight(B, a[4], div1), bmx_weight(B, a[5], div1),
                                         u01(rr[j]));
                    if ((mask >> d) & 1u) {
                        i += step_delta(d, g.sz, g.sy);
                        atomicAdd(&h[i], 1u);
                        ++acc;
                    }
                }
            }
        }
    }
    wave_add_u64(accepted, acc);
}

template <typename T>
__global__ __launch_bounds__(kWalkThreads) void k_walk_ref(
    const T *__restrict__ img, WalkGeom g, const bmx_seed *__restrict__ seeds, int64_t nseeds,
    int nbrw, int sorw, uint32_t *__restrict__ hits, unsigned long long *accepted)
{
    const int64_t s = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * kWalkThreads + threadIdx.x;
    uint64_t acc = 0;
    if (s < nseeds) {
        const bmx_seed sd = seeds[s];
        const uint32_t p0 = sd.pos;
        const int z0 = (int)(p0 / g.sz);
        const uint32_t r0 = p0 - (uint32_t)z0 * g.sz;
        const int y0 = (int)(r0 / g.sy);
        const int x0 = (int)(r0 - (uint32_t)y0 * g.sy);
        const float B = bmx_ld(img, (int64_t)p0);
        const float div1 = 1.0f / (2.0f * sd.var);
        uint32_t *__restrict__ h = hits + (int64_t)sd.label * g.V;
        const float norm = 2.328306549295728e-10f;
        const float m1 = 4294967087.0f, m2 = 4294944443.0f;
        const float a12 = 1403580.0f, a13n = 810728.0f, a21 = 527612.0f, a23n = 1370589.0f;
        const float fi = (float)(int32_t)p0;
        float s10 = fi, s11 = fi, s12 = fi, s20 = fi, s21 = fi, s22 = fi;
        uint32_t i = p0;
        int z = z0, y = y0, x = x0;
        int n = 0, o = 0, p = 0;
        for (int w = 0; w < nbrw; ++w) {
            for (int step = 0; step < sorw; ++step) {
                const float a0 = bmx_ld(img, (int64_t)(i + g.sz));
                const float a1 = bmx_ld(img, (int64_t)(i - g.sz));
                const float a2 = bmx_ld(img, (int64_t)(i + g.sy));
                const float a3 = bmx_ld(img, (int64_t)(i - g.sy));
                const float a4 = bmx_ld(img, (int64_t)(i + 1u));
                const float a5 = bmx_ld(img, (int64_t)(i - 1u));
                const float W0 = bmx_weight(B, a0, div1);
                const float W1 = bmx_weight(B, a1, div1) + W0;
                const float W2 = bmx_weight(B, a2, div1) + W1;
                const float W3 = bmx_weight(B, a3, div1) + W2;
                const float W4 = bmx_weight(B, a4, div1) + W3;
                const float W5 = bmx_weight(B, a5, div1) + W4;
                // the recurrence as the reference's compilers contract it
                // (FP_CONTRACT on): a*b - c*d -> fma(a, b, -(c*d)), p -= k*m -> fma(-k, m, p)
                float p1 = __builtin_fmaf(a12, s11, -(a13n * s10));
                int64_t kk = (int64_t)(p1 / m1);
                p1 = __builtin_fmaf(-(float)kk, m1, p1);
                if (p1 < 0.0f) p1 += m1;
                s10 = s11; s11 = s12; s12 = p1;
                float p2 = __builtin_fmaf(a21, s22, -(a23n * s20));
                kk = (int64_t)(p2 / m2);
                p2 = __builtin_fmaf(-(float)kk, m2, p2);
                if (p2 < 0.0f) p2 += m2;
                s20 = s21; s21 = s22; s22 = p2;
                const float rand = (p1 <= p2) ? W5 * ((p1 - p2 + m1) * norm) : W5 * ((p1 - p2) * norm);
                // Direction chain of :285-290; registers keep their value when
                // no branch matches.
                if (rand < W0 || rand == 0.0f) { n = 1; o = 0; p = 0; }
                else if (rand >= W0 && rand < W1) { n = -1; o = 0; p = 0; }
                else if (rand >= W1 && rand < W2) { n = 0; o = 1; p = 0; }
                else if (rand >= W2 && rand < W3) { n = 0; o = -1; p = 0; }
                else if (rand >= W3 && rand < W4) { n = 0; o = 0; p = 1; }
                else if (rand >= W4 && rand <= W5) { n = 0; o = 0; p = -1; }
                const int zn = z + n, yn = y + o, xn = x + p;
                if (zn > 0 && zn < g.Z - 1 && yn > 0 && yn < g.Y - 1 && xn > 0 && xn < g.X - 1) {
                    z = zn; y = yn; x = xn;
                    i = (uint32_t)z * g.sz + (uint32_t)y * g.sy + (uint32_t)x;
                    atomicAdd(&h[i], 1u);
                    ++acc;
                }
            }
            z = z0; y = y0; x = x0;
            i = p0;
        }
    }
    wave_add_u64(accepted, acc);
}

// ---------------------------------------------------------------------------
// Sorted lockstep walker.  Every (seed, walk) pair is a walker with state
// {pos, seed, walk, step}.  All walkers advance `seg` steps per launch; between
// launches they are counting-sorted by the spatial brick of their position
// (bricks of 2^b voxels per side, 4x4x4 bricks per super-tile, super-tiles in
// raster order, bricks in Morton order inside), so the walkers resident on one
// XCD at any time share a compact region of the image records and hit volume
// that fits in its L2.  The Philox counter depends only on (seed voxel, walk,
// step), so the order in which walkers run changes nothing in the hits.
// ---------------------------------------------------------------------------
struct SortGeom {
    uint32_t sz, sy;
    int b;            // log2 brick side
    uint32_t nsx, nsy; // super-tiles along x, y
};

__device__ __forceinline__ uint32_t morton3_2bit(uint32_t z, uint32_t y, uint32_t x)
{
    return ((z & 1u) << 2) | ((y & 1u) << 1) | (x & 1u) | ((z & 2u) << 4) | ((y & 2u) << 3) |
           ((x & 2u) << 2);
}

__device__ __forceinline__ uint32_t brick_key(uint32_t pos, const SortGeom &sg)
{
    const uint32_t z = pos / sg.sz;
    const uint32_t r = pos - z * sg.sz;
    const uint32_t y = r / sg.sy;
    const uint32_t x = r - y * sg.sy;
    const uint32_t bz = z >> sg.b, by = y >> sg.b, bx = x >> sg.b;
    const uint32_t st = ((bz >> 2) * sg.nsy + (by >> 2)) * sg.nsx + (bx >> 2);
    return st * 64u + morton3_2bit(bz & 3u, by & 3u, bx & 3u);
}

__global__ __launch_bounds__(256) void k_seg_init(const bmx_seed *__restrict__ seeds, int64_t nseeds,
                                                  int nbrw, uint4 *__restrict__ st)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseeds * nbrw) return;
    const int64_t s = i / nbrw;
    st[i] = make_uint4(seeds[s].pos, (uint32_t)s, (uint32_t)(i - s * nbrw), 0u);
}

template <typename T>
__global__ __launch_bounds__(kWalkThreads) void k_seg_walk(
    const void *__restrict__ recv, const T *__restrict__ img, WalkGeom g,
    const bmx_seed *__restrict__ seeds, uint4 *__restrict__ st, int64_t n, int sorw, int seg,
    uint32_t k0, uint32_t k1, uint32_t *__restrict__ hits, unsigned long long *accepted,
    uint32_t *__restrict__ counts, SortGeom sg)
{
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * kWalkThreads + threadIdx.x;
    uint64_t acc = 0;
    bool live = false;
    uint32_t key = 0;
    if (i < n) {
        uint4 w4 = st[i];
        uint32_t pos = w4.x;
        const uint32_t w = w4.z;
        int step = (int)w4.w;
        const bmx_seed sd = seeds[w4.y];
        const uint32_t p0 = sd.pos;
        const float B = bmx_ld(img, (int64_t)p0);
        const float div1 = 1.0f / (2.0f * sd.var);
        uint32_t *__restrict__ h = hits + (int64_t)sd.label * g.V;
        const int end = min(step + seg, sorw);
        for (int sb = step; sb < end; sb += 4) {
            const bmx_u4 r = bmx_philox(p0, 0u, w, (uint32_t)(sb >> 2), k0, k1);
            const uint32_t rr[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (sb + j < end) {
                    float a[6];
                    uint32_t mask;
                    if constexpr (sizeof(T) == 1) {
                        const uint64_t q = ((const uint64_t *)recv)[pos];
#pragma unroll
                        for (int k = 0; k < 6; ++k) a[k] = (float)(uint32_t)((q >> (8 * k)) & 255u);
                        mask = (uint32_t)(q >> 48);
                    } else {
                        const float4 *p = (const float4 *)((const RecF32 *)recv + pos);
                        const float4 q0 = p[0], q1 = p[1];
                        a[0] = q0.x; a[1] = q0.y; a[2] = q0.z; a[3] = q0.w;
                        a[4] = q1.x; a[5] = q1.y;
                        mask = __float_as_uint(q1.z);
                    }
                    const int d = choose(bmx_weight(B, a[0], div1), bmx_weight(B, a[1], div1),
                                         bmx_weight(B, a[2], div1), bmx_weight(B, a[3], div1),
                                         bmx_weight(B, a[4], div1), bmx_weight(B, a[5], div1),
                                         u01(rr[j]));
                    if ((mask >> d) & 1u) {
                        pos += step_delta(d, g.sz, g.sy);
                        atomicAdd(&h[pos], 1u);
                        ++acc;
                    }
                }
            }
        }
        st[i] = make_uint4(pos, w4.y, w, (uint32_t)end);
        live = end < sorw;
        key = brick_key(pos, sg);
    }
    wave_add_u64(accepted, acc);
    if (counts) (void)agg_atomic_inc(counts, key, live);
}

__global__ __launch_bounds__(256) void k_seg_scatter(const uint4 *__restrict__ in, int64_t n,
                                                     uint4 *__restrict__ out, uint32_t *cursor,
                                                     SortGeom sg, int sorw)
{
    const int64_t i = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    uint4 w4 = make_uint4(0, 0, 0, 0);
    bool live = false;
    uint32_t key = 0;
    if (i < n) {
        w4 = in[i];
        live = (int)w4.w < sorw;
        key = brick_key(w4.x, sg);
    }
    const uint32_t dst = agg_atomic_inc(cursor, key, live);
    if (live) out[dst] = w4;
}

// Packed neighbour records, cached in the context (grown on demand).
int ensure_records(bmx_ctx *ctx, size_t bytes)
{
    if (ctx->rec_cap >= bytes) return BMX_OK;
    if (ctx->d_rec) {
        BMX_HIP(hipStreamSynchronize(ctx->stream));
        BMX_HIP(hipFree(ctx->d_rec));
    }
    ctx->d_rec = nullptr;
    ctx->rec_cap = 0;
    BMX_HIP(hipMalloc(&ctx->d_rec, bytes));
    ctx->rec_cap = bytes;
    return BMX_OK;
}

constexpr int kMaxTimedLaunches = 1024;

int ensure_events(bmx_ctx *ctx)
{
    if (ctx->seg_ev) return BMX_OK;
    hipEvent_t *ev = (hipEvent_t *)calloc(2 * kMaxTimedLaunches, sizeof(hipEvent_t));
    if (!ev) return bmx_set_error(BMX_E_OOM, "bmx_walk: host allocation failed");
    for (int k = 0; k < 2 * kMaxTimedLaunches; ++k) {
        hipError_t e = hipEventCreate(&ev[k]);
        if (e != hipSuccess) {
            for (int j = 0; j < k; ++j) (void)hipEventDestroy(ev[j]);
            free(ev);
            return bmx_set_error(BMX_E_DEVICE, "bmx_walk: hipEventCreate: %s", hipGetErrorString(e));
        }
    }
    ctx->seg_ev = ev;
    return BMX_OK;
}

int walk_sorted(bmx_ctx *ctx, const void *img, int img_dtype, const WalkGeom &g, const bmx_seed *seeds,
                int64_t nseeds, int nbrw, int sorw, uint32_t k0, uint32_t k1, uint32_t *hits,
                unsigned long long *acc, int seg, int brick_log2, int *launches)
{
    const int64_t n = nseeds * (int64_t)nbrw;
    BMX_CHECK(n < ((int64_t)1 << 32), "bmx_walk: %lld walkers exceed 2^32", (long long)n);
    SortGeom sg;
    sg.sz = g.sz;
    sg.sy = g.sy;
    sg.b = brick_log2;
    const uint32_t tile = 4u << brick_log2;    // super-tile side in voxels
    sg.nsx = (uint32_t)((g.X + tile - 1) / tile);
    sg.nsy = (uint32_t)((g.Y + tile - 1) / tile);
    const uint32_t nsz = (uint32_t)((g.Z + tile - 1) / tile);
    const size_t nkeys = (size_t)nsz * sg.nsy * sg.nsx * 64;
    int rc = grow(ctx, &ctx->d_state, &ctx->state_cap, (size_t)n * 2 * sizeof(uint4));
    if (rc) return rc;
    rc = grow(ctx, &ctx->d_keys, &ctx->keys_cap, nkeys * sizeof(uint32_t));
    if (rc) return rc;
    rc = ensure_events(ctx);
    if (rc) return rc;
    uint4 *A = (uint4 *)ctx->d_state, *Bst = A + n;
    uint32_t *counts = (uint32_t *)ctx->d_keys;
    const unsigned blocks = bmx_blocks(n, kWalkThreads);
    k_seg_init<<<blocks, 256, 0, ctx->stream>>>(seeds, nseeds, nbrw, A);
    BMX_LAUNCH_CHECK();
    const int nseg = (sorw + seg - 1) / seg;
    for (int k = 0; k < nseg; ++k) {
        const bool last = (k == nseg - 1);
        if (!last) BMX_HIP(hipMemsetAsync(counts, 0, nkeys * sizeof(uint32_t), ctx->stream));
        if (k < kMaxTimedLaunches) BMX_HIP(hipEventRecord(ctx->seg_ev[2 * k], ctx->stream));
        if (img_dtype == BMX_IMG_U8)
            k_seg_walk<uint8_t><<<blocks, kWalkThreads, 0, ctx->stream>>>(
                ctx->d_rec, (const uint8_t *)img, g, seeds, A, n, sorw, seg, k0, k1, hits, acc,
                last ? nullptr : counts, sg);
        else
            k_seg_walk<float><<<blocks, kWalkThreads, 0, ctx->stream>>>(
                ctx->d_rec, (const float *)img, g, seeds, A, n, sorw, seg, k0, k1, hits, acc,
                last ? nullptr : counts, sg);
        BMX_LAUNCH_CHECK();
        if (k < kMaxTimedLaunches) BMX_HIP(hipEventRecord(ctx->seg_ev[2 * k + 1], ctx->stream));
        if (last) break;
        rc = scan_u32(ctx, counts, (int64_t)nkeys);
        if (rc) return rc;
        k_seg_scatter<<<blocks, 256, 0, ctx->stream>>>(A, n, Bst, counts, sg, sorw);
        BMX_LAUNCH_CHECK();
        uint4 *t = A; A = Bst; Bst = t;
    }
    *launches = nseg < kMaxTimedLaunches ? nseg : kMaxTimedLaunches;
    return BMX_OK;
}

}  // namespace

// Reads the accepted-move count and the walk timings of the launch bracketed
// by ctx->ev0 / ev1 (per-round events in ctx->seg_ev for `launches` rounds).
// Round pairs per graph replay of the tiled walker: graph_rounds -1 = host-driven
// rounds (0 pairs), 0 = the default of 8 rounds, k = k rounds rounded up to even.
static int graph_pairs_of(const bmx_walk_opts *opts)
{
    const int gr = opts ? opts->graph_rounds : 0;
    return gr < 0 ? 0 : (gr == 0 ? 4 : (gr + 1) / 2);
}

static int finish_walk(bmx_ctx *ctx, bool run, int launches, uint64_t *accepted, float graph_kernel_ms = -1.0f)
{
    if (accepted) {
        uint64_t a = 0;
        BMX_HIP(hipMemcpyAsync(&a, ctx->d_counter, sizeof(a), hipMemcpyDeviceToHost, ctx->stream));
        BMX_HIP(hipStreamSynchronize(ctx->stream));
        *accepted = a;
        float ms = 0.0f;
        BMX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->last_walk_ms = ms;
        ctx->last_walk_kernel_ms = ms;
        ctx->last_walk_launches = run ? 1 : 0;
        if (graph_kernel_ms >= 0.0f || graph_kernel_ms == -2.0f) {
            // graph-driven rounds: the summed walk kernels, or the whole phase
            if (graph_kernel_ms >= 0.0f) ctx->last_walk_kernel_ms = graph_kernel_ms;
            ctx->last_walk_launches = launches;
        } else if (launches > 0) {
            float sum = 0.0f;
            for (int k = 0; k < launches; ++k) {
                float t = 0.0f;
                BMX_HIP(hipEventElapsedTime(&t, ctx->seg_ev[2 * k], ctx->seg_ev[2 * k + 1]));
                sum += t;
            }
            ctx->last_walk_kernel_ms = sum;
            ctx->last_walk_launches = launches;
        }
    }
    return BMX_OK;
}


// Packed neighbour records of the direct Philox walkers (V * 8 bytes for
// uint8, V * sizeof(RecF32) for float32) into rec.
static int pack_records(bmx_ctx *ctx, const void *img, int img_dtype, const WalkGeom &g, void *rec)
{
    if (img_dtype == BMX_IMG_U8)
        k_pack_u8<<<bmx_blocks(g.V, 256), 256, 0, ctx->stream>>>((const uint8_t *)img, g, (uint64_t *)rec);
    else
        k_pack_f32<<<bmx_blocks(g.V, 256), 256, 0, ctx->stream>>>((const float *)img, g, (RecF32 *)rec);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

// The direct Philox walk over packed records: per-seed weight tables in LDS for
// uint8 images with 8..256 walks per seed, exact weights otherwise.
static int walk_direct_philox(bmx_ctx *ctx, const void *img, int img_dtype, const WalkGeom &g,
                              const bmx_seed *seeds, int64_t nseeds, int nbrw, int sorw, uint32_t k0, uint32_t k1,
                              uint32_t *hits, const void *rec, unsigned long long *acc)
{
    const int64_t nwalkers = nseeds * (int64_t)nbrw;
    if (img_dtype == BMX_IMG_U8 && nbrw >= kLutMinWalks && nbrw <= kWalkThreads) {
        const int spb = kWalkThreads / nbrw;
        const int64_t blocks = (nseeds + spb - 1) / spb;
        BMX_CHECK(blocks < ((int64_t)1 << 31), "bmx_walk: too many walkers");
        k_walk_philox_u8_lut<<<(unsigned)blocks, kWalkThreads, (size_t)spb * 256 * sizeof(float), ctx->stream>>>(
            (const uint64_t *)rec, (const uint8_t *)img, g, seeds, nseeds, nbrw, sorw, spb, k0, k1, hits, acc);
    } else {
        const int64_t blocks = (nwalkers + kWalkThreads - 1) / kWalkThreads;
        BMX_CHECK(blocks < ((int64_t)1 << 31), "bmx_walk: too many walkers");
        if (img_dtype == BMX_IMG_U8)
            k_walk_philox_exp<uint8_t><<<(unsigned)blocks, kWalkThreads, 0, ctx->stream>>>(
                rec, (const uint8_t *)img, g, seeds, nseeds, nbrw, sorw, k0, k1, hits, acc);
        else
            k_walk_philox_exp<float><<<(unsigned)blocks, kWalkThreads, 0, ctx->stream>>>(
                rec, (const float *)img, g, seeds, nseeds, nbrw, sorw, k0, k1, hits, acc);
    }
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

__global__ void k_zero_counter(unsigned long long *c)
{
    if (threadIdx.x == 0) *c = 0ull;
}

extern "C" int bmx_walk_ex(bmx_ctx *ctx, const void *img, int img_dtype, int64_t Z, int64_t Y,
                           int64_t X, const bmx_seed *seeds, int64_t nseeds, int nlabels, int nbrw,
                           int sorw, int rng, uint64_t key, uint32_t *hits, uint64_t *accepted,
                           const bmx_walk_opts *opts)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(img && hits, "bmx_walk: null pointer");
    BMX_CHECK(nseeds == 0 || seeds, "bmx_walk: null seeds");
    BMX_CHECK(img_dtype == BMX_IMG_U8 || img_dtype == BMX_IMG_F32, "bmx_walk: bad image dtype %d",
              img_dtype);
    BMX_CHECK(Z >= 3 && Y >= 3 && X >= 3 && Z * Y * X < (int64_t)1 << 32,
              "bmx_walk: bad dims %lld x %lld x %lld", (long long)Z, (long long)Y, (long long)X);
    BMX_CHECK(nlabels >= 1 && nlabels <= 256, "bmx_walk: nlabels %d not in [1,256]", nlabels);
    BMX_CHECK(nbrw >= 0 && sorw >= 0, "bmx_walk: negative nbrw/sorw");
    BMX_CHECK(nseeds >= 0, "bmx_walk: negative nseeds");
    BMX_CHECK(rng == BMX_RNG_PHILOX || rng == BMX_RNG_REF, "bmx_walk: bad rng %d", rng);
    int algo = opts ? opts->algorithm : BMX_WALK_AUTO;
    int seg = (opts && opts->segment_steps > 0) ? opts->segment_steps : 128;
    int blog = (opts && opts->brick_log2 > 0) ? opts->brick_log2 : 3;
    const int weight = (opts && opts->hit_weight > 0) ? opts->hit_weight : 1;
    BMX_CHECK(!opts || opts->hit_weight >= 0, "bmx_walk: negative hit_weight");
    const int overrides = (opts && opts->seed_overrides) ? 1 : 0;
    const int shape = opts ? opts->tile_shape : 0;
    BMX_CHECK(!opts || (opts->graph_rounds >= -1 && opts->graph_rounds <= 1024),
              "bmx_walk: graph_rounds %d not in [-1, 1024]", opts ? opts->graph_rounds : 0);
    const int graph_pairs = graph_pairs_of(opts);
    BMX_CHECK(algo == BMX_WALK_AUTO || algo == BMX_WALK_DIRECT || algo == BMX_WALK_SORTED ||
                  algo == BMX_WALK_TILED,
              "bmx_walk: bad algorithm %d", algo);
    BMX_CHECK(seg % 4 == 0, "bmx_walk: segment_steps %d must be a multiple of 4", seg);
    BMX_CHECK(blog >= 1 && blog <= 6, "bmx_walk: brick_log2 %d not in [1,6]", blog);
    WalkGeom g{(int)Z, (int)Y, (int)X, (uint32_t)X, (uint32_t)(X * Y), Z * Y * X};
    BMX_HIP(hipMemsetAsync(ctx->d_counter, 0, sizeof(uint64_t), ctx->stream));
    unsigned long long *acc = (unsigned long long *)ctx->d_counter;
    const bool run = nseeds > 0 && nbrw > 0 && sorw > 0;
    const int64_t nwalkers = nseeds * (int64_t)nbrw;
    // automatic choice: weights and per-seed overrides need the tiled walker
    if (algo == BMX_WALK_AUTO)
        algo = (weight > 1 || overrides || nwalkers >= ((int64_t)1 << 16)) ? BMX_WALK_TILED : BMX_WALK_DIRECT;
    if (rng != BMX_RNG_PHILOX) algo = BMX_WALK_DIRECT;
    BMX_CHECK((weight == 1 && !overrides) || (algo == BMX_WALK_TILED && rng == BMX_RNG_PHILOX),
              "bmx_walk: hit_weight / seed_overrides need the tiled walker and the Philox stream");
    BMX_CHECK(nbrw <= 65535, "bmx_walk: nbrw %d above 65535", nbrw);
    if (algo == BMX_WALK_TILED && !(opts && opts->segment_steps > 0)) seg = 256;
    if (run && rng == BMX_RNG_PHILOX && algo != BMX_WALK_TILED) {
        const size_t rec_bytes = (size_t)g.V * (img_dtype == BMX_IMG_U8 ? 8 : sizeof(RecF32));
        rc = ensure_records(ctx, rec_bytes);
        if (rc) return rc;
        rc = pack_records(ctx, img, img_dtype, g, ctx->d_rec);
        if (rc) return rc;
    }
    int launches = 0;
    float graph_ms = -1.0f;
    BMX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (run) {
        const uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
        if (rng == BMX_RNG_PHILOX && algo == BMX_WALK_TILED) {
            rc = ensure_events(ctx);
            if (rc) return rc;
            const bmx_volume vol{0, g.Z, g.Y, g.X, nlabels, 0, 0, (uint64_t)nseeds};
            rc = bmx_walk_tiled(ctx, img, img_dtype, &vol, 1, seeds, nseeds, nlabels, nbrw, sorw, seg,
                                (uint32_t)weight, overrides, k0, k1, hits, acc, shape, graph_pairs,
                                ctx->seg_ev, kMaxTimedLaunches, &launches, &graph_ms);
            if (rc) return rc;
        } else if (rng == BMX_RNG_PHILOX && algo == BMX_WALK_SORTED) {
            rc = walk_sorted(ctx, img, img_dtype, g, seeds, nseeds, nbrw, sorw, k0, k1, hits, acc, seg,
                             blog, &launches);
            if (rc) return rc;
        } else if (rng == BMX_RNG_PHILOX) {
            rc = walk_direct_philox(ctx, img, img_dtype, g, seeds, nseeds, nbrw, sorw, k0, k1, hits, ctx->d_rec, acc);
            if (rc) return rc;
        } else {
            const int64_t blocks = (nseeds + kWalkThreads - 1) / kWalkThreads;
            if (img_dtype == BMX_IMG_U8)
                k_walk_ref<uint8_t><<<(unsigned)blocks, kWalkThreads, 0, ctx->stream>>>(
                    (const uint8_t *)img, g, seeds, nseeds, nbrw, sorw, hits, acc);
            else
                k_walk_ref<float><<<(unsigned)blocks, kWalkThreads, 0, ctx->stream>>>(
                    (const float *)img, g, seeds, nseeds, nbrw, sorw, hits, acc);
            BMX_LAUNCH_CHECK();
        }
    }
    BMX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    return finish_walk(ctx, run, launches, accepted, graph_ms);
}

// The direct Philox walk with its packed records and move counter in the
// library's static scratch buffer (bmx_noalloc_scratch): no device memory is
// allocated.  Layout of the buffer: the move counter at offset 0, the records
// from kNoAllocRecOffset.
extern "C" int bmx_walk_noalloc(bmx_ctx *ctx, const void *img, int img_dtype, int64_t Z, int64_t Y, int64_t X,
                                const bmx_seed *seeds, int64_t nseeds, int nlabels, int nbrw, int sorw,
                                uint64_t key, uint32_t *hits, uint64_t *accepted)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(img && hits, "bmx_walk_noalloc: null pointer");
    BMX_CHECK(nseeds == 0 || seeds, "bmx_walk_noalloc: null seeds");
    BMX_CHECK(img_dtype == BMX_IMG_U8 || img_dtype == BMX_IMG_F32, "bmx_walk_noalloc: bad image dtype %d",
              img_dtype);
    BMX_CHECK(Z >= 3 && Y >= 3 && X >= 3 && Z * Y * X < (int64_t)1 << 32,
              "bmx_walk_noalloc: bad dims %lld x %lld x %lld", (long long)Z, (long long)Y, (long long)X);
    BMX_CHECK(nlabels >= 1 && nlabels <= 256, "bmx_walk_noalloc: nlabels %d not in [1,256]", nlabels);
    BMX_CHECK(nbrw >= 0 && sorw >= 0 && nbrw <= 65535, "bmx_walk_noalloc: nbrw/sorw out of range");
    BMX_CHECK(nseeds >= 0, "bmx_walk_noalloc: negative nseeds");
    WalkGeom g{(int)Z, (int)Y, (int)X, (uint32_t)X, (uint32_t)(X * Y), Z * Y * X};
    const size_t rec_bytes = (size_t)g.V * (img_dtype == BMX_IMG_U8 ? 8 : sizeof(RecF32));
    BMX_CHECK(rec_bytes <= (size_t)BMX_NOALLOC_SCRATCH_BYTES - kNoAllocRecOffset,
              "bmx_walk_noalloc: a %lld-voxel volume needs %zu bytes of scratch, above the %u-byte maximum "
              "(%llu voxels for uint8 images, %llu for float32)", (long long)g.V, rec_bytes + kNoAllocRecOffset,
              (unsigned)BMX_NOALLOC_SCRATCH_BYTES,
              (unsigned long long)((BMX_NOALLOC_SCRATCH_BYTES - kNoAllocRecOffset) / 8),
              (unsigned long long)((BMX_NOALLOC_SCRATCH_BYTES - kNoAllocRecOffset) / sizeof(RecF32)));
    unsigned char *scr = (unsigned char *)bmx_noalloc_scratch(ctx->device);
    BMX_CHECK(scr != nullptr, "bmx_walk_noalloc: no scratch buffer on device %d", ctx->device);
    unsigned long long *acc = (unsigned long long *)scr;
    void *rec = scr + kNoAllocRecOffset;
    k_zero_counter<<<1, 64, 0, ctx->stream>>>(acc);
    BMX_LAUNCH_CHECK();
    const bool run = nseeds > 0 && nbrw > 0 && sorw > 0;
    if (run) {
        rc = pack_records(ctx, img, img_dtype, g, rec);
        if (rc) return rc;
    }
    BMX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (run) {
        rc = walk_direct_philox(ctx, img, img_dtype, g, seeds, nseeds, nbrw, sorw, (uint32_t)key,
                                (uint32_t)(key >> 32), hits, rec, acc);
        if (rc) return rc;
    }
    BMX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    if (accepted) {
        uint64_t a = 0;
        BMX_HIP(hipMemcpyAsync(&a, acc, sizeof(a), hipMemcpyDeviceToHost, ctx->stream));
        BMX_HIP(hipStreamSynchronize(ctx->stream));
        *accepted = a;
        float ms = 0.0f;
        BMX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->last_walk_ms = ms;
        ctx->last_walk_kernel_ms = ms;
        ctx->last_walk_launches = run ? 1 : 0;
    }
    return BMX_OK;
}

extern "C" int bmx_walk(bmx_ctx *ctx, const void *img, int img_dtype, int64_t Z, int64_t Y, int64_t X,
                        const bmx_seed *seeds, int64_t nseeds, int nlabels, int nbrw, int sorw, int rng,
                        uint64_t key, uint32_t *hits, uint64_t *accepted)
{
    return bmx_walk_ex(ctx, img, img_dtype, Z, Y, X, seeds, nseeds, nlabels, nbrw, sorw, rng, key, hits,
                       accepted, nullptr);
}

extern "C" int bmx_walk_batch(bmx_ctx *ctx, const void *img, int img_dtype, const bmx_volume *vols, int nvol,
                              const bmx_seed *seeds, int nbrw, int sorw, uint64_t key, uint32_t *hits,
                              uint64_t *accepted, const bmx_walk_opts *opts)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(img && hits && vols, "bmx_walk_batch: null pointer");
    BMX_CHECK(img_dtype == BMX_IMG_U8 || img_dtype == BMX_IMG_F32, "bmx_walk_batch: bad image dtype %d",
              img_dtype);
    BMX_CHECK(nvol >= 1 && nvol <= (1 << 20), "bmx_walk_batch: nvol %d not in [1, 2^20]", nvol);
    BMX_CHECK(nbrw >= 0 && nbrw <= 65535 && sorw >= 0, "bmx_walk_batch: nbrw / sorw out of range");
    const int algo = opts ? opts->algorithm : BMX_WALK_AUTO;
    BMX_CHECK(algo == BMX_WALK_AUTO || algo == BMX_WALK_TILED, "bmx_walk_batch: algorithm must be tiled");
    BMX_CHECK(!opts || opts->hit_weight >= 0, "bmx_walk_batch: negative hit_weight");
    const int seg = (opts && opts->segment_steps > 0) ? opts->segment_steps : 256;
    BMX_CHECK(seg % 4 == 0, "bmx_walk_batch: segment_steps %d must be a multiple of 4", seg);
    const int weight = (opts && opts->hit_weight > 0) ? opts->hit_weight : 1;
    const int overrides = (opts && opts->seed_overrides) ? 1 : 0;
    const int shape = opts ? opts->tile_shape : 0;
    BMX_CHECK(!opts || (opts->graph_rounds >= -1 && opts->graph_rounds <= 1024),
              "bmx_walk_batch: graph_rounds %d not in [-1, 1024]", opts ? opts->graph_rounds : 0);
    const int graph_pairs = graph_pairs_of(opts);
    uint64_t img_end = 0, seed_end = 0;
    int nlabels = 1;
    for (int v = 0; v < nvol; ++v) {
        const bmx_volume &b = vols[v];
        BMX_CHECK(b.Z >= 3 && b.Y >= 3 && b.X >= 3, "bmx_walk_batch: volume %d: bad dims %d x %d x %d", v, b.Z,
                  b.Y, b.X);
        BMX_CHECK(b.nlabels >= 1 && b.nlabels <= 256, "bmx_walk_batch: volume %d: nlabels %d not in [1,256]", v,
                  b.nlabels);
        BMX_CHECK(b.offset >= img_end, "bmx_walk_batch: volume %d overlaps or precedes volume %d", v, v - 1);
        BMX_CHECK(b.seed_offset == seed_end, "bmx_walk_batch: volume %d: seeds must follow volume %d's", v, v - 1);
        img_end = b.offset + (uint64_t)b.Z * (uint64_t)b.Y * (uint64_t)b.X;
        seed_end = b.seed_offset + b.nseeds;
        BMX_CHECK(img_end < ((uint64_t)1 << 32), "bmx_walk_batch: batch image exceeds 2^32 elements");
        if (b.nlabels > nlabels) nlabels = b.nlabels;
    }
    const int64_t nseeds = (int64_t)seed_end;
    BMX_CHECK(nseeds == 0 || seeds, "bmx_walk_batch: null seeds");
    BMX_HIP(hipMemsetAsync(ctx->d_counter, 0, sizeof(uint64_t), ctx->stream));
    const bool run = nseeds > 0 && nbrw > 0 && sorw > 0;
    int launches = 0;
    float graph_ms = -1.0f;
    BMX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (run) {
        rc = ensure_events(ctx);
        if (rc) return rc;
        rc = bmx_walk_tiled(ctx, img, img_dtype, vols, nvol, seeds, nseeds, nlabels, nbrw, sorw, seg,
                            (uint32_t)weight, overrides, (uint32_t)key, (uint32_t)(key >> 32), hits,
                            (unsigned long long *)ctx->d_counter, shape, graph_pairs, ctx->seg_ev,
                            kMaxTimedLaunches, &launches, &graph_ms);
        if (rc) return rc;
    }
    BMX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    return finish_walk(ctx, run, launches, accepted, graph_ms);
}

extern "C" int bmx_last_walk_stats(bmx_ctx *ctx, float *phase_ms, float *kernel_ms, int *launches)
{
    BMX_CHECK(ctx != nullptr, "bmx_last_walk_stats: null context");
    if (phase_ms) *phase_ms = ctx->last_walk_ms;
    if (kernel_ms) *kernel_ms = ctx->last_walk_kernel_ms;
    if (launches) *launches = ctx->last_walk_launches;
    return BMX_OK;
}
