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
// Hits are uint32 [label][V], one global atomic add per accepted move.
#include "bmx_internal.h"

namespace {

constexpr int kWalkThreads = 256;

struct WalkGeom {
    int Z, Y, X;
    uint32_t sy, sz;
    int64_t V;
};

__device__ __forceinline__ void wave_add_u64(unsigned long long *dst, uint64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

// One step of the walk shared by both streams.  Returns true if the move is
// accepted (and then has updated i, z, y, x).
template <typename T>
__device__ __forceinline__ int direction(const T *__restrict__ img, const WalkGeom &g, uint32_t i,
                                         float B, float div1, float u)
{
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
    const float rand = W5 * u;
    // First d with rand < W_d; the prefix sums are non-decreasing, so this is
    // the number of prefixes <= rand.  rand <= W5 holds because u < 1.
    const int cnt = (rand >= W0) + (rand >= W1) + (rand >= W2) + (rand >= W3) + (rand >= W4);
    return (rand == 0.0f) ? 0 : cnt;
}

template <typename T>
__global__ __launch_bounds__(kWalkThreads) void k_walk_philox(
    const T *__restrict__ img, WalkGeom g, const bmx_seed *__restrict__ seeds, int64_t nseeds,
    int nbrw, int sorw, uint32_t k0, uint32_t k1, uint32_t *__restrict__ hits,
    unsigned long long *accepted)
{
    const int64_t tid = (int64_t)blockIdx.x * kWalkThreads + threadIdx.x;
    const int64_t s = tid / nbrw;
    uint64_t acc = 0;
    if (s < nseeds) {
        const int w = (int)(tid - s * nbrw);
        const bmx_seed sd = seeds[s];
        const uint32_t p0 = sd.pos;
        const int z0 = (int)(p0 / g.sz);
        const uint32_t r0 = p0 - (uint32_t)z0 * g.sz;
        const int y0 = (int)(r0 / g.sy);
        const int x0 = (int)(r0 - (uint32_t)y0 * g.sy);
        const float B = bmx_ld(img, (int64_t)p0);
        const float div1 = 1.0f / (2.0f * sd.var);
        uint32_t *__restrict__ h = hits + (int64_t)sd.label * g.V;
        uint32_t i = p0;
        int z = z0, y = y0, x = x0;
        for (int sb = 0; sb < sorw; sb += 4) {
            const bmx_u4 r = bmx_philox(p0, 0u, (uint32_t)w, (uint32_t)(sb >> 2), k0, k1);
            const uint32_t rr[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (sb + j < sorw) {
                    const float u = (float)(rr[j] >> 8) * 5.9604644775390625e-08f;
                    const int d = direction(img, g, i, B, div1, u);
                    const int dz = (d == 0) - (d == 1);
                    const int dy = (d == 2) - (d == 3);
                    const int dx = (d == 4) - (d == 5);
                    const int zn = z + dz, yn = y + dy, xn = x + dx;
                    if (zn > 0 && zn < g.Z - 1 && yn > 0 && yn < g.Y - 1 && xn > 0 && xn < g.X - 1) {
                        z = zn; y = yn; x = xn;
                        i = (uint32_t)z * g.sz + (uint32_t)y * g.sy + (uint32_t)x;
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
    const int64_t s = (int64_t)blockIdx.x * kWalkThreads + threadIdx.x;
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
                float p1 = a12 * s11 - a13n * s10;
                int64_t kk = (int64_t)(p1 / m1);
                p1 -= (float)kk * m1;
                if (p1 < 0.0f) p1 += m1;
                s10 = s11; s11 = s12; s12 = p1;
                float p2 = a21 * s22 - a23n * s20;
                kk = (int64_t)(p2 / m2);
                p2 -= (float)kk * m2;
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

}  // namespace

extern "C" int bmx_walk(bmx_ctx *ctx, const void *img, int img_dtype, int64_t Z, int64_t Y, int64_t X,
                        const bmx_seed *seeds, int64_t nseeds, int nlabels, int nbrw, int sorw, int rng,
                        uint64_t key, uint32_t *hits, uint64_t *accepted)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(img && hits, "bmx_walk: null pointer");
    BMX_CHECK(nseeds == 0 || seeds, "bmx_walk: null seeds");
    BMX_CHECK(img_dtype == BMX_IMG_U8 || img_dtype == BMX_IMG_F32, "bmx_walk: bad image dtype %d",
              img_dtype);
    BMX_CHECK(Z >= 3 && Y >= 3 && X >= 3 && Z * Y * X < (int64_t)1 << 32,
              "bmx_walk: bad dims %lld x %lld x %lld", (long long)Z, (long long)Y, (long long)X);
    BMX_CHECK(nlabels >= 1 && nlabels <= 255, "bmx_walk: nlabels %d not in [1,255]", nlabels);
    BMX_CHECK(nbrw >= 0 && sorw >= 0, "bmx_walk: negative nbrw/sorw");
    BMX_CHECK(nseeds >= 0, "bmx_walk: negative nseeds");
    BMX_CHECK(rng == BMX_RNG_PHILOX || rng == BMX_RNG_REF, "bmx_walk: bad rng %d", rng);
    WalkGeom g{(int)Z, (int)Y, (int)X, (uint32_t)X, (uint32_t)(X * Y), Z * Y * X};
    BMX_HIP(hipMemsetAsync(ctx->d_counter, 0, sizeof(uint64_t), ctx->stream));
    unsigned long long *acc = (unsigned long long *)ctx->d_counter;
    BMX_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (nseeds > 0 && nbrw > 0 && sorw > 0) {
        if (rng == BMX_RNG_PHILOX) {
            const int64_t nthreads = nseeds * (int64_t)nbrw;
            const int64_t blocks = (nthreads + kWalkThreads - 1) / kWalkThreads;
            BMX_CHECK(blocks < ((int64_t)1 << 31), "bmx_walk: too many walkers");
            const uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
            if (img_dtype == BMX_IMG_U8)
                k_walk_philox<uint8_t><<<(unsigned)blocks, kWalkThreads, 0, ctx->stream>>>(
                    (const uint8_t *)img, g, seeds, nseeds, nbrw, sorw, k0, k1, hits, acc);
            else
                k_walk_philox<float><<<(unsigned)blocks, kWalkThreads, 0, ctx->stream>>>(
                    (const float *)img, g, seeds, nseeds, nbrw, sorw, k0, k1, hits, acc);
        } else {
            const int64_t blocks = (nseeds + kWalkThreads - 1) / kWalkThreads;
            if (img_dtype == BMX_IMG_U8)
                k_walk_ref<uint8_t><<<(unsigned)blocks, kWalkThreads, 0, ctx->stream>>>(
                    (const uint8_t *)img, g, seeds, nseeds, nbrw, sorw, hits, acc);
            else
                k_walk_ref<float><<<(unsigned)blocks, kWalkThreads, 0, ctx->stream>>>(
                    (const float *)img, g, seeds, nseeds, nbrw, sorw, hits, acc);
        }
        BMX_LAUNCH_CHECK();
    }
    BMX_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    if (accepted) {
        uint64_t a = 0;
        BMX_HIP(hipMemcpyAsync(&a, ctx->d_counter, sizeof(a), hipMemcpyDeviceToHost, ctx->stream));
        BMX_HIP(hipStreamSynchronize(ctx->stream));
        *accepted = a;
        float ms = 0.0f;
        BMX_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->last_walk_ms = ms;
    }
    return BMX_OK;
}
