// bmx_walk_common.h -- device helpers shared by the walker kernels
// (bmx_walk.hip: direct and sorted walkers; bmx_tile.hip: tile-owner walker).
#pragma once
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

// Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Give XCD k one contiguous range of logical blocks so its L2 sees a compact
// region of the volume.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n)
{
    const uint32_t q = n / 8, r = n % 8, xcd = b % 8, idx = b / 8;
    return xcd * q + min(xcd, r) + idx;
}


// Index increment of direction d (+z,-z,+y,-y,+x,-x).
__device__ __forceinline__ uint32_t step_delta(int d, uint32_t sz, uint32_t sy)
{
    const uint32_t mag = d < 2 ? sz : (d < 4 ? sy : 1u);
    return (d & 1) ? (uint32_t)(-(int32_t)mag) : mag;
}

// Direction from the six weights (first d with rand < prefix_d, rand == 0 -> 0).
// The prefix sums are non-decreasing, so the first such d is the number of
// prefixes <= rand; u < 1 keeps rand <= W5.  The count is taken in two
// halves: if rand >= W2, the first three prefixes count and W3, W4 decide
// the rest, else only W0 and W1 can count.  A NaN prefix makes every later
// one and rand NaN, so both forms then count the same leading prefixes.
__device__ __forceinline__ int choose(float w0, float w1, float w2, float w3, float w4, float w5,
                                      float u)
{
    const float W0 = w0;
    const float W1 = w1 + W0;
    const float W2 = w2 + W1;
    const float W3 = w3 + W2;
    const float W4 = w4 + W3;
    const float W5 = w5 + W4;
    const float rand = W5 * u;
    const bool hi = rand >= W2;
    const float a = hi ? W3 : W0, b = hi ? W4 : W1;
    const int cnt = (hi ? 3 : 0) + (rand >= a) + (rand >= b);
    return (rand == 0.0f) ? 0 : cnt;
}

__device__ __forceinline__ float u01(uint32_t word)
{
    return (float)(word >> 8) * 5.9604644775390625e-08f;
}


// Wave-aggregated atomicAdd(&base[key], 1) for every active lane; returns the
// lane's old value plus its rank among the active lanes with the same key.
// Must be reached by all lanes of the wave.
__device__ __forceinline__ uint32_t agg_atomic_inc(uint32_t *base, uint32_t key, bool active)
{
    const int lane = threadIdx.x & 63;
    uint64_t todo = __ballot(active);
    uint32_t result = 0;
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t lkey = __shfl(key, leader, 64);
        const uint64_t grp = __ballot(active && key == lkey) & todo;
        uint32_t old = 0;
        if (lane == leader) old = atomicAdd(base + lkey, (uint32_t)__popcll(grp));
        old = __shfl(old, leader, 64);
        if ((grp >> lane) & 1ull) result = old + (uint32_t)__popcll(grp & ((1ull << lane) - 1ull));
        todo &= ~grp;
    }
    return result;
}


// Exclusive scan of n uint32 counts in place by one block of 1024 threads.
// In-place exclusive scan of counts[0..n) in three passes: per-block totals of
// kScanTile elements, a scan of those totals (one block), and the block-local
// scans plus their offsets.  Loads are coalesced (4 consecutive per thread).
constexpr int kScanThreads = 1024;
constexpr int kScanTile = 4 * kScanThreads;

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *part, uint32_t *total)
{
    // inclusive scan within each wave, then across the 16 wave totals
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) part[wave] = x;
    __syncthreads();
    if (threadIdx.x < 64) {
        const int nw = kScanThreads / 64;
        uint32_t w = threadIdx.x < nw ? part[threadIdx.x] : 0u;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(w, off, 64);
            if ((int)threadIdx.x >= off) w += y;
        }
        if (threadIdx.x < nw) part[threadIdx.x] = w;
        if (threadIdx.x == nw - 1) *total = w;
    }
    __syncthreads();
    const uint32_t before = wave > 0 ? part[wave - 1] : 0u;
    return before + x - v;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_totals(const uint32_t *__restrict__ counts, int64_t n,
                                                              uint32_t *__restrict__ totals)
{
    __shared__ uint32_t part[kScanThreads / 64];
    __shared__ uint32_t total;
    const int64_t base = (int64_t)blockIdx.x * kScanTile + 4 * threadIdx.x;
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (base + q < n) s += counts[base + q];
    (void)block_excl_scan(s, part, &total);
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_top(uint32_t *__restrict__ totals, int64_t nb)
{
    __shared__ uint32_t part[kScanThreads / 64];
    __shared__ uint32_t total;
    uint32_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += kScanThreads) {
        const int64_t b = b0 + threadIdx.x;
        const uint32_t v = b < nb ? totals[b] : 0u;
        const uint32_t ex = block_excl_scan(v, part, &total);
        if (b < nb) totals[b] = carry + ex;
        carry += total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_apply(uint32_t *__restrict__ counts, int64_t n,
                                                             const uint32_t *__restrict__ totals)
{
    __shared__ uint32_t part[kScanThreads / 64];
    __shared__ uint32_t total;
    const int64_t base = (int64_t)blockIdx.x * kScanTile + 4 * threadIdx.x;
    uint32_t v[4];
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[q] = base + q < n ? counts[base + q] : 0u;
        s += v[q];
    }
    uint32_t run = totals[blockIdx.x] + block_excl_scan(s, part, &total);
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (base + q < n) {
            counts[base + q] = run;
            run += v[q];
        }
}


// Exclusive scan of counts[0..n) on the context stream (block partials in ctx->d_scan).
inline int scan_u32(bmx_ctx *ctx, uint32_t *counts, int64_t n)
{
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    int rc = grow(ctx, &ctx->d_scan, &ctx->scan_cap, (size_t)(nb > 0 ? nb : 1) * sizeof(uint32_t));
    if (rc) return rc;
    if (n <= 0) return BMX_OK;
    uint32_t *totals = (uint32_t *)ctx->d_scan;
    k_scan_totals<<<(unsigned)nb, kScanThreads, 0, ctx->stream>>>(counts, n, totals);
    BMX_LAUNCH_CHECK();
    k_scan_top<<<1, kScanThreads, 0, ctx->stream>>>(totals, nb);
    BMX_LAUNCH_CHECK();
    k_scan_apply<<<(unsigned)nb, kScanThreads, 0, ctx->stream>>>(counts, n, totals);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

}  // namespace
