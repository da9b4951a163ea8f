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
// prefixes <= rand; u < 1 keeps rand <= W5.
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
    const int cnt = (rand >= W0) + (rand >= W1) + (rand >= W2) + (rand >= W3) + (rand >= W4);
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
__global__ __launch_bounds__(1024) void k_scan_u32(uint32_t *counts, int64_t n)
{
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (n + 1023) / 1024;
    const int64_t lo = t * per, hi = (lo + per < n) ? lo + per : n;
    uint32_t s = 0;
    for (int64_t i = lo; i < hi; ++i) s += counts[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t v = (t >= off) ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (int64_t i = lo; i < hi; ++i) {
        const uint32_t c = counts[i];
        counts[i] = run;
        run += c;
    }
}


template <typename P>
int grow(bmx_ctx *ctx, P **ptr, size_t *cap, size_t bytes)
{
    if (*cap >= bytes) return BMX_OK;
    if (*ptr) {
        BMX_HIP(hipStreamSynchronize(ctx->stream));
        BMX_HIP(hipFree(*ptr));
    }
    *ptr = nullptr;
    *cap = 0;
    BMX_HIP(hipMalloc((void **)ptr, bytes));
    *cap = bytes;
    return BMX_OK;
}


}  // namespace
