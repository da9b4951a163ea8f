// bmx_internal.h -- shared pieces of libbmx_hip: context, error reporting and
// the device arithmetic every walker uses (Philox4x32-10, exp, weights).
//
// Arithmetic contract: all HIP sources are built with -ffp-contract=off, so
// every float expression is evaluated as the IEEE binary32 operations written,
// in the order written; fused multiply-adds appear only as __builtin_fmaf.
// The CPU restatement in oracle/bmx_oracle.c follows the same contract, which
// is what makes walker hit counts bit-exact between the two.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/bmx.h"

struct bmx_ctx {
    int device;
    hipStream_t stream;
    hipEvent_t ev0, ev1;     // bracket the last walker launch (bmx_last_walk_ms)
    float last_walk_ms;
    uint64_t *d_counter;     // 16 x uint64 scratch: accepted moves, seed counts, tiled-walk round control
    void *d_small;           // kSmallBytes of scratch: label-class counts and boxes, min/max, lookup tables
    int64_t *d_block_counts; // seed compaction scratch
    int64_t block_counts_cap;
    void *d_rec;             // packed neighbour records of the walker
    size_t rec_cap;
    void *d_state;           // sorted walker: two walker-state arrays
    size_t state_cap;
    void *d_keys;            // sorted / tiled walker: per-key counts / cursors
    size_t keys_cap;
    void *d_buckets;         // tiled walker: non-empty (key, offset, count) lists
    size_t buckets_cap;
    void *d_scan;            // exclusive-scan block partials
    size_t scan_cap;
    void *d_vols;            // tiled walker: volume table of the walk
    size_t vols_cap;
    void *d_wseeds;          // tiled walker: per-seed constants of the walk
    size_t wseeds_cap;
    void *d_netpart;         // conv3d: float32 partial sums of a two-launch convolution
    size_t netpart_cap;
    hipEvent_t *seg_ev;      // sorted walker: event pairs around each walk launch
    float last_walk_kernel_ms;
    int last_walk_launches;
    hipStream_t aux[2];      // smoothing: two more streams (default priority), made on first use
    hipEvent_t aux_ev[3];    // smoothing: ordering events between the streams
};

// Creates ctx->aux and ctx->aux_ev on first use.
int bmx_aux_streams(bmx_ctx *ctx);

// Layout of bmx_ctx::d_small (bmx_ingest.hip): counts and boxes of the label
// classes from offset 0 (BMX_LABEL_CLASSES x 8 + x 24 bytes), two 256-byte
// tables from kSmallLutOffset.
constexpr size_t kSmallLutOffset = 8448;
constexpr size_t kSmallBytes = 16384;
static_assert(BMX_LABEL_CLASSES * 32 <= kSmallLutOffset && kSmallLutOffset + 512 <= kSmallBytes,
              "d_small layout");

// Error reporting ----------------------------------------------------------
int bmx_set_error(int code, const char *fmt, ...);

#define BMX_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess)                                                      \
            return bmx_set_error(e_ == hipErrorOutOfMemory ? BMX_E_OOM : BMX_E_DEVICE, \
                                 "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                 __FILE__, __LINE__);                              \
    } while (0)

#define BMX_CHECK(cond, ...)                                                       \
    do {                                                                           \
        if (!(cond)) return bmx_set_error(BMX_E_INVALID, __VA_ARGS__);             \
    } while (0)

#define BMX_LAUNCH_CHECK() BMX_HIP(hipGetLastError())

int bmx_set_device(bmx_ctx *ctx);

// Grows a context buffer to at least `bytes` (contents not kept).
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

// Device address of the static scratch buffer (BMX_NOALLOC_SCRATCH_BYTES) on
// `device`, resolved once per device by bmx_open; nullptr before that.  Used
// by bmx_walk_noalloc: the move counter at offset 0, records from
// kNoAllocRecOffset.
void *bmx_noalloc_scratch(int device);
constexpr size_t kNoAllocRecOffset = 256;

// Tile-owner walker (bmx_tile.hip), called by bmx_walk_ex (one volume) and
// bmx_walk_batch.  vols is a host array of nvol validated volumes.  On return
// *launches = walk rounds; *kernel_ms = -1 for host-driven rounds (walk
// launches bracketed by ev[2k], ev[2k+1]), else the summed walk-kernel time of
// the graph-driven rounds (device wall clock), or -2 where it is unknown.
int bmx_walk_tiled(bmx_ctx *ctx, const void *img, int img_dtype, const bmx_volume *vols, int nvol,
                   const bmx_seed *seeds, int64_t nseeds, int nlabels, int nbrw, int sorw, int cap,
                   uint32_t weight, int overrides, uint32_t k0, uint32_t k1, uint32_t *hits,
                   unsigned long long *acc, int shape, int graph_pairs, hipEvent_t *ev, int max_timed,
                   int *launches, float *kernel_ms);

// Device arithmetic -----------------------------------------------------------

// Philox4x32-10 (Salmon et al., SC'11): 10 rounds of two 32x32->64 multiplies
// with the published multipliers and Weyl key increments.
struct bmx_u4 { uint32_t x, y, z, w; };

__device__ __forceinline__ bmx_u4 bmx_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                             uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32->64 product per multiplier (v_mad_u64_u32) gives both halves
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return bmx_u4{c0, c1, c2, c3};
}

// exp for binary32: Cody-Waite reduction x = n ln2 + r (ln2 split in a hi part
// with 16 significant bits and a lo part, applied with fused multiply-adds),
// e^r by the degree-5 polynomial 1 + r + c2 r^2 + ... + c5 r^5 (near-minimax
// relative-error fit on |r| <= ln2/2, fit error 1.06e-7 with these float32
// coefficients; Horner with fused multiply-adds), 2^n applied with ldexp
// (IEEE gradual underflow below FLT_MIN).  n = round(x log2 e) comes from one
// fused multiply-add onto 1.5 * 2^23, whose low mantissa bits then hold n as
// an integer (the float n is that sum minus the shift).  Max error on
// [-104, 0] is 2.2 ulp against a long double exp.  Valid for x <= 88 and
// x >= -1000 (weights: x >= -65025/2, where every result below -104 is 0);
// the CPU oracle clamps outside.
__device__ __forceinline__ float bmx_expf(float x)
{
    const float LOG2E = 1.44269502925872802734375f;
    const float LN2_HI = 0.693145751953125f;
    const float LN2_LO = 1.428606765330187045e-06f;
    const float SHIFT = 12582912.0f;
    const float s = __builtin_fmaf(x, LOG2E, SHIFT);
    const float n = s - SHIFT;
    const int ni = (int)(__float_as_uint(s) - __float_as_uint(SHIFT));
    float r = __builtin_fmaf(-n, LN2_HI, x);
    r = __builtin_fmaf(-n, LN2_LO, r);
    float p = 0.00831252523f;
    p = __builtin_fmaf(p, r, 0.0418901145f);
    p = __builtin_fmaf(p, r, 0.166671142f);
    p = __builtin_fmaf(p, r, 0.499992311f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    return __builtin_amdgcn_ldexpf(p, ni);
}

// Transition weight of the reference (pyopencl_small.py:187-190):
// exp(-(B-A)^2 * div1), evaluated as ((-t)*t)*div1.
__device__ __forceinline__ float bmx_weight(float B, float A, float div1)
{
    const float t = B - A;
    return bmx_expf(((-t) * t) * div1);
}

template <typename T> __device__ __forceinline__ float bmx_ld(const T *p, int64_t i);
template <> __device__ __forceinline__ float bmx_ld<uint8_t>(const uint8_t *p, int64_t i)
{
    return (float)p[i];
}
template <> __device__ __forceinline__ float bmx_ld<float>(const float *p, int64_t i)
{
    return p[i];
}

// Grid helpers
static inline unsigned bmx_blocks(int64_t n, int threads)
{
    return (unsigned)((n + threads - 1) / threads);
}

// Element-wise kernels over ranges that may pass 2^32 elements stride over
// them: one dispatch covers at most 2^32-1 work-items per grid dimension (the
// dispatch packet's grid size is 32-bit), so such launches are capped at
// kMaxStrideBlocks blocks and each thread loops with bmx_gtid/bmx_gstride.
constexpr int64_t kMaxStrideBlocks = (int64_t)1 << 20;
static inline unsigned bmx_stride_blocks(int64_t n, int threads)
{
    const int64_t b = (n + threads - 1) / threads;
    return (unsigned)(b < 1 ? 1 : (b < kMaxStrideBlocks ? b : kMaxStrideBlocks));
}
__device__ __forceinline__ int64_t bmx_gtid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ int64_t bmx_gstride() { return (int64_t)gridDim.x * blockDim.x; }
