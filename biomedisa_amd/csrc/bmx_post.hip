// bmx_post.hip -- post-processing of a label volume: outlier removal (clean)
// and hole filling (fill), biomedisa/features/remove_outlier.py:41-139.
//
// Per label k present in the volume (in ascending order, skipping the smallest
// present value, as `unique(...)[1:]` does), on the box of its voxels grown by
// one voxel and clamped (reduce_blocksize, :41-58):
//   clean: 6-connected components of (v == k); components larger than
//          threshold x the largest keep k, the rest of k becomes 0 (:60-95).
//   fill:  6-connected components of (v != k) in the box; the largest entry of
//          [voxels of k, component sizes...] (first maximum, components in the
//          raster order of their first voxel) is "outside"; every other
//          component becomes k unless it is larger than threshold x (voxels of
//          k); the voxels of k outside the result become 0 (:97-139).  fill
//          works on the volume as changed by the labels before it.
//
// Components: union-find over voxel indices of the box with parents that only
// decrease (a union links the larger root below the smaller with atomicMin),
// so a component's root is its first voxel in raster order - the order in which
// scipy.ndimage.label numbers components.  Every chain strictly decreases, so
// finds end; a union's pair of roots strictly decreases in sum at each retry,
// so unions end.  Parent reads inside the merge kernel are relaxed atomics at
// agent scope (the XCDs' L2 caches are not coherent with each other).
#include "bmx_internal.h"

#include <climits>
#include <vector>

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;

struct BoxG {
    int z0, y0, x0;          // box origin in the volume
    int Zb, Yb, Xb;          // box extents
    int64_t Vb;
    int64_t sy, sz;          // volume strides
};

__device__ __forceinline__ int64_t vol_index(const BoxG &b, int64_t i)
{
    const int z = (int)(i / ((int64_t)b.Yb * b.Xb));
    const int64_t r = i - (int64_t)z * b.Yb * b.Xb;
    const int y = (int)(r / b.Xb);
    const int x = (int)(r - (int64_t)y * b.Xb);
    return (int64_t)(b.z0 + z) * b.sz + (int64_t)(b.y0 + y) * b.sy + (b.x0 + x);
}

__device__ __forceinline__ uint32_t ld_parent(const uint32_t *par, uint32_t i)
{
    return __hip_atomic_load(par + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t find_root(const uint32_t *par, uint32_t i)
{
    uint32_t p = ld_parent(par, i);
    while (p != i) {
        i = p;
        p = ld_parent(par, i);
    }
    return i;
}

// find_root with path halving: each visited voxel's parent is lowered to its
// grandparent with atomicMin.  A grandparent is an ancestor, parents only
// decrease, and an ancestor stays an ancestor, so the tree keeps its root and
// no concurrent link is undone; chains shorten as the merge proceeds.
__device__ __forceinline__ uint32_t find_halve(uint32_t *par, uint32_t i)
{
    uint32_t p = ld_parent(par, i);
    while (p != i) {
        const uint32_t gp = ld_parent(par, p);
        if (gp < p) atomicMin(par + i, gp);
        i = p;
        p = ld_parent(par, i);
    }
    return i;
}

__device__ __forceinline__ void unite(uint32_t *par, uint32_t a, uint32_t b)
{
    while (true) {
        a = find_halve(par, a);
        b = find_halve(par, b);
        if (a == b) return;
        if (a < b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        const uint32_t old = atomicMin(par + a, b);
        if (old == a) return;
        a = old;
    }
}

// Foreground of the box: v == k (clean) or v != k (fill).
template <bool INV>
__device__ __forceinline__ bool fg(const uint8_t *v, int64_t vi, uint8_t k)
{
    return INV ? v[vi] != k : v[vi] == k;
}

// One wave per x-row of the box, 64 voxels at a time: every foreground voxel
// points at the first voxel of its run along x (the run's smallest index, a
// root), found from the wave's ballot of non-foreground lanes below it, or
// carried over from the previous 64 voxels when the run crosses them.
template <bool INV>
__global__ __launch_bounds__(256) void k_cc_init(const uint8_t *__restrict__ vol, BoxG b, uint8_t k,
                                                 uint32_t *__restrict__ par, uint32_t *__restrict__ size)
{
    const int lane = threadIdx.x & 63;
    const int64_t rows = (int64_t)b.Zb * b.Yb;
    const int64_t wave0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    for (int64_t row = wave0; row < rows; row += nwaves) {   // the same row in every lane of the wave
        const int64_t i0 = row * b.Xb;
        const int64_t v0 = vol_index(b, i0);
        uint32_t carry = kNone;
        for (int x0 = 0; x0 < b.Xb; x0 += 64) {
            const int x = x0 + lane;
            const bool f = x < b.Xb && fg<INV>(vol, v0 + x, k);
            const uint64_t gaps = ~__ballot(f);                    // lanes that end a run
            const uint64_t below = gaps & ((1ull << lane) - 1ull);  // gaps below this lane
            uint32_t start;
            if (below) start = (uint32_t)(i0 + x0 + (63 - __clzll(below)) + 1);
            else start = carry != kNone ? carry : (uint32_t)(i0 + x0);
            if (x < b.Xb) {
                par[i0 + x] = f ? start : kNone;
                size[i0 + x] = 0u;
            }
            // the run through lane 63 continues into the next 64 voxels
            const uint32_t last = __shfl(start, 63, 64);
            carry = (gaps >> 63) & 1ull ? kNone : last;
        }
    }
}

// Unions with the backward neighbours in y and z that are foreground (x
// neighbours are already joined by the row runs); a voxel whose -y (-z)
// neighbour continues the same pair of runs as its -x neighbour skips it.
__global__ __launch_bounds__(256) void k_cc_merge(BoxG b, uint32_t *par)
{
    const int64_t plane = (int64_t)b.Yb * b.Xb;
    for (int64_t i = bmx_gtid(); i < b.Vb; i += bmx_gstride()) {
        if (par[i] == kNone) continue;
        const int z = (int)(i / plane);
        const int64_t r = i - (int64_t)z * plane;
        const int y = (int)(r / b.Xb);
        const int x = (int)(r - (int64_t)y * b.Xb);
        const bool left = x > 0 && par[i - 1] != kNone;
        if (y > 0 && par[i - b.Xb] != kNone && !(left && par[i - b.Xb - 1] != kNone))
            unite(par, (uint32_t)i, (uint32_t)(i - b.Xb));
        if (z > 0 && par[i - plane] != kNone && !(left && par[i - plane - 1] != kNone))
            unite(par, (uint32_t)i, (uint32_t)(i - plane));
    }
}

// Every parent is a run start (a voxel's parent starts as its run's first
// voxel; unions link roots, which are run starts; halving moves a parent to a
// grandparent).  Run starts are first pointed straight at their roots, so that
// every voxel then reaches its root in two steps.
__global__ __launch_bounds__(256) void k_cc_compress(BoxG b, uint32_t *par)
{
    for (int64_t i = bmx_gtid(); i < b.Vb; i += bmx_gstride()) {
        if (par[i] == kNone) continue;
        const int x = (int)(i % b.Xb);
        if (x > 0 && par[i - 1] != kNone) continue;   // not a run start
        const uint32_t r = find_halve(par, (uint32_t)i);
        atomicMin(par + i, r);
    }
}

// Roots count their voxels.  Neighbouring voxels mostly share a root, so the
// lanes of a wave add per distinct root, one atomic each (a single root of a
// large component would otherwise take one atomic per voxel, serialised on one
// address).  The loop bound is uniform per block, as the ballots need every
// lane of the wave.
__global__ __launch_bounds__(256) void k_cc_flatten(BoxG b, uint32_t *par, uint32_t *size)
{
    const int lane = threadIdx.x & 63;
    for (int64_t b0 = (int64_t)blockIdx.x * blockDim.x; b0 < b.Vb; b0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = b0 + threadIdx.x;
        const bool f = i < b.Vb && par[i] != kNone;
        const uint32_t r = f ? find_root(par, (uint32_t)i) : kNone;
        uint64_t todo = __ballot(f);
        while (todo) {
            const int leader = __ffsll((unsigned long long)todo) - 1;
            const uint32_t lr = __shfl(r, leader, 64);
            const uint64_t grp = __ballot(f && r == lr) & todo;
            if (lane == leader) atomicAdd(size + lr, (uint32_t)__popcll(grp));
            todo &= ~grp;
        }
    }
}

__global__ __launch_bounds__(256) void k_cc_relabel(BoxG b, uint32_t *par)
{
    for (int64_t i = bmx_gtid(); i < b.Vb; i += bmx_gstride()) {
        if (par[i] == kNone) continue;
        par[i] = find_root(par, (uint32_t)i);
    }
}

// stats[0] = largest component, stats[1] = smallest root among the largest
// components, stats[2] = foreground voxels.
__global__ __launch_bounds__(256) void k_cc_max(BoxG b, const uint32_t *__restrict__ size, uint32_t *stats)
{
    for (int64_t i = bmx_gtid(); i < b.Vb; i += bmx_gstride()) {
        const uint32_t s = size[i];
        if (s) {
            atomicMax(stats + 0, s);
            atomicAdd(stats + 2, s);
        }
    }
}

__global__ __launch_bounds__(256) void k_cc_argmax(BoxG b, const uint32_t *__restrict__ size, uint32_t *stats)
{
    const uint32_t m = stats[0];
    for (int64_t i = bmx_gtid(); i < b.Vb; i += bmx_gstride())
        if (m && size[i] == m) atomicMin(stats + 1, (uint32_t)i);
}

// Decisions, in double as the reference compares in float64.
//   clean: keep a component of k when size > t * largest.
//   fill:  n_k = box voxels of k; "outside" = the first maximum of
//          [n_k, component sizes in root order]; a component becomes k when it
//          is not outside and size <= t * n_k; when outside is the voxels of k
//          themselves, they leave the result.
template <bool FILL>
__global__ __launch_bounds__(256) void k_cc_apply(uint8_t *__restrict__ out, BoxG b,
                                                  uint8_t k, double t, const uint32_t *__restrict__ par,
                                                  const uint32_t *__restrict__ size,
                                                  const uint32_t *__restrict__ stats)
{
    const uint32_t largest = stats[0], first_largest = stats[1], nfg = stats[2];
    for (int64_t i = bmx_gtid(); i < b.Vb; i += bmx_gstride()) {
        const int64_t vi = vol_index(b, i);
        const uint32_t p = par[i];
        if (!FILL) {
            if (p == kNone) continue;   // not k
            out[vi] = ((double)size[p] > t * (double)largest) ? k : (uint8_t)0;
        } else {
            const uint32_t nk = (uint32_t)b.Vb - nfg;   // voxels of k in the box
            const bool k_outside = nk >= largest;       // first maximum: index 0 wins ties
            if (p == kNone) {                            // a voxel of k
                if (k_outside) out[vi] = 0;
                continue;
            }
            const bool outside = !k_outside && p == first_largest;
            const bool large = (double)size[p] > t * (double)nk;
            if (!outside && !large) out[vi] = k;
        }
    }
}

// Histogram of the label values (for the set of present labels).
__global__ __launch_bounds__(256) void k_hist256(const uint8_t *__restrict__ v, int64_t V, uint32_t *hist)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0u;
    __syncthreads();
    for (int64_t i = bmx_gtid(); i < V; i += bmx_gstride()) atomicAdd(&h[v[i]], 1u);
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(hist + threadIdx.x, h[threadIdx.x]);
}

// Box of the voxels equal to k: [zmin, zmax, ymin, ymax, xmin, xmax].
__global__ __launch_bounds__(256) void k_label_box(const uint8_t *__restrict__ v, int Z, int Y, int X, uint8_t k,
                                                   int *box)
{
    const int64_t sz = (int64_t)X * Y, V = sz * Z;
    int b[6] = {INT_MAX, -1, INT_MAX, -1, INT_MAX, -1};
    for (int64_t i = bmx_gtid(); i < V; i += bmx_gstride()) {
        if (v[i] != k) continue;
        const int z = (int)(i / sz);
        const int64_t r = i - (int64_t)z * sz;
        const int y = (int)(r / X), x = (int)(r - (int64_t)y * X);
        b[0] = min(b[0], z); b[1] = max(b[1], z);
        b[2] = min(b[2], y); b[3] = max(b[3], y);
        b[4] = min(b[4], x); b[5] = max(b[5], x);
    }
    if (b[1] >= 0) {
        atomicMin(box + 0, b[0]); atomicMax(box + 1, b[1]);
        atomicMin(box + 2, b[2]); atomicMax(box + 3, b[3]);
        atomicMin(box + 4, b[4]); atomicMax(box + 5, b[5]);
    }
}

__global__ void k_post_init(int *box, uint32_t *stats)
{
    const int i = threadIdx.x;
    if (i < 6) box[i] = (i & 1) ? -1 : INT_MAX;
    if (i == 0) {
        stats[0] = 0u;
        stats[1] = kNone;
        stats[2] = 0u;
    }
}

struct PostBufs {
    uint32_t *par = nullptr, *size = nullptr, *stats = nullptr, *hist = nullptr;
    int *box = nullptr;
};

template <bool FILL>
int post_label(bmx_ctx *ctx, uint8_t *vol, int Z, int Y, int X, uint8_t k, double t, PostBufs &pb)
{
    const int64_t V = (int64_t)Z * Y * X;
    k_post_init<<<1, 64, 0, ctx->stream>>>(pb.box, pb.stats);
    BMX_LAUNCH_CHECK();
    k_label_box<<<bmx_stride_blocks(V, 256), 256, 0, ctx->stream>>>(vol, Z, Y, X, k, pb.box);
    BMX_LAUNCH_CHECK();
    int h[6];
    BMX_HIP(hipMemcpyAsync(h, pb.box, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    BMX_HIP(hipStreamSynchronize(ctx->stream));
    if (h[1] < 0) return BMX_OK;   // k no longer present (fill may have covered it)
    BoxG b;
    b.z0 = max(h[0] - 1, 0);
    b.y0 = max(h[2] - 1, 0);
    b.x0 = max(h[4] - 1, 0);
    b.Zb = min(h[1] + 1, Z - 1) + 1 - b.z0;
    b.Yb = min(h[3] + 1, Y - 1) + 1 - b.y0;
    b.Xb = min(h[5] + 1, X - 1) + 1 - b.x0;
    b.Vb = (int64_t)b.Zb * b.Yb * b.Xb;
    b.sy = X;
    b.sz = (int64_t)X * Y;
    const unsigned grid = bmx_stride_blocks(b.Vb, 256);
    k_cc_init<FILL><<<grid, 256, 0, ctx->stream>>>(vol, b, k, pb.par, pb.size);
    BMX_LAUNCH_CHECK();
    k_cc_merge<<<grid, 256, 0, ctx->stream>>>(b, pb.par);
    BMX_LAUNCH_CHECK();
    k_cc_compress<<<grid, 256, 0, ctx->stream>>>(b, pb.par);
    BMX_LAUNCH_CHECK();
    k_cc_flatten<<<grid, 256, 0, ctx->stream>>>(b, pb.par, pb.size);
    BMX_LAUNCH_CHECK();
    k_cc_relabel<<<grid, 256, 0, ctx->stream>>>(b, pb.par);
    BMX_LAUNCH_CHECK();
    k_cc_max<<<grid, 256, 0, ctx->stream>>>(b, pb.size, pb.stats);
    BMX_LAUNCH_CHECK();
    k_cc_argmax<<<grid, 256, 0, ctx->stream>>>(b, pb.size, pb.stats);
    BMX_LAUNCH_CHECK();
    k_cc_apply<FILL><<<grid, 256, 0, ctx->stream>>>(vol, b, k, t, pb.par, pb.size, pb.stats);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

template <bool FILL>
int post_volume(bmx_ctx *ctx, const uint8_t *in, int64_t Z, int64_t Y, int64_t X, double threshold, uint8_t *out)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(in && out, "bmx_%s: null pointer", FILL ? "fill" : "clean");
    BMX_CHECK(Z >= 1 && Y >= 1 && X >= 1 && Z * Y * X < ((int64_t)1 << 32) - 1, "bmx_%s: bad dims",
              FILL ? "fill" : "clean");
    BMX_CHECK(threshold >= 0.0, "bmx_%s: negative threshold", FILL ? "fill" : "clean");
    const int64_t V = Z * Y * X;
    if (out != in) BMX_HIP(hipMemcpyAsync(out, in, (size_t)V, hipMemcpyDeviceToDevice, ctx->stream));
    PostBufs pb;
    hipError_t e = hipMalloc((void **)&pb.par, (size_t)V * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&pb.size, (size_t)V * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&pb.stats, 4 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&pb.hist, 256 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&pb.box, 8 * sizeof(int));
    if (e == hipSuccess) e = hipMemsetAsync(pb.hist, 0, 256 * sizeof(uint32_t), ctx->stream);
    std::vector<uint32_t> hist(256, 0u);
    if (e == hipSuccess) {
        k_hist256<<<bmx_stride_blocks(V, 256) < 1024 ? bmx_stride_blocks(V, 256) : 1024, 256, 0, ctx->stream>>>(
            out, V, pb.hist);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(hist.data(), pb.hist, 256 * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
        rc = bmx_set_error(e == hipErrorOutOfMemory ? BMX_E_OOM : BMX_E_DEVICE, "bmx_%s: %s",
                           FILL ? "fill" : "clean", hipGetErrorString(e));
    // the labels present, skipping the smallest (remove_outlier.py:62, :99)
    bool first = true;
    for (int k = 0; k < 256 && rc == BMX_OK; ++k) {
        if (!hist[k]) continue;
        if (first) {
            first = false;
            continue;
        }
        rc = post_label<FILL>(ctx, out, (int)Z, (int)Y, (int)X, (uint8_t)k, threshold, pb);
    }
    const hipError_t s = hipStreamSynchronize(ctx->stream);
    if (s != hipSuccess && rc == BMX_OK) rc = bmx_set_error(BMX_E_DEVICE, "bmx_post: %s", hipGetErrorString(s));
    (void)hipFree(pb.par);
    (void)hipFree(pb.size);
    (void)hipFree(pb.stats);
    (void)hipFree(pb.hist);
    (void)hipFree(pb.box);
    return rc;
}

}  // namespace

extern "C" int bmx_clean(bmx_ctx *ctx, const uint8_t *in, int64_t Z, int64_t Y, int64_t X, double threshold,
                         uint8_t *out)
{
    return post_volume<false>(ctx, in, Z, Y, X, threshold, out);
}

extern "C" int bmx_fill(bmx_ctx *ctx, const uint8_t *in, int64_t Z, int64_t Y, int64_t X, double threshold,
                        uint8_t *out)
{
    return post_volume<true>(ctx, in, Z, Y, X, threshold, out);
}

extern "C" int bmx_histogram(bmx_ctx *ctx, const uint8_t *vol, int64_t n, uint64_t *counts)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(vol && counts, "bmx_histogram: null pointer");
    BMX_CHECK(n >= 0, "bmx_histogram: negative size");
    uint32_t *d = nullptr;
    BMX_HIP(hipMalloc((void **)&d, 256 * sizeof(uint32_t)));
    hipError_t e = hipMemsetAsync(d, 0, 256 * sizeof(uint32_t), ctx->stream);
    if (e == hipSuccess && n > 0) {
        const unsigned grid = bmx_stride_blocks(n, 256);
        k_hist256<<<grid < 1024 ? grid : 1024, 256, 0, ctx->stream>>>(vol, n, d);
        e = hipGetLastError();
    }
    uint32_t h[256];
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return bmx_set_error(BMX_E_DEVICE, "bmx_histogram: %s", hipGetErrorString(e));
    for (int k = 0; k < 256; ++k) counts[k] = h[k];
    return BMX_OK;
}

namespace {

// smooth_img_3x3 (features/biomedisa_helper.py:309-323): every voxel becomes the mean of
// its 3x3x3 neighbourhood clipped to the volume, summed in (dz, dy, dx) order
// in double and stored with the image's type (uint8: truncated).
template <typename T>
__global__ __launch_bounds__(256) void k_denoise(const T *__restrict__ in, int Z, int Y, int X, T *__restrict__ out)
{
    const int64_t sy = X, sz = (int64_t)X * Y, V = sz * Z;
    for (int64_t i = bmx_gtid(); i < V; i += bmx_gstride()) {
        const int z = (int)(i / sz);
        const int64_t r = i - (int64_t)z * sz;
        const int y = (int)(r / sy), x = (int)(r - (int64_t)y * sy);
        double tmp = 0.0;
        int n = 0;
        for (int k = -1; k < 2; ++k)
            for (int l = -1; l < 2; ++l)
                for (int m = -1; m < 2; ++m)
                    if (z + k >= 0 && z + k < Z && y + l >= 0 && y + l < Y && x + m >= 0 && x + m < X) {
                        tmp += (double)in[i + k * sz + l * sy + m];
                        ++n;
                    }
        out[i] = (T)(tmp / (double)n);
    }
}

}  // namespace

extern "C" int bmx_denoise(bmx_ctx *ctx, const void *img, int img_dtype, int64_t Z, int64_t Y, int64_t X, void *out)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(img && out && img != out, "bmx_denoise: null or aliased pointers");
    BMX_CHECK(img_dtype == BMX_IMG_U8 || img_dtype == BMX_IMG_F32, "bmx_denoise: bad image dtype");
    BMX_CHECK(Z >= 1 && Y >= 1 && X >= 1 && Z * Y * X < ((int64_t)1 << 32), "bmx_denoise: bad dims");
    const unsigned grid = bmx_stride_blocks(Z * Y * X, 256);
    if (img_dtype == BMX_IMG_U8)
        k_denoise<uint8_t><<<grid, 256, 0, ctx->stream>>>((const uint8_t *)img, (int)Z, (int)Y, (int)X,
                                                          (uint8_t *)out);
    else
        k_denoise<float><<<grid, 256, 0, ctx->stream>>>((const float *)img, (int)Z, (int)Y, (int)X, (float *)out);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}
