// bmx_seeds.hip -- labelled-plane detection and seed extraction.
//
// Reference semantics:
//   plane labelled (XY):      biomedisa_helper.py:1015-1022 (np.any(slice))
//   plane labelled (allaxis): biomedisa_helper.py:1024-1066 (1-voxel-thin
//                             lines removed along the first, then the second
//                             in-plane axis; labelled iff anything survives)
//   seed mask:                pyopencl_small.py:89-107 (labelled planes keep
//                             their label, everything else is ignore_value)
//   seed test:                pyopencl_small.py:205-207 (interior voxel whose
//                             mask value is a label)
//   variance:                 _calc_var, pyopencl_small.py:140-185
// Large-volume variants (bmx_plane_flags_ex large=1, bmx_adaptive_mask):
//   planes (XY):              biomedisa_helper.py:1068-1080 (a labelled plane
//                             hides the next 4 planes)
//   planes (allaxis):         biomedisa_helper.py:1082-1103 (as above, but only
//                             positive labels count)
//   adaptive walks:           pyopencl_large.py:570-590 (a seed walks if it lies
//                             on the sub-lattice (x+z, y+z even) or another label
//                             is seen on the stride-5 grid within +-100 voxels)
#include "bmx_internal.h"

namespace {

constexpr int kSeedThreads = 256;
constexpr int kSeedPerThread = 16;
constexpr int kSeedChunk = kSeedThreads * kSeedPerThread;

// Label volumes: label index per voxel, 0 = background; the index `none`
// (BMX_LABEL_NONE, or 256 when all 256 indices are labels) marks ignored
// voxels (negative labels).  Inside the kernels kNone16 stands for "no label"
// (ignored, or off the labelled planes), which no index equals.
constexpr int kNone16 = 0x100;

// Voxel counts for plane detection: any non-zero label index (ignored voxels
// included), or with POS only positive labels.
template <bool POS>
__device__ __forceinline__ bool nz(const uint8_t *lab, int64_t i, int none)
{
    const int v = lab[i];
    return POS ? (v != 0 && v != none) : v != 0;
}

// Voxel survives the first thin-line pass along an axis of extent na, stride sa.
template <bool POS>
__device__ __forceinline__ bool pass1(const uint8_t *lab, int64_t i, int a, int na, int64_t sa, int none)
{
    if (!nz<POS>(lab, i, none)) return false;
    if (a > 0 && a < na - 1 && !nz<POS>(lab, i - sa, none) && !nz<POS>(lab, i + sa, none)) return false;
    return true;
}

template <bool POS>
__device__ __forceinline__ bool pass2(const uint8_t *lab, int64_t i, int a, int na, int64_t sa,
                                      int b, int nb, int64_t sb, int none)
{
    if (!pass1<POS>(lab, i, a, na, sa, none)) return false;
    if (b > 0 && b < nb - 1 && !pass1<POS>(lab, i - sb, a, na, sa, none) &&
        !pass1<POS>(lab, i + sb, a, na, sa, none))
        return false;
    return true;
}

template <bool POS>
__global__ __launch_bounds__(256) void k_plane_flags(const uint8_t *__restrict__ lab, int none, int Z, int Y,
                                                     int X, int allaxis, uint8_t *fz, uint8_t *fy,
                                                     uint8_t *fx)
{
    const int64_t sy = X, sz = (int64_t)X * Y, V = sz * Z;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += stride) {
        if (!nz<false>(lab, i, none)) continue;
        const int z = (int)(i / sz);
        const int64_t r = i - z * sz;
        const int y = (int)(r / sy);
        const int x = (int)(r - y * sy);
        if (!allaxis) {
            if (!fz[z]) fz[z] = 1;
            continue;
        }
        if (!fz[z] && pass2<POS>(lab, i, y, Y, sy, x, X, 1, none)) fz[z] = 1;
        if (!fy[y] && pass2<POS>(lab, i, z, Z, sz, x, X, 1, none)) fy[y] = 1;
        if (!fx[x] && pass2<POS>(lab, i, z, Z, sz, y, Y, sy, none)) fx[x] = 1;
    }
}

// XY planes of the large-volume path: scanning z upwards, a labelled plane is
// kept and the 4 planes after it are skipped (biomedisa_helper.py:1068-1080).
__global__ void k_skip_planes(uint8_t *fz, int Z)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int z = 0;
    while (z < Z) {
        if (fz[z]) {
            for (int k = z + 1; k < z + 5 && k < Z; ++k) fz[k] = 0;
            z += 5;
        } else {
            ++z;
        }
    }
}

struct Geom {
    int Z, Y, X;
    int64_t sy, sz, V;
    int none;   // the label volume's index for ignored voxels (256: none)
};

// The label the reference's rebuilt slice volume holds at a voxel
// (pyopencl_small.py:89-107): its label on a labelled plane, else kNone16.
__device__ __forceinline__ int mask_at(const uint8_t *lab, const uint8_t *fz, const uint8_t *fy,
                                       const uint8_t *fx, const Geom &g, int z, int y, int x, int64_t i)
{
    if (!(fz[z] | fy[y] | fx[x])) return kNone16;
    const int v = lab[i];
    return v == g.none ? kNone16 : v;
}

__device__ __forceinline__ bool is_seed(const uint8_t *lab, const uint8_t *fz, const uint8_t *fy,
                                        const uint8_t *fx, const Geom &g, int z, int y, int x,
                                        int64_t i)
{
    if (z <= 0 || z >= g.Z - 1 || y <= 0 || y >= g.Y - 1 || x <= 0 || x >= g.X - 1) return false;
    return mask_at(lab, fz, fy, fx, g, z, y, x, i) != kNone16;
}

__device__ __forceinline__ void coords(const Geom &g, int64_t i, int &z, int &y, int &x)
{
    z = (int)(i / g.sz);
    const int64_t r = i - (int64_t)z * g.sz;
    y = (int)(r / g.sy);
    x = (int)(r - (int64_t)y * g.sy);
}

__device__ __forceinline__ void advance(const Geom &g, int &z, int &y, int &x)
{
    if (++x == g.X) {
        x = 0;
        if (++y == g.Y) { y = 0; ++z; }
    }
}

// Seeds among this thread's kSeedPerThread consecutive voxels.
__device__ int thread_seed_count(const uint8_t *lab, const uint8_t *fz, const uint8_t *fy,
                                 const uint8_t *fx, const Geom &g, int64_t first)
{
    int cnt = 0;
    if (first >= g.V) return 0;
    int z, y, x;
    coords(g, first, z, y, x);
    for (int k = 0; k < kSeedPerThread; ++k) {
        const int64_t i = first + k;
        if (i >= g.V) break;
        cnt += is_seed(lab, fz, fy, fx, g, z, y, x, i) ? 1 : 0;
        advance(g, z, y, x);
    }
    return cnt;
}

// Exclusive block scan of one int per thread (256 threads = 4 waves of 64).
__device__ int block_exclusive_scan(int v, int *total)
{
    __shared__ int wave_sums[kSeedThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) wave_sums[wave] = incl;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kSeedThreads / 64; ++w) {
        if (w < wave) base += wave_sums[w];
        tot += wave_sums[w];
    }
    __syncthreads();
    *total = tot;
    return base + incl - v;
}

__global__ __launch_bounds__(kSeedThreads) void k_seed_count(const uint8_t *__restrict__ lab,
                                                             const uint8_t *__restrict__ flags,
                                                             Geom g, int64_t *block_counts)
{
    const uint8_t *fz = flags, *fy = flags + g.Z, *fx = flags + g.Z + g.Y;
    const int64_t first = (int64_t)blockIdx.x * kSeedChunk + (int64_t)threadIdx.x * kSeedPerThread;
    const int c = thread_seed_count(lab, fz, fy, fx, g, first);
    int tot;
    (void)block_exclusive_scan(c, &tot);
    if (threadIdx.x == 0) block_counts[blockIdx.x] = tot;
}

// Exclusive scan of n int64 counts in place by one block; total -> out_total.
__global__ __launch_bounds__(1024) void k_scan_counts(int64_t *counts, int64_t n, uint64_t *out_total)
{
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (n + 1023) / 1024;
    const int64_t lo = t * per, hi = (lo + per < n) ? lo + per : n;
    int64_t s = 0;
    for (int64_t i = lo; i < hi; ++i) s += counts[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t v = (t >= off) ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t run = part[t] - s;
    for (int64_t i = lo; i < hi; ++i) {
        const int64_t c = counts[i];
        counts[i] = run;
        run += c;
    }
    if (t == 1023) *out_total = (uint64_t)part[1023];
}

// _calc_var (pyopencl_small.py:140-185) with dev += tmp * tmp contracted to
// fma(tmp, tmp, dev), as the reference's compilers build it (contraction on
// by default); exact either way for uint8 images.
template <typename T>
__device__ float calc_var(const uint8_t *lab, const uint8_t *fz, const uint8_t *fy, const uint8_t *fx,
                          const T *img, const Geom &g, int z, int y, int x, int64_t i0)
{
    const int seg = mask_at(lab, fz, fy, fx, g, z, y, x, i0);
    const float B = bmx_ld(img, i0);
    float dev = 0.0f, summe = 0.0f;
    for (int n = -1; n < 2; ++n)
        for (int o = -1; o < 2; ++o) {
            const int64_t i = i0 + n * g.sy + o;
            if (mask_at(lab, fz, fy, fx, g, z, y + n, x + o, i) == seg) {
                const float t = B - bmx_ld(img, i);
                dev = __builtin_fmaf(t, t, dev);
                summe += 1.0f;
            }
        }
    for (int n = -1; n <= 1; n += 2)
        for (int o = -1; o < 2; ++o) {
            const int64_t i = i0 + n * g.sz + o;
            if (summe < 9.0f && mask_at(lab, fz, fy, fx, g, z + n, y, x + o, i) == seg) {
                const float t = B - bmx_ld(img, i);
                dev = __builtin_fmaf(t, t, dev);
                summe += 1.0f;
            }
        }
    for (int n = -1; n <= 1; n += 2)
        for (int o = -1; o <= 1; o += 2) {
            const int64_t i = i0 + n * g.sz + o * g.sy;
            if (summe < 9.0f && mask_at(lab, fz, fy, fx, g, z + n, y + o, x, i) == seg) {
                const float t = B - bmx_ld(img, i);
                dev = __builtin_fmaf(t, t, dev);
                summe += 1.0f;
            }
        }
    float var = dev / summe;
    if (var < 1.0f) var = 1.0f;
    return var;
}

template <typename T>
__global__ __launch_bounds__(kSeedThreads) void k_seed_write(const uint8_t *__restrict__ lab,
                                                             const T *__restrict__ img,
                                                             const uint8_t *__restrict__ flags,
                                                             Geom g, const int64_t *block_offsets,
                                                             bmx_seed *seeds, int64_t capacity)
{
    const uint8_t *fz = flags, *fy = flags + g.Z, *fx = flags + g.Z + g.Y;
    const int64_t first = (int64_t)blockIdx.x * kSeedChunk + (int64_t)threadIdx.x * kSeedPerThread;
    const int c = thread_seed_count(lab, fz, fy, fx, g, first);
    int tot;
    const int off = block_exclusive_scan(c, &tot);
    if (c == 0) return;
    int64_t o = block_offsets[blockIdx.x] + off;
    int z, y, x;
    coords(g, first, z, y, x);
    for (int k = 0; k < kSeedPerThread; ++k) {
        const int64_t i = first + k;
        if (i >= g.V) break;
        if (is_seed(lab, fz, fy, fx, g, z, y, x, i)) {
            if (o < capacity) {
                bmx_seed s;
                s.pos = (uint32_t)i;
                s.label = lab[i];
                s.var = calc_var(lab, fz, fy, fx, img, g, z, y, x, i);
                s.reserved = 0;
                seeds[o] = s;
            }
            ++o;
        }
        advance(g, z, y, x);
    }
}

}  // namespace

extern "C" {

int bmx_plane_flags_lab(bmx_ctx *ctx, const uint8_t *lab, int label_none, int64_t Z, int64_t Y, int64_t X,
                        int allaxis, int large, uint8_t *flags)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(lab && flags, "bmx_plane_flags: null pointer");
    BMX_CHECK(label_none >= 1 && label_none <= 256, "bmx_plane_flags: label_none %d not in [1,256]", label_none);
    BMX_CHECK(Z >= 1 && Y >= 1 && X >= 1 && Z * Y * X < (int64_t)1 << 32,
              "bmx_plane_flags: bad dims %lld x %lld x %lld", (long long)Z, (long long)Y,
              (long long)X);
    BMX_HIP(hipMemsetAsync(flags, 0, (size_t)(Z + Y + X), ctx->stream));
    const int64_t V = Z * Y * X;
    int64_t blocks = (V + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (large)
        k_plane_flags<true><<<(unsigned)blocks, 256, 0, ctx->stream>>>(lab, label_none, (int)Z, (int)Y, (int)X,
                                                                      allaxis, flags, flags + Z, flags + Z + Y);
    else
        k_plane_flags<false><<<(unsigned)blocks, 256, 0, ctx->stream>>>(lab, label_none, (int)Z, (int)Y, (int)X,
                                                                       allaxis, flags, flags + Z, flags + Z + Y);
    BMX_LAUNCH_CHECK();
    if (large && !allaxis) {
        k_skip_planes<<<1, 64, 0, ctx->stream>>>(flags, (int)Z);
        BMX_LAUNCH_CHECK();
    }
    return BMX_OK;
}

int bmx_plane_flags_ex(bmx_ctx *ctx, const uint8_t *lab, int64_t Z, int64_t Y, int64_t X, int allaxis,
                       int large, uint8_t *flags)
{
    return bmx_plane_flags_lab(ctx, lab, BMX_LABEL_NONE, Z, Y, X, allaxis, large, flags);
}

int bmx_plane_flags(bmx_ctx *ctx, const uint8_t *lab, int64_t Z, int64_t Y, int64_t X, int allaxis,
                    uint8_t *flags)
{
    return bmx_plane_flags_lab(ctx, lab, BMX_LABEL_NONE, Z, Y, X, allaxis, 0, flags);
}

int bmx_seeds_lab(bmx_ctx *ctx, const uint8_t *lab, int label_none, const void *img, int img_dtype, int64_t Z,
                  int64_t Y, int64_t X, const uint8_t *flags, bmx_seed *seeds, int64_t capacity, int64_t *nseeds)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(lab && flags && nseeds, "bmx_seeds: null pointer");
    BMX_CHECK(label_none >= 1 && label_none <= 256, "bmx_seeds: label_none %d not in [1,256]", label_none);
    BMX_CHECK(img_dtype == BMX_IMG_U8 || img_dtype == BMX_IMG_F32, "bmx_seeds: bad image dtype %d",
              img_dtype);
    BMX_CHECK(seeds == nullptr || img != nullptr, "bmx_seeds: image needed to write seeds");
    BMX_CHECK(Z >= 1 && Y >= 1 && X >= 1 && Z * Y * X < (int64_t)1 << 32,
              "bmx_seeds: bad dims %lld x %lld x %lld", (long long)Z, (long long)Y, (long long)X);
    Geom g{(int)Z, (int)Y, (int)X, X, X * Y, X * Y * Z, label_none};
    const int64_t nblk = (g.V + kSeedChunk - 1) / kSeedChunk;
    if (ctx->block_counts_cap < nblk) {
        if (ctx->d_block_counts) BMX_HIP(hipFree(ctx->d_block_counts));
        ctx->d_block_counts = nullptr;
        ctx->block_counts_cap = 0;
        BMX_HIP(hipMalloc((void **)&ctx->d_block_counts, (size_t)nblk * sizeof(int64_t)));
        ctx->block_counts_cap = nblk;
    }
    k_seed_count<<<(unsigned)nblk, kSeedThreads, 0, ctx->stream>>>(lab, flags, g, ctx->d_block_counts);
    BMX_LAUNCH_CHECK();
    k_scan_counts<<<1, 1024, 0, ctx->stream>>>(ctx->d_block_counts, nblk, ctx->d_counter);
    BMX_LAUNCH_CHECK();
    if (seeds) {
        if (img_dtype == BMX_IMG_U8)
            k_seed_write<uint8_t><<<(unsigned)nblk, kSeedThreads, 0, ctx->stream>>>(
                lab, (const uint8_t *)img, flags, g, ctx->d_block_counts, seeds, capacity);
        else
            k_seed_write<float><<<(unsigned)nblk, kSeedThreads, 0, ctx->stream>>>(
                lab, (const float *)img, flags, g, ctx->d_block_counts, seeds, capacity);
        BMX_LAUNCH_CHECK();
    }
    uint64_t total = 0;
    BMX_HIP(hipMemcpyAsync(&total, ctx->d_counter, sizeof(total), hipMemcpyDeviceToHost, ctx->stream));
    BMX_HIP(hipStreamSynchronize(ctx->stream));
    *nseeds = (int64_t)total;
    if (seeds && (int64_t)total > capacity)
        return bmx_set_error(BMX_E_INVALID, "bmx_seeds: %lld seeds exceed capacity %lld",
                             (long long)total, (long long)capacity);
    return BMX_OK;
}

int bmx_seeds(bmx_ctx *ctx, const uint8_t *lab, const void *img, int img_dtype, int64_t Z, int64_t Y,
              int64_t X, const uint8_t *flags, bmx_seed *seeds, int64_t capacity, int64_t *nseeds)
{
    return bmx_seeds_lab(ctx, lab, BMX_LABEL_NONE, img, img_dtype, Z, Y, X, flags, seeds, capacity, nseeds);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Seed selection by a voxel mask (surrogate-guided walking): keeps the seeds
// whose voxel has mask != 0.  Output order is unspecified (the walk does not
// depend on seed order).
// ---------------------------------------------------------------------------
namespace {

// A block takes kSelItems seeds per thread (loads strided by the block for
// coalescing), counts its kept seeds with a block-wide scan and reserves their
// output range with one atomic: one atomic per 4096 seeds, where one per wave
// made the shared counter the bottleneck.
constexpr int kSelItems = 16;

__global__ __launch_bounds__(256) void k_select_seeds(const bmx_seed *__restrict__ in, int64_t n,
                                                      const uint8_t *__restrict__ mask,
                                                      bmx_seed *__restrict__ out,
                                                      unsigned long long *count)
{
    __shared__ uint32_t wsum[4];
    __shared__ unsigned long long sbase;
    const int tid = (int)threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t b0 = (int64_t)blockIdx.x * 256 * kSelItems;
    bmx_seed s[kSelItems];
    uint32_t keep = 0;
#pragma unroll
    for (int j = 0; j < kSelItems; ++j) {
        const int64_t i = b0 + (int64_t)j * 256 + tid;
        if (i < n) {
            s[j] = in[i];
            if (mask[s[j].pos] != 0) keep |= 1u << j;
        }
    }
    // exclusive prefix of the kept counts over the block
    const uint32_t c = (uint32_t)__popc(keep);
    uint32_t incl = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(incl, (unsigned)o, 64);
        if (lane >= o) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        before += w < wave ? wsum[w] : 0u;
        total += wsum[w];
    }
    if (tid == 0) sbase = total ? atomicAdd(count, (unsigned long long)total) : 0ull;
    __syncthreads();
    unsigned long long at = sbase + before + incl - c;
#pragma unroll
    for (int j = 0; j < kSelItems; ++j)
        if ((keep >> j) & 1u) out[at++] = s[j];
}

// Adaptive walks of the large-volume path (pyopencl_large.py:570-590): the
// seed at p (label L) walks if x+z and y+z are even, or if some interior voxel
// q = p + 5*(i,j,k), |i|,|j|,|k| <= 20, carries a label other than L on a
// labelled plane.  p itself is in that window with label L, so "another label
// is seen" is "the window's smallest and largest label differ".  The window is
// a product of one stride-5 set per axis (the interior test is per axis too),
// so min and max are taken in three separable passes: off-plane voxels are
// skipped and ignored ones (index `none`) count as no label: 0xFF is the
// identity of min (the seed's own label is always in its window) and such a
// voxel counts as 0 for max.
constexpr int kAdaptStep = 5, kAdaptTaps = 20;   // offsets -100..100 in steps of 5

struct MinMax {
    uint8_t lo, hi;
};

// Pass along x: the label seen at q (labelled plane, else none), min/max over
// the stride-5 window of x.
__global__ __launch_bounds__(256) void k_adapt_x(const uint8_t *__restrict__ lab, int none,
                                                 const uint8_t *__restrict__ fz, const uint8_t *__restrict__ fy,
                                                 const uint8_t *__restrict__ fx, int Z, int Y, int X,
                                                 MinMax *__restrict__ out)
{
    const int64_t sz = (int64_t)X * Y, V = sz * Z;
    for (int64_t i = bmx_gtid(); i < V; i += bmx_gstride()) {
        const int z = (int)(i / sz);
        const int64_t r = i - z * sz;
        const int y = (int)(r / X);
        const int x = (int)(r - (int64_t)y * X);
        uint8_t lo = 0xFF, hi = 0;
        if (fz[z] | fy[y]) {
            const int64_t row = i - x;
            for (int k = -kAdaptTaps; k <= kAdaptTaps; ++k) {
                const int qx = x + kAdaptStep * k;
                if (qx <= 0 || qx >= X - 1) continue;
                const uint8_t v = lab[row + qx];
                lo = (int)v == none ? lo : min(lo, v);
                hi = max(hi, (int)v == none ? (uint8_t)0 : v);
            }
        } else {
            // only x-planes can be labelled on this row
            for (int k = -kAdaptTaps; k <= kAdaptTaps; ++k) {
                const int qx = x + kAdaptStep * k;
                if (qx <= 0 || qx >= X - 1 || !fx[qx]) continue;
                const uint8_t v = lab[i - x + qx];
                lo = (int)v == none ? lo : min(lo, v);
                hi = max(hi, (int)v == none ? (uint8_t)0 : v);
            }
        }
        out[i] = MinMax{lo, hi};
    }
}

// Pass along y (stride X) or z (stride X*Y) of the (min, max) pairs.  The z
// pass writes the mask: 1 where the seed walks.
template <bool FINAL>
__global__ __launch_bounds__(256) void k_adapt_axis(const MinMax *__restrict__ in, int Z, int Y, int X, int axis,
                                                    MinMax *__restrict__ out, uint8_t *__restrict__ mask)
{
    const int64_t sz = (int64_t)X * Y, V = sz * Z;
    const int64_t st = axis == 1 ? (int64_t)X : sz;
    const int n = axis == 1 ? Y : Z;
    for (int64_t i = bmx_gtid(); i < V; i += bmx_gstride()) {
        const int z = (int)(i / sz);
        const int64_t r = i - z * sz;
        const int y = (int)(r / X);
        const int x = (int)(r - (int64_t)y * X);
        const int c = axis == 1 ? y : z;
        uint8_t lo = 0xFF, hi = 0;
        for (int k = -kAdaptTaps; k <= kAdaptTaps; ++k) {
            const int q = c + kAdaptStep * k;
            if (q <= 0 || q >= n - 1) continue;
            const MinMax m = in[i + (int64_t)(kAdaptStep * k) * st];
            lo = min(lo, m.lo);
            hi = max(hi, m.hi);
        }
        if (FINAL) {
            const bool lattice = ((x + z) % 2 == 0) && ((y + z) % 2 == 0);
            mask[i] = (lattice || lo != hi) ? 1 : 0;
        } else {
            out[i] = MinMax{lo, hi};
        }
    }
}

}  // namespace

extern "C" int bmx_adaptive_mask_lab(bmx_ctx *ctx, const uint8_t *lab, int label_none, int64_t Z, int64_t Y,
                                     int64_t X, const uint8_t *flags, void *tmp, uint8_t *mask)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(lab && flags && tmp && mask, "bmx_adaptive_mask: null pointer");
    BMX_CHECK(label_none >= 1 && label_none <= 256, "bmx_adaptive_mask: label_none %d not in [1,256]", label_none);
    BMX_CHECK(Z >= 1 && Y >= 1 && X >= 1 && Z * Y * X < (int64_t)1 << 32, "bmx_adaptive_mask: bad dims");
    const int64_t V = Z * Y * X;
    MinMax *a = (MinMax *)tmp, *b = a + V;
    const unsigned grid = bmx_stride_blocks(V, 256);
    k_adapt_x<<<grid, 256, 0, ctx->stream>>>(lab, label_none, flags, flags + Z, flags + Z + Y, (int)Z, (int)Y, (int)X,
                                             a);
    BMX_LAUNCH_CHECK();
    k_adapt_axis<false><<<grid, 256, 0, ctx->stream>>>(a, (int)Z, (int)Y, (int)X, 1, b, nullptr);
    BMX_LAUNCH_CHECK();
    k_adapt_axis<true><<<grid, 256, 0, ctx->stream>>>(b, (int)Z, (int)Y, (int)X, 2, nullptr, mask);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

extern "C" int bmx_adaptive_mask(bmx_ctx *ctx, const uint8_t *lab, int64_t Z, int64_t Y, int64_t X,
                                 const uint8_t *flags, void *tmp, uint8_t *mask)
{
    return bmx_adaptive_mask_lab(ctx, lab, BMX_LABEL_NONE, Z, Y, X, flags, tmp, mask);
}

extern "C" int bmx_select_seeds(bmx_ctx *ctx, const bmx_seed *seeds, int64_t nseeds, const uint8_t *mask,
                                bmx_seed *out, int64_t *nout)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(nout != nullptr, "bmx_select_seeds: nout is null");
    BMX_CHECK(nseeds == 0 || (seeds && mask && out), "bmx_select_seeds: null pointer");
    BMX_CHECK(nseeds >= 0 && nseeds < ((int64_t)1 << 32) - 256, "bmx_select_seeds: nseeds out of range");
    BMX_HIP(hipMemsetAsync(ctx->d_counter + 2, 0, sizeof(uint64_t), ctx->stream));
    if (nseeds > 0) {
        k_select_seeds<<<bmx_blocks(nseeds, 256 * kSelItems), 256, 0, ctx->stream>>>(
            seeds, nseeds, mask, out, (unsigned long long *)(ctx->d_counter + 2));
        BMX_LAUNCH_CHECK();
    }
    uint64_t c = 0;
    BMX_HIP(hipMemcpyAsync(&c, ctx->d_counter + 2, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
    BMX_HIP(hipStreamSynchronize(ctx->stream));
    *nout = (int64_t)c;
    return BMX_OK;
}
