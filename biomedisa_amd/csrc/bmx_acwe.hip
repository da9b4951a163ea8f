// bmx_acwe.hip -- active-contour post-processing of a label volume
// (biomedisa/features/active_contour.py:158-169 with the evolution and curvop
// operators of curvop_numba.py:32-148).
//
// `steps` times:
//   1. mean intensity of every label (float32 of the float64 mean, as
//      np.mean of uint8 values stored into a float32 array);
//   2. evolution: every interior voxel takes, among its 27 neighbours' labels
//      (dz, dy, dx outer to inner, -1..1), the first with the strictly smallest
//      cost |mean[l] - img|, alpha x that for label 0, starting from its own;
//   3. for every label in ascending order, `smooth` rounds of a morphological
//      pair on the volume as changed by the labels before it: even rounds an
//      erosion then a dilation, odd rounds the reverse.
//        erosion: an interior voxel of the label with a 26-neighbour of another
//                 label, and no one of 9 planar 3x3 patterns through it wholly
//                 of the label, takes the last other label among its
//                 26 neighbours (dz, dy, dx order);
//        dilation: an interior voxel not of the label with a 26-neighbour of
//                 the label takes the label (the reference's pattern test there
//                 counts at least 18 of 27 positions and never passes).
// Every operator reads one volume and writes another (the reference's
// start/final copies), so each is one data-parallel pass; costs are compared in
// double, where |float32 mean - integer| is exact.
#include "bmx_internal.h"

#include <vector>

namespace {

struct LabelMeans {
    float m[256];
};

struct Geo {
    int Z, Y, X;
    int64_t sy, sz, V;
};

__device__ __forceinline__ bool interior(const Geo &g, int z, int y, int x)
{
    return z > 0 && z < g.Z - 1 && y > 0 && y < g.Y - 1 && x > 0 && x < g.X - 1;
}

__device__ __forceinline__ void coords(const Geo &g, int64_t i, int &z, int &y, int &x)
{
    z = (int)(i / g.sz);
    const int64_t r = i - (int64_t)z * g.sz;
    y = (int)(r / g.sy);
    x = (int)(r - (int64_t)y * g.sy);
}

// Per-label sums and counts of the image (uint64; the float64 mean of uint8
// values is their exact sum divided by the count).
__global__ __launch_bounds__(256) void k_label_sums(const uint8_t *__restrict__ img, const uint8_t *__restrict__ lab,
                                                    int64_t V, unsigned long long *__restrict__ sums,
                                                    unsigned long long *__restrict__ counts)
{
    __shared__ unsigned long long s[256], c[256];
    s[threadIdx.x] = 0ull;
    c[threadIdx.x] = 0ull;
    __syncthreads();
    for (int64_t i = bmx_gtid(); i < V; i += bmx_gstride()) {
        const uint8_t l = lab[i];
        atomicAdd(&s[l], (unsigned long long)img[i]);
        atomicAdd(&c[l], 1ull);
    }
    __syncthreads();
    if (c[threadIdx.x]) {
        atomicAdd(sums + threadIdx.x, s[threadIdx.x]);
        atomicAdd(counts + threadIdx.x, c[threadIdx.x]);
    }
}

__global__ __launch_bounds__(256) void k_evolution(const uint8_t *__restrict__ A, const uint8_t *__restrict__ img,
                                                   Geo g, LabelMeans mean, double alpha, uint8_t *__restrict__ B)
{
    for (int64_t i = bmx_gtid(); i < g.V; i += bmx_gstride()) {
        int z, y, x;
        coords(g, i, z, y, x);
        const uint8_t ref_label = A[i];
        if (!interior(g, z, y, x)) {
            B[i] = ref_label;
            continue;
        }
        const double v = (double)img[i];
        double ref = fabs((double)mean.m[ref_label] - v);
        if (ref_label == 0) ref = alpha * ref;
        uint8_t fin = ref_label;
        for (int n = -1; n <= 1; ++n)
            for (int o = -1; o <= 1; ++o)
                for (int p = -1; p <= 1; ++p) {
                    const uint8_t t = A[i + n * g.sz + o * g.sy + p];
                    if (t != ref_label) {
                        double val = fabs((double)mean.m[t] - v);
                        if (t == 0) val = alpha * val;
                        if (val < ref) {
                            ref = val;
                            fin = t;
                        }
                    }
                }
        B[i] = fin;
    }
}

// The 9 planar patterns through the centre of a 3x3x3 block, as bit masks over
// the block index (dz+1)*9 + (dy+1)*3 + (dx+1): the planes dx = 0, dy = 0,
// dz = 0, and the six diagonal planes dy = dx, dy = -dx, dz = dx, dz = -dx,
// dz = dy, dz = -dy (curvop_numba.py:130-138).
__host__ __device__ constexpr uint32_t pattern_bit(int k, int l, int m) { return 1u << (k * 9 + l * 3 + m); }

struct Patterns {
    uint32_t p[9];
};

Patterns make_patterns()
{
    Patterns P{};
    for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l)
            for (int m = 0; m < 3; ++m) {
                const uint32_t b = pattern_bit(k, l, m);
                if (m == 1) P.p[0] |= b;
                if (l == 1) P.p[1] |= b;
                if (k == 1) P.p[2] |= b;
                if (l == m) P.p[3] |= b;
                if (l == 2 - m) P.p[4] |= b;
                if (k == m) P.p[5] |= b;
                if (k == 2 - m) P.p[6] |= b;
                if (k == l) P.p[7] |= b;
                if (k == 2 - l) P.p[8] |= b;
            }
    return P;
}

template <bool EROSION>
__global__ __launch_bounds__(256) void k_morph(const uint8_t *__restrict__ in, Geo g, uint8_t label, Patterns P,
                                               uint8_t *__restrict__ out)
{
    for (int64_t i = bmx_gtid(); i < g.V; i += bmx_gstride()) {
        int z, y, x;
        coords(g, i, z, y, x);
        const uint8_t c = in[i];
        uint8_t r = c;
        if (interior(g, z, y, x) && (EROSION ? c == label : c != label)) {
            uint32_t same = 0;        // bit set where the block voxel equals the label
            uint8_t last_other = label;
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l)
                    for (int m = 0; m < 3; ++m) {
                        const uint8_t t = in[i + (k - 1) * g.sz + (l - 1) * g.sy + (m - 1)];
                        if (t == label) same |= pattern_bit(k, l, m);
                        else last_other = t;
                    }
            constexpr uint32_t kAll = (1u << 27) - 1u;
            if (EROSION) {
                if (same != kAll) {   // some neighbour of another label
                    bool whole = false;
                    for (int t = 0; t < 9; ++t) whole |= (same & P.p[t]) == P.p[t];
                    if (!whole) r = last_other;
                }
            } else {
                if (same != 0u) r = label;
            }
        }
        out[i] = r;
    }
}

int launch_morph(bmx_ctx *ctx, bool erosion, const uint8_t *in, uint8_t *out, const Geo &g, uint8_t label,
                 const Patterns &P)
{
    const unsigned grid = bmx_stride_blocks(g.V, 256);
    if (erosion)
        k_morph<true><<<grid, 256, 0, ctx->stream>>>(in, g, label, P, out);
    else
        k_morph<false><<<grid, 256, 0, ctx->stream>>>(in, g, label, P, out);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

// --- refinement (active_contour.py:193-225) --------------------------------
//
// geodis (active_contour.py:42-89): one forward and one backward raster scan of
// a chamfer distance with weights 1, sqrt2, sqrt3 over float32 values, each
// voxel's new value min(neighbour + weight, own) computed in float64 and stored
// as float32.  A scan reads voxels it has already updated, so it is sequential;
// in the forward scan voxel (z, y, x) reads only voxels of smaller
// T = 4z + 2y + x, so all voxels of one T can be updated together and the
// wavefronts T = 0, 1, ... reproduce the scan bit for bit (the backward scan
// likewise with T taken from the far corner).
constexpr double kSqrt2 = 1.4142135623730951, kSqrt3 = 1.7320508075688772;   // np.sqrt(2), np.sqrt(3)

__global__ __launch_bounds__(256) void k_geo_init(const uint8_t *__restrict__ lab, int64_t V, float *__restrict__ c)
{
    for (int64_t i = bmx_gtid(); i < V; i += bmx_gstride()) c[i] = lab[i] == 0 ? 100.0f : 0.0f;
}

template <bool FORWARD>
__global__ __launch_bounds__(256) void k_geo_front(float *__restrict__ c, Geo g, int T)
{
    const int64_t n = (int64_t)g.Z * g.Y;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    int z = (int)(j / g.Y), y = (int)(j - (int64_t)z * g.Y);
    int x;
    if (FORWARD) {
        // z in [1, Z-1], y in [1, Y-2], x in [1, X-2]
        x = T - 4 * z - 2 * y;
        if (z < 1 || y < 1 || y > g.Y - 2 || x < 1 || x > g.X - 2) return;
    } else {
        // z in [0, Z-2], y in [1, Y-2], x in [1, X-2]; T counted from the far corner
        x = (g.X - 1) - (T - 4 * (g.Z - 1 - z) - 2 * (g.Y - 1 - y));
        if (z > g.Z - 2 || y < 1 || y > g.Y - 2 || x < 1 || x > g.X - 2) return;
    }
    const int64_t i = (int64_t)z * g.sz + (int64_t)y * g.sy + x;
    const int dz = FORWARD ? -1 : 1;
    const int64_t pz = i + dz * g.sz;
    double m = (double)c[i];
    // neighbours in the plane z -/+ 1
    for (int o = -1; o <= 1; ++o)
        for (int p = -1; p <= 1; ++p) {
            const int k = (o != 0) + (p != 0);
            const double w = k == 0 ? 1.0 : (k == 1 ? kSqrt2 : kSqrt3);
            m = fmin(m, (double)c[pz + o * g.sy + p] + w);
        }
    // neighbours in the same plane already scanned
    const int s = FORWARD ? -1 : 1;
    m = fmin(m, (double)c[i + s * g.sy - 1] + kSqrt2);
    m = fmin(m, (double)c[i + s * g.sy] + 1.0);
    m = fmin(m, (double)c[i + s * g.sy + 1] + kSqrt2);
    m = fmin(m, (double)c[i + s] + 1.0);
    c[i] = (float)m;
}

// d = 0 where the distance is at most 10, else 1 (active_contour.py:206-208).
__global__ __launch_bounds__(256) void k_geo_threshold(const float *__restrict__ c, int64_t V, uint8_t *__restrict__ d)
{
    for (int64_t i = bmx_gtid(); i < V; i += bmx_gstride()) d[i] = c[i] <= 10.0f ? 0 : 1;
}

// Per z-plane, over the voxels with d == 0: count, sum and sum of squares of
// the image per label value.  One block per plane.
__global__ __launch_bounds__(256) void k_refine_stats(const uint8_t *__restrict__ img, const uint8_t *__restrict__ lab,
                                                      const uint8_t *__restrict__ d, Geo g,
                                                      unsigned long long *__restrict__ stats)
{
    __shared__ unsigned long long h[256 * 3];
    for (int k = threadIdx.x; k < 256 * 3; k += blockDim.x) h[k] = 0ull;
    __syncthreads();
    const int z = blockIdx.x;
    const int64_t base = (int64_t)z * g.sz;
    for (int64_t r = threadIdx.x; r < g.sz; r += blockDim.x) {
        if (d[base + r]) continue;
        const uint8_t l = lab[base + r];
        const unsigned long long v = img[base + r];
        atomicAdd(&h[3 * l + 0], 1ull);
        atomicAdd(&h[3 * l + 1], v);
        atomicAdd(&h[3 * l + 2], v * v);
    }
    __syncthreads();
    unsigned long long *out = stats + (int64_t)z * 256 * 3;
    for (int k = threadIdx.x; k < 256 * 3; k += blockDim.x) out[k] = h[k];
}

// out = (d == 0 and lut[z][img]) ? label[z] : 0 (label[z] = 0: none).
__global__ __launch_bounds__(256) void k_refine_apply(const uint8_t *__restrict__ img, const uint8_t *__restrict__ d,
                                                      Geo g, const uint8_t *__restrict__ lut,
                                                      const uint8_t *__restrict__ label, uint8_t *__restrict__ out)
{
    for (int64_t i = bmx_gtid(); i < g.V; i += bmx_gstride()) {
        const int z = (int)(i / g.sz);
        const uint8_t l = label[z];
        out[i] = (l != 0 && d[i] == 0 && lut[(int64_t)z * 256 + img[i]]) ? l : (uint8_t)0;
    }
}

}  // namespace

extern "C" int bmx_refine_distance(bmx_ctx *ctx, const uint8_t *lab, int64_t Z, int64_t Y, int64_t X, float *dist,
                                   uint8_t *d)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(lab && dist && d, "bmx_refine_distance: null pointer");
    BMX_CHECK(Z >= 1 && Y >= 1 && X >= 1 && Z * Y * X < ((int64_t)1 << 32) && Z * Y < ((int64_t)1 << 31),
              "bmx_refine_distance: bad dims");
    Geo g{(int)Z, (int)Y, (int)X, X, X * Y, Z * Y * X};
    k_geo_init<<<bmx_stride_blocks(g.V, 256), 256, 0, ctx->stream>>>(lab, g.V, dist);
    BMX_LAUNCH_CHECK();
    const unsigned blocks = bmx_blocks(Z * Y, 256);
    const int Tmax = 4 * (int)(Z - 1) + 2 * (int)(Y - 1) + (int)(X - 1);
    if (Y >= 3 && X >= 3) {
        for (int T = 0; T <= Tmax; ++T) {
            k_geo_front<true><<<blocks, 256, 0, ctx->stream>>>(dist, g, T);
            BMX_LAUNCH_CHECK();
        }
        for (int T = 0; T <= Tmax; ++T) {
            k_geo_front<false><<<blocks, 256, 0, ctx->stream>>>(dist, g, T);
            BMX_LAUNCH_CHECK();
        }
    }
    k_geo_threshold<<<bmx_stride_blocks(g.V, 256), 256, 0, ctx->stream>>>(dist, g.V, d);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

extern "C" int bmx_refine_stats(bmx_ctx *ctx, const uint8_t *img, const uint8_t *lab, const uint8_t *d, int64_t Z,
                                int64_t Y, int64_t X, uint64_t *stats)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(img && lab && d && stats, "bmx_refine_stats: null pointer");
    BMX_CHECK(Z >= 1 && Y >= 1 && X >= 1 && Z * Y * X < ((int64_t)1 << 32), "bmx_refine_stats: bad dims");
    Geo g{(int)Z, (int)Y, (int)X, X, X * Y, Z * Y * X};
    k_refine_stats<<<(unsigned)Z, 256, 0, ctx->stream>>>(img, lab, d, g, (unsigned long long *)stats);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

extern "C" int bmx_refine_apply(bmx_ctx *ctx, const uint8_t *img, const uint8_t *d, int64_t Z, int64_t Y,
                                int64_t X, const uint8_t *lut, const uint8_t *label, uint8_t *out)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(img && d && lut && label && out, "bmx_refine_apply: null pointer");
    BMX_CHECK(Z >= 1 && Y >= 1 && X >= 1 && Z * Y * X < ((int64_t)1 << 32), "bmx_refine_apply: bad dims");
    Geo g{(int)Z, (int)Y, (int)X, X, X * Y, Z * Y * X};
    k_refine_apply<<<bmx_stride_blocks(g.V, 256), 256, 0, ctx->stream>>>(img, d, g, lut, label, out);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

extern "C" int bmx_acwe(bmx_ctx *ctx, const uint8_t *img, const uint8_t *lab, int64_t Z, int64_t Y, int64_t X,
                        const uint8_t *labels, int nlabels, double alpha, int smooth, int steps, uint8_t *out,
                        uint8_t *tmp)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(img && lab && labels && out && tmp, "bmx_acwe: null pointer");
    BMX_CHECK(Z >= 1 && Y >= 1 && X >= 1 && Z * Y * X < ((int64_t)1 << 32), "bmx_acwe: bad dims");
    BMX_CHECK(nlabels >= 1 && nlabels <= 256 && smooth >= 0 && steps >= 0, "bmx_acwe: bad arguments");
    Geo g{(int)Z, (int)Y, (int)X, X, X * Y, Z * Y * X};
    const Patterns P = make_patterns();
    BMX_HIP(hipMemcpyAsync(out, lab, (size_t)g.V, hipMemcpyDeviceToDevice, ctx->stream));
    unsigned long long *d_sc = nullptr;
    BMX_HIP(hipMalloc((void **)&d_sc, 512 * sizeof(unsigned long long)));
    std::vector<unsigned long long> sc(512);
    const unsigned grid = bmx_stride_blocks(g.V, 256);
    for (int s = 0; s < steps && rc == BMX_OK; ++s) {
        // 1. means of the labels listed (others 0)
        if (hipMemsetAsync(d_sc, 0, 512 * sizeof(unsigned long long), ctx->stream) != hipSuccess ||
            (k_label_sums<<<grid < 2048 ? grid : 2048, 256, 0, ctx->stream>>>(img, out, g.V, d_sc, d_sc + 256),
             hipGetLastError() != hipSuccess) ||
            hipMemcpyAsync(sc.data(), d_sc, 512 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream) !=
                hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess) {
            rc = bmx_set_error(BMX_E_DEVICE, "bmx_acwe: label means failed");
            break;
        }
        LabelMeans mean{};
        for (int j = 0; j < nlabels; ++j) {
            const int k = labels[j];
            if (sc[256 + k]) mean.m[k] = (float)((double)sc[k] / (double)sc[256 + k]);
        }
        // 2. evolution out -> tmp, then back
        k_evolution<<<grid, 256, 0, ctx->stream>>>(out, img, g, mean, alpha, tmp);
        if (hipGetLastError() != hipSuccess) {
            rc = bmx_set_error(BMX_E_DEVICE, "bmx_acwe: launch failed");
            break;
        }
        uint8_t *cur = tmp, *nxt = out;
        // 3. curvop per label
        for (int j = 0; j < nlabels && rc == BMX_OK; ++j) {
            const uint8_t k = labels[j];
            for (int r = 0; r < smooth && rc == BMX_OK; ++r) {
                const bool first_erosion = (r % 2 == 0);
                rc = launch_morph(ctx, first_erosion, cur, nxt, g, k, P);
                if (rc == BMX_OK) rc = launch_morph(ctx, !first_erosion, nxt, cur, g, k, P);
            }
        }
        if (rc == BMX_OK && cur != out && hipMemcpyAsync(out, cur, (size_t)g.V, hipMemcpyDeviceToDevice,
                                                         ctx->stream) != hipSuccess)
            rc = bmx_set_error(BMX_E_DEVICE, "bmx_acwe: copy failed");
    }
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess && rc == BMX_OK) rc = bmx_set_error(BMX_E_DEVICE, "bmx_acwe: %s", hipGetErrorString(e));
    (void)hipFree(d_sc);
    return rc;
}
