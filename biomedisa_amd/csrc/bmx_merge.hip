// bmx_merge.hip -- per-voxel reductions over the label axis of the hit volume:
// the argmax merge (rw_small.py:200-201 + get_labels,
// biomedisa_helper.py:1209-1213) and the top-3 uncertainty
// (MAX_SRC + UNCERTAINTY_SRC, opencl_kernels.py:54-133; rw_small.py:152-176).
#include "bmx_internal.h"

namespace {

struct LabelValues {
    uint8_t v[256];
};

// Each thread handles 4 consecutive voxels with 16-byte loads per label.
__global__ __launch_bounds__(256) void k_argmax4(const uint32_t *__restrict__ hits, int nlabels,
                                                 int64_t V, LabelValues lv, uint8_t *__restrict__ out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t v0 = q * 4;
    if (v0 >= V) return;
    if (v0 + 4 <= V) {
        uint4 best = *(const uint4 *)(hits + v0);
        uint32_t bi0 = 0, bi1 = 0, bi2 = 0, bi3 = 0;
        for (int l = 1; l < nlabels; ++l) {
            const uint4 h = *(const uint4 *)(hits + (int64_t)l * V + v0);
            if (h.x > best.x) { best.x = h.x; bi0 = l; }
            if (h.y > best.y) { best.y = h.y; bi1 = l; }
            if (h.z > best.z) { best.z = h.z; bi2 = l; }
            if (h.w > best.w) { best.w = h.w; bi3 = l; }
        }
        const uint32_t packed = (uint32_t)lv.v[bi0] | ((uint32_t)lv.v[bi1] << 8) |
                                ((uint32_t)lv.v[bi2] << 16) | ((uint32_t)lv.v[bi3] << 24);
        *(uint32_t *)(out + v0) = packed;
    } else {
        for (int64_t v = v0; v < V; ++v) {
            uint32_t b = hits[v];
            int bi = 0;
            for (int l = 1; l < nlabels; ++l) {
                const uint32_t h = hits[(int64_t)l * V + v];
                if (h > b) { b = h; bi = l; }
            }
            out[v] = lv.v[bi];
        }
    }
}

__global__ __launch_bounds__(256) void k_argmax1(const uint32_t *__restrict__ hits, int nlabels,
                                                 int64_t V, LabelValues lv, uint8_t *__restrict__ out)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    uint32_t b = hits[v];
    int bi = 0;
    for (int l = 1; l < nlabels; ++l) {
        const uint32_t h = hits[(int64_t)l * V + v];
        if (h > b) { b = h; bi = l; }
    }
    out[v] = lv.v[bi];
}

__global__ __launch_bounds__(256) void k_uncertainty(const uint32_t *__restrict__ hits, int nlabels,
                                                     int Z, int Y, int X, uint8_t *__restrict__ out)
{
    const int64_t V = (int64_t)Z * Y * X;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const int64_t flat = (int64_t)X * Y;
    const int z = (int)(v / flat);
    const int64_t r = v - z * flat;
    const int y = (int)(r / X);
    const int x = (int)(r - (int64_t)y * X);
    if (z <= 0 || z >= Z - 1 || y <= 0 || y >= Y - 1 || x <= 0 || x >= X - 1) {
        out[v] = 0;
        return;
    }
    // maxbuf[0] <= maxbuf[1] <= maxbuf[2] maintained as in MAX_SRC, labels in order.
    float m0 = 0.0f, m1 = 0.0f, m2 = 0.0f;
    for (int l = 0; l < nlabels; ++l) {
        const float t = (float)hits[(int64_t)l * V + v];
        if (t > m2) { m0 = m1; m1 = m2; m2 = t; }
        else if (t > m1) { m0 = m1; m1 = t; }
        else if (t > m0) { m0 = t; }
    }
    float max2 = m2;
    if (max2 == 0.0f) max2 = 1.0f;
    const float t0 = m0 / max2;
    const float t1 = m1 / max2;
    const float u = 1.0f - (1.0f - t0) * (1.0f - t1);
    out[v] = (uint8_t)(u * 255.0f);
}

}  // namespace

extern "C" {

int bmx_argmax(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int64_t V, const uint8_t *values,
               uint8_t *out)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(hits && values && out, "bmx_argmax: null pointer");
    BMX_CHECK(nlabels >= 1 && nlabels <= 256, "bmx_argmax: nlabels %d not in [1,256]", nlabels);
    BMX_CHECK(V >= 0 && V < ((int64_t)1 << 32) - 256, "bmx_argmax: V %lld not in [0, 2^32)", (long long)V);
    if (V == 0) return BMX_OK;
    LabelValues lv;
    for (int i = 0; i < 256; ++i) lv.v[i] = i < nlabels ? values[i] : 0;
    const bool aligned = ((uintptr_t)hits % 16 == 0) && ((uintptr_t)out % 4 == 0) && (V % 4 == 0);
    if (aligned) {
        const int64_t nq = (V + 3) / 4;
        k_argmax4<<<bmx_blocks(nq, 256), 256, 0, ctx->stream>>>(hits, nlabels, V, lv, out);
    } else {
        k_argmax1<<<bmx_blocks(V, 256), 256, 0, ctx->stream>>>(hits, nlabels, V, lv, out);
    }
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_uncertainty(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int64_t Z, int64_t Y, int64_t X,
                    uint8_t *out)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(hits && out, "bmx_uncertainty: null pointer");
    BMX_CHECK(nlabels >= 1 && nlabels <= 256, "bmx_uncertainty: nlabels %d not in [1,256]", nlabels);
    BMX_CHECK(Z >= 1 && Y >= 1 && X >= 1 && Z * Y * X < (int64_t)1 << 32, "bmx_uncertainty: bad dims");
    const int64_t V = Z * Y * X;
    k_uncertainty<<<bmx_blocks(V, 256), 256, 0, ctx->stream>>>(hits, nlabels, (int)Z, (int)Y, (int)X,
                                                              out);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

}  // extern "C"
