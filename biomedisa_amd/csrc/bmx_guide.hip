// bmx_guide.hip -- surrogate-guided walking: encode the network input from
// the image and the sparse labels, turn the network output into a label
// proposal and an uncertain set, select the region whose seeds must walk, and
// merge walker and network labels.  No counterpart in the reference.
//
// Guided rule: U = {voxels whose top softmax probability < thr} united with
// the predicted label boundaries dilated by `brad` voxels (Chebyshev).  Seeds
// within `srad` voxels (Chebyshev) of U walk; the result takes the walker's
// argmax on U and the network's label elsewhere.  Far field: seeds in the ring
// between srad and a wider radius walk k times fewer walks with hits weighted by
// k (bmx_walk_opts.hit_weight), so the hits on U also carry an unbiased estimate
// of the contributions of the seeds beyond srad.
#include "bmx_internal.h"

#include <algorithm>

namespace {

__device__ __forceinline__ uint16_t f2bf(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Network input on a padded grid (Zp,Yp,Xp) >= (Z,Y,X), 32 bf16 channels:
// 0: image/127.5 - 1; 1: 1 on seed voxels (labelled planes, label not
// ignored); 2 + l: one-hot of label index l on seed voxels (l < 30).
constexpr int kNetIn = 32;
template <typename T>
__global__ void k_net_input(const T *__restrict__ img, const uint8_t *__restrict__ lab,
                            const uint8_t *__restrict__ flags, int Z, int Y, int X, int Zp, int Yp, int Xp,
                            uint16_t *__restrict__ out)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np_ = (int64_t)Zp * Yp * Xp;
    if (v >= np_) return;
    const int x = (int)(v % Xp), y = (int)((v / Xp) % Yp), z = (int)(v / ((int64_t)Xp * Yp));
    uint16_t ch[kNetIn];
#pragma unroll
    for (int c = 0; c < kNetIn; ++c) ch[c] = 0;
    if (z < Z && y < Y && x < X) {
        const int64_t i = ((int64_t)z * Y + y) * X + x;
        ch[0] = f2bf((float)img[i] * (1.0f / 127.5f) - 1.0f);
        const bool plane = flags[z] | flags[Z + y] | flags[Z + Y + x];
        const uint8_t l = lab[i];
        if (plane && l != BMX_LABEL_NONE) {
            ch[1] = 0x3F80;  // 1.0
            if (l < kNetIn - 2) ch[2 + l] = 0x3F80;
        }
    }
    uint4 *o = (uint4 *)(out + v * kNetIn);
#pragma unroll
    for (int q = 0; q < kNetIn / 8; ++q)
        o[q] = make_uint4(ch[8 * q] | (ch[8 * q + 1] << 16), ch[8 * q + 2] | (ch[8 * q + 3] << 16),
                          ch[8 * q + 4] | (ch[8 * q + 5] << 16), ch[8 * q + 6] | (ch[8 * q + 7] << 16));
}

// 1x1x1 head on the padded grid -> per voxel of the (Z,Y,X) volume: argmax
// label index (first maximum) and an uncertainty flag (top probability < thr).
// Logits accumulate in channel order in FP32 (no contraction), the softmax
// denominator in label order with the engine's exp, so the result is
// reproducible on the CPU.
__global__ void k_head_predict(const uint16_t *__restrict__ x, int Z, int Y, int X, int Zp, int Yp, int Xp,
                               int L, const float *__restrict__ W, const float *__restrict__ b, float thr,
                               uint8_t *__restrict__ pred, uint8_t *__restrict__ unc)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (int64_t)Z * Y * X) return;
    const int xx = (int)(v % X), yy = (int)((v / X) % Y), zz = (int)(v / ((int64_t)X * Y));
    const int64_t p = ((int64_t)zz * Yp + yy) * Xp + xx;
    float xin[16];
    for (int c = 0; c < 16; ++c) xin[c] = bf2f(x[p * 16 + c]);
    float best = -1e30f, den = 0.0f;
    int bl = 0;
    float lg[32];
    for (int l = 0; l < L; ++l) {
        float s = b[l];
        for (int c = 0; c < 16; ++c) s += W[l * 16 + c] * xin[c];
        lg[l] = s;
        if (s > best) { best = s; bl = l; }
    }
    for (int l = 0; l < L; ++l) den += bmx_expf(lg[l] - best);
    pred[v] = (uint8_t)bl;
    unc[v] = (1.0f / den) < thr ? 1 : 0;
}

// U seed: voxels whose predicted label differs from a 6-neighbour's.
__global__ void k_boundary(const uint8_t *__restrict__ pred, int Z, int Y, int X, uint8_t *__restrict__ out)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (int64_t)Z * Y * X) return;
    const int x = (int)(v % X), y = (int)((v / X) % Y), z = (int)(v / ((int64_t)X * Y));
    const uint8_t p = pred[v];
    const int64_t sy = X, sz = (int64_t)X * Y;
    bool b = false;
    if (x > 0) b |= pred[v - 1] != p;
    if (x < X - 1) b |= pred[v + 1] != p;
    if (y > 0) b |= pred[v - sy] != p;
    if (y < Y - 1) b |= pred[v + sy] != p;
    if (z > 0) b |= pred[v - sz] != p;
    if (z < Z - 1) b |= pred[v + sz] != p;
    out[v] = b ? 1 : 0;
}

// Dilation by r along one axis: out[j] = 1 where some in[k] != 0 with
// |k - j| <= r on the same line.  One pass per line with a read-ahead of r:
// after in[0..i] have been read, `last` is the last set index so far, and
// out[i - r] = (last >= i - 2r), since every index up to (i - r) + r has been
// read.  Each voxel is read once and written once.
//
// Along z (axis 0) and y (axis 1): one thread per line, adjacent threads on
// adjacent x, so every step is a coalesced row access.  out != in.
__global__ void k_dilate_lines(const uint8_t *__restrict__ in, int Z, int Y, int X, int axis, int r,
                               uint8_t *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nlines = axis == 0 ? (int64_t)Y * X : (int64_t)Z * X;
    if (t >= nlines) return;
    const int n = axis == 0 ? Z : Y;
    const int64_t st = axis == 0 ? (int64_t)X * Y : (int64_t)X;
    // first voxel of the line: (0, y, x) for z lines, (z, 0, x) for y lines
    const int64_t base = axis == 0 ? t : (t / X) * (int64_t)X * Y + t % X;
    int last = -2 * r - 1;
#pragma unroll 8
    for (int i = 0; i < n + r; ++i) {
        if (i < n && in[base + (int64_t)i * st]) last = i;
        const int j = i - r;
        if (j >= 0) out[base + (int64_t)j * st] = last >= j - r ? 1 : 0;
    }
}

// Along x (rows contiguous in memory): a block stages `rpb` consecutive rows
// (at most kDilRows, rpb * X <= kDilLds bytes) in LDS with coalesced 4-byte
// copies, one thread per row runs the one-pass dilation on its row in place
// (out[j] is written after in[j] has been read, and nothing reads in[j]
// again), and the block copies the rows back.  In place on `buf`.
constexpr int kDilRows = 64, kDilThreads = 256, kDilLds = 64 * 516;
__global__ __launch_bounds__(kDilThreads) void k_dilate_rows(uint8_t *__restrict__ buf, int64_t nrows, int X, int r,
                                                             int rpb)
{
    __shared__ uint32_t s4[kDilLds / 4];
    uint8_t *s = (uint8_t *)s4;
    const int tid = (int)threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * rpb;
    const int rows = (int)min((int64_t)rpb, nrows - r0);
    const int bytes = rows * X;
    uint8_t *g = buf + r0 * X;
    // 4-byte copies where the block's range is 4-aligned (all blocks when
    // rpb * X % 4 == 0, the buffer being 256-aligned), bytes otherwise
    const bool al = ((uintptr_t)g & 3u) == 0;
    const int words = al ? bytes / 4 : 0;
    for (int k = tid; k < words; k += kDilThreads) s4[k] = ((const uint32_t *)g)[k];
    for (int k = 4 * words + tid; k < bytes; k += kDilThreads) s[k] = g[k];
    __syncthreads();
    if (tid < rows) {
        uint8_t *row = s + tid * X;
        int last = -2 * r - 1;
#pragma unroll 8
        for (int i = 0; i < X + r; ++i) {
            if (i < X && row[i]) last = i;
            const int j = i - r;
            if (j >= 0) row[j] = last >= j - r ? 1 : 0;
        }
    }
    __syncthreads();
    for (int k = tid; k < words; k += kDilThreads) ((uint32_t *)g)[k] = s4[k];
    for (int k = 4 * words + tid; k < bytes; k += kDilThreads) g[k] = s[k];
}

// Rows longer than kDilLds bytes: one thread per row, the same one pass, into out.
__global__ void k_dilate_rows_long(const uint8_t *__restrict__ in, int64_t nrows, int X, int r,
                                   uint8_t *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows) return;
    const uint8_t *row = in + t * X;
    uint8_t *orow = out + t * X;
    int last = -2 * r - 1;
    for (int i = 0; i < X + r; ++i) {
        if (i < X && row[i]) last = i;
        const int j = i - r;
        if (j >= 0) orow[j] = last >= j - r ? 1 : 0;
    }
}

__global__ void k_or(uint8_t *__restrict__ a, const uint8_t *__restrict__ b, int64_t n)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) a[v] = (a[v] | b[v]) ? 1 : 0;
}

__global__ void k_andnot(uint8_t *__restrict__ a, const uint8_t *__restrict__ b, int64_t n)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) a[v] = (a[v] && !b[v]) ? 1 : 0;
}

struct LabelValues {
    uint8_t v[256];
};

// out = U ? values[argmax_l hits] : values[pred]
__global__ void k_merge_guided(const uint32_t *__restrict__ hits, int nlabels, int64_t V,
                               const uint8_t *__restrict__ pred, const uint8_t *__restrict__ U, LabelValues lv,
                               uint8_t *__restrict__ out)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    if (!U[v]) {
        out[v] = lv.v[pred[v]];
        return;
    }
    uint32_t b = hits[v];
    int bi = 0;
    for (int l = 1; l < nlabels; ++l) {
        const uint32_t h = hits[(int64_t)l * V + v];
        if (h > b) { b = h; bi = l; }
    }
    out[v] = lv.v[bi];
}

// Repair test of the guided result: a voxel outside U whose walker hits (at
// least minhits in all) have a first-maximum label other than the proposal's.
__global__ void k_guided_disagree(const uint32_t *__restrict__ hits, int nlabels, int64_t V,
                                  const uint8_t *__restrict__ pred, const uint8_t *__restrict__ U,
                                  uint32_t minhits, uint8_t *__restrict__ D)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    uint8_t d = 0;
    if (!U[v]) {
        uint64_t tot = 0;
        uint32_t b = 0;
        int bi = 0;
        for (int l = 0; l < nlabels; ++l) {
            const uint32_t h = hits[(int64_t)l * V + v];
            tot += h;
            if (h > b) { b = h; bi = l; }
        }
        d = (tot >= minhits && tot > 0 && bi != pred[v]) ? 1 : 0;
    }
    D[v] = d;
}

// hits[:, v] = hits2[:, v] and U[v] = 1 on the repaired voxels U2.
__global__ void k_repair_hits(uint32_t *__restrict__ hits, const uint32_t *__restrict__ hits2, int nlabels,
                              int64_t V, const uint8_t *__restrict__ U2, uint8_t *__restrict__ U)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V || !U2[v]) return;
    for (int l = 0; l < nlabels; ++l) hits[(int64_t)l * V + v] = hits2[(int64_t)l * V + v];
    U[v] = 1;
}

// out = U ? a : values[pred]  (a already holds label values)
__global__ void k_select_labels(const uint8_t *__restrict__ a, const uint8_t *__restrict__ pred,
                                const uint8_t *__restrict__ U, LabelValues lv, int64_t V,
                                uint8_t *__restrict__ out)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    out[v] = U[v] ? a[v] : lv.v[pred[v]];
}

// Soft training targets from walker hits: t[v][l] = hits[l][v] / sum_l hits,
// on a padded grid (zeros and mask 0 outside the volume or where no walker
// arrived); L >= nlabels outputs (extra labels get 0).
__global__ void k_hits_target(const uint32_t *__restrict__ hits, int nlabels, int Z, int Y, int X, int Zp,
                              int Yp, int Xp, int L, int hard, float *__restrict__ target,
                              uint8_t *__restrict__ vmask)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np_ = (int64_t)Zp * Yp * Xp;
    if (v >= np_) return;
    const int x = (int)(v % Xp), y = (int)((v / Xp) % Yp), z = (int)(v / ((int64_t)Xp * Yp));
    const int64_t V = (int64_t)Z * Y * X;
    const bool in = z < Z && y < Y && x < X;
    const int64_t i = in ? ((int64_t)z * Y + y) * X + x : 0;
    if (hard) {
        // one-hot of the walker's argmax: first maximum, index 0 when every count is 0
        int best = 0;
        uint32_t bh = 0;
        if (in)
            for (int l = 0; l < nlabels; ++l) {
                const uint32_t h = hits[(int64_t)l * V + i];
                if (h > bh) { bh = h; best = l; }
            }
        for (int l = 0; l < L; ++l) target[v * L + l] = (in && l == best) ? 1.0f : 0.0f;
        vmask[v] = in ? 1 : 0;
        return;
    }
    float tot = 0.0f;
    if (in)
        for (int l = 0; l < nlabels; ++l) tot += (float)hits[(int64_t)l * V + i];
    const bool ok = in && tot > 0.0f;
    for (int l = 0; l < L; ++l)
        target[v * L + l] = (ok && l < nlabels) ? (float)hits[(int64_t)l * V + i] / tot : 0.0f;
    vmask[v] = ok ? 1 : 0;
}

// Training label of every grid voxel: the first argmax index of the walker's
// hits (0 when all counts are 0, as bmx_argmax), and the in-volume mask.
__global__ void k_hits_index(const uint32_t *__restrict__ hits, int nlabels, int Z, int Y, int X, int Zp, int Yp,
                             int Xp, uint8_t *__restrict__ idx, uint8_t *__restrict__ vmask)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np_ = (int64_t)Zp * Yp * Xp;
    if (v >= np_) return;
    const int x = (int)(v % Xp), y = (int)((v / Xp) % Yp), z = (int)(v / ((int64_t)Xp * Yp));
    const bool in = z < Z && y < Y && x < X;
    int best = 0;
    if (in) {
        const int64_t V = (int64_t)Z * Y * X, i = ((int64_t)z * Y + y) * X + x;
        uint32_t bh = 0;
        for (int l = 0; l < nlabels; ++l) {
            const uint32_t h = hits[(int64_t)l * V + i];
            if (h > bh) { bh = h; best = l; }
        }
    }
    idx[v] = (uint8_t)best;
    vmask[v] = in ? 1 : 0;
}

int dilate(bmx_ctx *ctx, uint8_t *buf, uint8_t *tmp, int Z, int Y, int X, int r)
{
    if (r <= 0) return BMX_OK;
    const int64_t V = (int64_t)Z * Y * X, nrows = (int64_t)Z * Y;
    // x in place (or via tmp for long rows), then y into tmp, z back into buf
    if (X <= kDilLds) {
        const int rpb = std::min(kDilRows, kDilLds / X);
        k_dilate_rows<<<bmx_blocks(nrows, rpb), kDilThreads, 0, ctx->stream>>>(buf, nrows, X, r, rpb);
    } else {
        k_dilate_rows_long<<<bmx_blocks(nrows, 256), 256, 0, ctx->stream>>>(buf, nrows, X, r, tmp);
        BMX_HIP(hipMemcpyAsync(buf, tmp, (size_t)V, hipMemcpyDeviceToDevice, ctx->stream));
    }
    k_dilate_lines<<<bmx_blocks((int64_t)Z * X, 256), 256, 0, ctx->stream>>>(buf, Z, Y, X, 1, r, tmp);
    k_dilate_lines<<<bmx_blocks((int64_t)Y * X, 256), 256, 0, ctx->stream>>>(tmp, Z, Y, X, 0, r, buf);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

}  // namespace

extern "C" {

int bmx_net_input(bmx_ctx *ctx, const void *img, int img_dtype, const uint8_t *lab, const uint8_t *flags,
                  int Z, int Y, int X, int Zp, int Yp, int Xp, uint16_t *out)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(img && lab && flags && out, "bmx_net_input: null pointer");
    BMX_CHECK(Zp >= Z && Yp >= Y && Xp >= X, "bmx_net_input: padded grid smaller than the volume");
    const int64_t np_ = (int64_t)Zp * Yp * Xp;
    BMX_CHECK(np_ < ((int64_t)1 << 32) - 256, "bmx_net_input: grid exceeds one dispatch");
    if (img_dtype == BMX_IMG_U8)
        k_net_input<uint8_t><<<bmx_blocks(np_, 256), 256, 0, ctx->stream>>>((const uint8_t *)img, lab, flags, Z,
                                                                          Y, X, Zp, Yp, Xp, out);
    else
        k_net_input<float><<<bmx_blocks(np_, 256), 256, 0, ctx->stream>>>((const float *)img, lab, flags, Z, Y,
                                                                        X, Zp, Yp, Xp, out);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_head_predict(bmx_ctx *ctx, const uint16_t *x, int Z, int Y, int X, int Zp, int Yp, int Xp, int L,
                     const float *W, const float *b, float thr, uint8_t *pred, uint8_t *unc)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(x && W && b && pred && unc && L >= 1 && L <= 32, "bmx_head_predict: bad arguments");
    const int64_t V = (int64_t)Z * Y * X;
    BMX_CHECK(V < ((int64_t)1 << 32) - 256, "bmx_head_predict: volume exceeds one dispatch");
    k_head_predict<<<bmx_blocks(V, 256), 256, 0, ctx->stream>>>(x, Z, Y, X, Zp, Yp, Xp, L, W, b, thr, pred, unc);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_guided_masks(bmx_ctx *ctx, const uint8_t *pred, const uint8_t *unc, int Z, int Y, int X, int brad,
                     int srad, uint8_t *U, uint8_t *need, uint8_t *tmp)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(pred && unc && U && need && tmp, "bmx_guided_masks: null pointer");
    BMX_CHECK(brad >= 0 && srad >= 0, "bmx_guided_masks: negative radius");
    const int64_t V = (int64_t)Z * Y * X;
    BMX_CHECK(V < ((int64_t)1 << 32) - 256, "bmx_guided_masks: volume exceeds one dispatch");
    k_boundary<<<bmx_blocks(V, 256), 256, 0, ctx->stream>>>(pred, Z, Y, X, U);
    BMX_LAUNCH_CHECK();
    rc = dilate(ctx, U, tmp, Z, Y, X, brad);
    if (rc) return rc;
    k_or<<<bmx_blocks(V, 256), 256, 0, ctx->stream>>>(U, unc, V);
    BMX_LAUNCH_CHECK();
    BMX_HIP(hipMemcpyAsync(need, U, (size_t)V, hipMemcpyDeviceToDevice, ctx->stream));
    return dilate(ctx, need, tmp, Z, Y, X, srad);
}

int bmx_ring_mask(bmx_ctx *ctx, const uint8_t *need, int Z, int Y, int X, int r, uint8_t *ring, uint8_t *tmp)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(need && ring && tmp, "bmx_ring_mask: null pointer");
    BMX_CHECK(r >= 0, "bmx_ring_mask: negative radius");
    const int64_t V = (int64_t)Z * Y * X;
    BMX_CHECK(V < ((int64_t)1 << 32) - 256, "bmx_ring_mask: volume exceeds one dispatch");
    BMX_HIP(hipMemcpyAsync(ring, need, (size_t)V, hipMemcpyDeviceToDevice, ctx->stream));
    rc = dilate(ctx, ring, tmp, Z, Y, X, r);
    if (rc) return rc;
    k_andnot<<<bmx_blocks(V, 256), 256, 0, ctx->stream>>>(ring, need, V);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_mask_andnot(bmx_ctx *ctx, uint8_t *a, const uint8_t *b, int64_t n)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(a && b && a != b, "bmx_mask_andnot: null or aliased pointers");
    BMX_CHECK(n >= 0 && n < ((int64_t)1 << 32) - 256, "bmx_mask_andnot: n out of range");
    if (n == 0) return BMX_OK;
    k_andnot<<<bmx_blocks(n, 256), 256, 0, ctx->stream>>>(a, b, n);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_guided_disagree(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int64_t V, const uint8_t *pred,
                        const uint8_t *U, uint32_t minhits, uint8_t *D)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(hits && pred && U && D && nlabels >= 1 && nlabels <= 255, "bmx_guided_disagree: bad arguments");
    BMX_CHECK(V >= 0 && V < ((int64_t)1 << 32) - 256, "bmx_guided_disagree: V out of range");
    if (V == 0) return BMX_OK;
    k_guided_disagree<<<bmx_blocks(V, 256), 256, 0, ctx->stream>>>(hits, nlabels, V, pred, U, minhits, D);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_repair_masks(bmx_ctx *ctx, uint8_t *D, const uint8_t *U, int Z, int Y, int X, int r, int srad,
                     uint8_t *need, uint8_t *tmp)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(D && U && need && tmp, "bmx_repair_masks: null pointer");
    BMX_CHECK(r >= 0 && srad >= 0, "bmx_repair_masks: negative radius");
    const int64_t V = (int64_t)Z * Y * X;
    BMX_CHECK(V < ((int64_t)1 << 32) - 256, "bmx_repair_masks: volume exceeds one dispatch");
    rc = dilate(ctx, D, tmp, Z, Y, X, r);
    if (rc) return rc;
    k_andnot<<<bmx_blocks(V, 256), 256, 0, ctx->stream>>>(D, U, V);
    BMX_LAUNCH_CHECK();
    BMX_HIP(hipMemcpyAsync(need, D, (size_t)V, hipMemcpyDeviceToDevice, ctx->stream));
    return dilate(ctx, need, tmp, Z, Y, X, srad);
}

int bmx_repair_hits(bmx_ctx *ctx, uint32_t *hits, const uint32_t *hits2, int nlabels, int64_t V,
                    const uint8_t *U2, uint8_t *U)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(hits && hits2 && U2 && U && nlabels >= 1 && nlabels <= 255, "bmx_repair_hits: bad arguments");
    BMX_CHECK(V >= 0 && V < ((int64_t)1 << 32) - 256, "bmx_repair_hits: V out of range");
    if (V == 0) return BMX_OK;
    k_repair_hits<<<bmx_blocks(V, 256), 256, 0, ctx->stream>>>(hits, hits2, nlabels, V, U2, U);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_merge_guided(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int64_t V, const uint8_t *values,
                     const uint8_t *pred, const uint8_t *U, uint8_t *out)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(hits && values && pred && U && out && nlabels >= 1 && nlabels <= 255,
              "bmx_merge_guided: bad arguments");
    BMX_CHECK(V >= 0 && V < ((int64_t)1 << 32) - 256, "bmx_merge_guided: V out of range");
    LabelValues lv;
    for (int i = 0; i < 256; ++i) lv.v[i] = i < nlabels ? values[i] : 0;
    k_merge_guided<<<bmx_blocks(V, 256), 256, 0, ctx->stream>>>(hits, nlabels, V, pred, U, lv, out);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_hits_target(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int Z, int Y, int X, int Zp, int Yp,
                    int Xp, int L, int hard, float *target, uint8_t *vmask)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(hits && target && vmask && L >= nlabels, "bmx_hits_target: bad arguments");
    const int64_t np_ = (int64_t)Zp * Yp * Xp;
    BMX_CHECK(np_ < ((int64_t)1 << 32) - 256, "bmx_hits_target: grid exceeds one dispatch");
    k_hits_target<<<bmx_blocks(np_, 256), 256, 0, ctx->stream>>>(hits, nlabels, Z, Y, X, Zp, Yp, Xp, L, hard,
                                                                target, vmask);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_hits_index(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int Z, int Y, int X, int Zp, int Yp, int Xp,
                   uint8_t *idx, uint8_t *vmask)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(hits && idx && vmask && nlabels >= 1 && nlabels <= 255, "bmx_hits_index: bad arguments");
    BMX_CHECK(Zp >= Z && Yp >= Y && Xp >= X, "bmx_hits_index: padded grid smaller than the volume");
    const int64_t np_ = (int64_t)Zp * Yp * Xp;
    BMX_CHECK(np_ < ((int64_t)1 << 32) - 256, "bmx_hits_index: grid exceeds one dispatch");
    k_hits_index<<<bmx_blocks(np_, 256), 256, 0, ctx->stream>>>(hits, nlabels, Z, Y, X, Zp, Yp, Xp, idx, vmask);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_select_labels(bmx_ctx *ctx, const uint8_t *a, const uint8_t *pred, const uint8_t *U, int nlabels,
                      const uint8_t *values, int64_t V, uint8_t *out)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(a && pred && U && values && out, "bmx_select_labels: null pointer");
    BMX_CHECK(V >= 0 && V < ((int64_t)1 << 32) - 256, "bmx_select_labels: V out of range");
    LabelValues lv;
    for (int i = 0; i < 256; ++i) lv.v[i] = i < nlabels ? values[i] : 0;
    k_select_labels<<<bmx_blocks(V, 256), 256, 0, ctx->stream>>>(a, pred, U, lv, V, out);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

}  // extern "C"
