// bmx_net.hip -- kernels of the learned proposal (a 3-D conv encoder-decoder):
// BF16 3x3x3 convolution forward / input-gradient / weight-gradient on MFMA
// with FP32 accumulation, and the small ops around it (2x average pooling,
// 2x nearest upsampling, channel concatenation, 1x1x1 softmax head with
// soft-target cross-entropy, Adam).  No counterpart in the reference: this
// is the surrogate that lets the walker skip seeds in confident regions.
//
// Layout: activations are channels-last bf16 [N][D][H][W][C]; conv weights
// are bf16 [Co][27][Ci] (tap = (kz*3 + ky)*3 + kx, offsets k-1), with an FP32
// master copy kept by the optimiser; gradients of weights are FP32.
//
// MFMA tiling (v_mfma_f32_16x16x16bf16_1k): lane l supplies A[l%16][4(l/16)..+3]
// and B[4(l/16)..+3][l%16] and receives D[4(l/16)+i][l%16], i = 0..3.  A wave
// computes 16*RB voxels (RB 16-voxel row blocks along x; RB = 2 in the
// gradient kernels, 4 or 8 in the forward) by 16*NB channels; the K loop runs
// over the 27 taps and 16-channel chunks of the input.
#include "bmx_internal.h"

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// float -> bf16, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t f2bf(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ short4_t ld4(const uint16_t *p)
{
    const uint2 v = *(const uint2 *)p;
    short4_t r;
    r[0] = (short)(v.x & 0xFFFF);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xFFFF);
    r[3] = (short)(v.y >> 16);
    return r;
}

__device__ __forceinline__ float4_t mfma16(short4_t a, short4_t b, float4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct Shape {
    int N, D, H, W;
};

// Forward: y = act(conv(x, w) + b).  grid.x: voxel groups of 64*RB (4 waves x
// 16*RB), grid.y: output-channel groups of 16*NB.  Each lane decodes its own
// voxel's coordinates, so any grid shape works; requires Ci % 16 == 0 and
// Co % (16*NB) == 0.  With a second source xb (Cb channels on the half-size
// grid), the input is concat(x, up2(xb)) with Ci = Ca + Cb, read in place: a
// 16-channel chunk comes from x below Ca and from xb at the voxel's parent
// (z/2, y/2, x/2) above, so neither the upsampled tensor nor the concatenation
// is materialised.
template <int NB, bool CAT, int RB = 2>
__global__ __launch_bounds__(256) void k_conv_fwd(const uint16_t *__restrict__ x, Shape s, int Ci,
                                                  const uint16_t *__restrict__ w,
                                                  const float *__restrict__ bias,
                                                  uint16_t *__restrict__ y, int Co, int relu,
                                                  const uint16_t *__restrict__ xb, int Ca)
{
    const int Cb = Ci - Ca;
    const int Dh = s.D / 2, Hh = s.H / 2, Wh = s.W / 2;
    const int lane = threadIdx.x & 63;
    const int64_t nvox = (int64_t)s.N * s.D * s.H * s.W;
    const int64_t v0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * (16 * RB);
    if (v0 >= nvox) return;
    const int co0 = blockIdx.y * 16 * NB;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    int cz[RB], cy[RB], cx[RB], cn[RB];
    bool vin[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
        int64_t v = v0 + rb * 16 + r;
        vin[rb] = v < nvox;
        cx[rb] = (int)(v % s.W); v /= s.W;
        cy[rb] = (int)(v % s.H); v /= s.H;
        cz[rb] = (int)(v % s.D);
        cn[rb] = (int)(v / s.D);
    }
    float4_t acc[RB][NB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[rb][nb] = float4_t{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 27; ++tap) {
        const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
        const uint16_t *src[RB], *srcb[RB];
        bool ok[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            const int z = cz[rb] + dz, yy = cy[rb] + dy, xx = cx[rb] + dx;
            ok[rb] = vin[rb] && (unsigned)z < (unsigned)s.D && (unsigned)yy < (unsigned)s.H &&
                     (unsigned)xx < (unsigned)s.W;
            const int zc = ok[rb] ? z : 0, yc = ok[rb] ? yy : 0, xc = ok[rb] ? xx : 0;
            const int64_t nv = (((int64_t)cn[rb] * s.D + zc) * s.H + yc) * s.W + xc;
            src[rb] = x + nv * (CAT ? Ca : Ci) + kq;
            if (CAT) {
                const int64_t nh = (((int64_t)cn[rb] * Dh + zc / 2) * Hh + yc / 2) * Wh + xc / 2;
                srcb[rb] = xb + nh * Cb + kq;
            }
        }
        const uint16_t *wt = w + ((int64_t)(co0 + r) * 27 + tap) * Ci + kq;
        for (int c0 = 0; c0 < Ci; c0 += 16) {
            short4_t a[RB];
#pragma unroll
            for (int rb = 0; rb < RB; ++rb) {
                a[rb] = (CAT && c0 >= Ca) ? ld4(srcb[rb] + (c0 - Ca)) : ld4(src[rb] + c0);
                if (!ok[rb]) a[rb] = short4_t{0, 0, 0, 0};
            }
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const short4_t b = ld4(wt + (int64_t)nb * 16 * 27 * Ci + c0);
#pragma unroll
                for (int rb = 0; rb < RB; ++rb) acc[rb][nb] = mfma16(a[rb], b, acc[rb][nb]);
            }
        }
    }
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int co = co0 + nb * 16 + r;
            const float bb = bias ? bias[co] : 0.0f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t v = v0 + rb * 16 + kq + i;
                if (v >= nvox) continue;
                float val = acc[rb][nb][i] + bb;
                if (relu) val = fmaxf(val, 0.0f);
                y[v * Co + co] = f2bf(val);
            }
        }
}

// Forward for 16 output channels from a plain input of at most 48 channels
// (the first encoder layer) with the input staged in LDS: a block owns a 32 (x)
// by 4 (y) tile of output rows and marches along ZC planes of z; the input
// planes z-1, z, z+1 of the tile plus a 1-voxel halo rotate through three LDS
// slots, so each input voxel is read from memory about 1.6 times instead of
// once per tap.  Wave w computes the 32 voxels of row y0 + w with the same MFMA
// sequence as k_conv_fwd<1, false> (taps outer, 16-channel chunks inner), so
// the results are bit-identical to it.
constexpr int kZmTX = 32, kZmTY = 4, kZmZC = 16;
constexpr int kZmPX = kZmTX + 2, kZmPY = kZmTY + 2, kZmPlane = kZmPX * kZmPY;
constexpr int kZmMaxCi = 48;

__host__ __device__ constexpr int zm_pitch(int Ci) { return Ci + 4; }   // bf16 per LDS voxel (bank spread)

__global__ __launch_bounds__(256) void k_conv_fwd_zm(const uint16_t *__restrict__ x, Shape s, int Ci,
                                                     const uint16_t *__restrict__ w,
                                                     const float *__restrict__ bias,
                                                     uint16_t *__restrict__ y, int relu)
{
    extern __shared__ __align__(16) uint16_t zm_lds[];
    const int P = zm_pitch(Ci);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    const int x0 = blockIdx.x * kZmTX, y0 = blockIdx.y * kZmTY;
    const int nzc = (s.D + kZmZC - 1) / kZmZC;
    const int n = blockIdx.z / nzc;
    const int zb = (blockIdx.z % nzc) * kZmZC, ze = min(s.D, zb + kZmZC);
    const int groups = Ci / 4;                       // 8-byte channel groups per voxel

    // input plane z into LDS slot (z + 3) % 3: zeros outside the grid
    auto load_plane = [&](int z) {
        uint16_t *dst = zm_lds + ((z + 3) % 3) * kZmPlane * P;
        for (int u = tid; u < kZmPlane * groups; u += 256) {
            const int p = u / groups, c = (u - p * groups) * 4;
            const int ly = p / kZmPX, lx = p - ly * kZmPX;
            const int yy = y0 - 1 + ly, xx = x0 - 1 + lx;
            uint2 v = make_uint2(0u, 0u);
            if ((unsigned)z < (unsigned)s.D && (unsigned)yy < (unsigned)s.H && (unsigned)xx < (unsigned)s.W) {
                const int64_t vox = (((int64_t)n * s.D + z) * s.H + yy) * s.W + xx;
                v = *(const uint2 *)(x + vox * Ci + c);
            }
            *(uint2 *)(dst + p * P + c) = v;
        }
    };

    const int yrow = y0 + wv;
    const float bb = bias ? bias[r] : 0.0f;
    load_plane(zb - 1);
    load_plane(zb);
    for (int z = zb; z < ze; ++z) {
        load_plane(z + 1);
        __syncthreads();
        float4_t acc[2] = {float4_t{0.f, 0.f, 0.f, 0.f}, float4_t{0.f, 0.f, 0.f, 0.f}};
        for (int tap = 0; tap < 27; ++tap) {
            const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
            const uint16_t *pl = zm_lds + ((z + dz + 3) % 3) * kZmPlane * P;
            const uint16_t *a0 = pl + ((wv + 1 + dy) * kZmPX + (1 + r + dx)) * P + kq;
            const uint16_t *a1 = a0 + 16 * P;
            const uint16_t *wt = w + ((int64_t)r * 27 + tap) * Ci + kq;
            for (int c0 = 0; c0 < Ci; c0 += 16) {
                const short4_t av0 = ld4(a0 + c0), av1 = ld4(a1 + c0);
                const short4_t b = ld4(wt + c0);
                acc[0] = mfma16(av0, b, acc[0]);
                acc[1] = mfma16(av1, b, acc[1]);
            }
        }
        if (yrow < s.H) {
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int xx = x0 + rb * 16 + kq + i;
                    if (xx >= s.W) continue;
                    float val = acc[rb][i] + bb;
                    if (relu) val = fmaxf(val, 0.0f);
                    const int64_t v = (((int64_t)n * s.D + z) * s.H + yrow) * s.W + xx;
                    y[v * 16 + r] = f2bf(val);
                }
        }
        __syncthreads();
    }
}

// Input gradient: dx[v][ci] = sum_tap sum_co dy[v - off(tap)][co] * w[co][tap][ci],
// times (xact[v][ci] > 0) when xact is given (ReLU of the layer below).
// wt: weights re-laid out as [Ci][27][Co].
template <int NB>
__global__ __launch_bounds__(256) void k_conv_dgrad(const uint16_t *__restrict__ dy, Shape s, int Co,
                                                    const uint16_t *__restrict__ wt,
                                                    const uint16_t *__restrict__ xact,
                                                    uint16_t *__restrict__ dx, int Ci)
{
    const int lane = threadIdx.x & 63;
    const int64_t nvox = (int64_t)s.N * s.D * s.H * s.W;
    const int64_t v0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
    if (v0 >= nvox) return;
    const int ci0 = blockIdx.y * 16 * NB;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    int cz[2], cy[2], cx[2], cn[2];
    bool vin[2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
        int64_t v = v0 + rb * 16 + r;
        vin[rb] = v < nvox;
        cx[rb] = (int)(v % s.W); v /= s.W;
        cy[rb] = (int)(v % s.H); v /= s.H;
        cz[rb] = (int)(v % s.D);
        cn[rb] = (int)(v / s.D);
    }
    float4_t acc[2][NB];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[rb][nb] = float4_t{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 27; ++tap) {
        // output voxel u = v - off contributes through tap
        const int dz = 1 - tap / 9, dyo = 1 - (tap / 3) % 3, dxo = 1 - tap % 3;
        const uint16_t *src[2];
        bool ok[2];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb) {
            const int z = cz[rb] + dz, yy = cy[rb] + dyo, xx = cx[rb] + dxo;
            ok[rb] = vin[rb] && (unsigned)z < (unsigned)s.D && (unsigned)yy < (unsigned)s.H &&
                     (unsigned)xx < (unsigned)s.W;
            const int64_t nv = (((int64_t)cn[rb] * s.D + (ok[rb] ? z : 0)) * s.H + (ok[rb] ? yy : 0)) * s.W +
                               (ok[rb] ? xx : 0);
            src[rb] = dy + (ok[rb] ? nv : 0) * Co + kq;
        }
        const uint16_t *wr = wt + ((int64_t)(ci0 + r) * 27 + tap) * Co + kq;
        for (int c0 = 0; c0 < Co; c0 += 16) {
            short4_t a[2];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) {
                a[rb] = ld4(src[rb] + c0);
                if (!ok[rb]) a[rb] = short4_t{0, 0, 0, 0};
            }
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) {
                const short4_t b = ld4(wr + (int64_t)nb * 16 * 27 * Co + c0);
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) acc[rb][nb] = mfma16(a[rb], b, acc[rb][nb]);
            }
        }
    }
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int ci = ci0 + nb * 16 + r;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t v = v0 + rb * 16 + kq + i;
                if (v >= nvox) continue;
                float val = acc[rb][nb][i];
                if (xact && !(bf2f(xact[v * Ci + ci]) > 0.0f)) val = 0.0f;
                dx[v * Ci + ci] = f2bf(val);
            }
        }
}

// Weight gradient: dw[co][tap][ci] += sum_v dy[v][co] * x[v + off(tap)][ci];
// db[co] += sum_v dy[v][co].  One wave: 16 output channels x 16 input
// channels of one tap, reducing over a chunk of `chunk` voxels (multiple of
// 16); partial sums are added to dw with float atomics.
// grid: (ceil(nvox / chunk), 27 * (Ci/16), Co/16), block 64.
__global__ __launch_bounds__(64) void k_conv_wgrad(const uint16_t *__restrict__ x, Shape s, int Ci,
                                                   const uint16_t *__restrict__ dy, int Co, int64_t chunk,
                                                   float *__restrict__ dw, float *__restrict__ db)
{
    const int lane = threadIdx.x;
    const int64_t nvox = (int64_t)s.N * s.D * s.H * s.W;
    const int64_t vb = (int64_t)blockIdx.x * chunk;
    const int tap = blockIdx.y % 27, cib = blockIdx.y / 27;
    const int co0 = blockIdx.z * 16, ci0 = cib * 16;
    const int dz = tap / 9 - 1, dyo = (tap / 3) % 3 - 1, dxo = tap % 3 - 1;
    const int r = lane & 15, kq = (lane >> 4) * 4;
    float4_t acc = float4_t{0.f, 0.f, 0.f, 0.f};
    float bacc = 0.0f;
    const int64_t vend = min(vb + chunk, nvox);
    for (int64_t k0 = vb; k0 < vend; k0 += 16) {
        // A[m=co][k=v] = dy[v][co]; B[k=v][n=ci] = x[v+off][ci]
        short4_t a, b;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t v = k0 + kq + i;
            uint16_t av = 0, bv = 0;
            if (v < vend) {
                av = dy[v * Co + co0 + r];
                int64_t t = v;
                const int xx = (int)(t % s.W) + dxo; t /= s.W;
                const int yy = (int)(t % s.H) + dyo; t /= s.H;
                const int z = (int)(t % s.D) + dz;
                const int n = (int)(t / s.D);
                if ((unsigned)z < (unsigned)s.D && (unsigned)yy < (unsigned)s.H && (unsigned)xx < (unsigned)s.W)
                    bv = x[((((int64_t)n * s.D + z) * s.H + yy) * s.W + xx) * Ci + ci0 + r];
            }
            a[i] = (short)av;
            b[i] = (short)bv;
        }
        acc = mfma16(a, b, acc);
        // the bias gradient rides on one (tap, ci-block) of each co-block
        if (tap == 13 && cib == 0 && lane < 16) {
            for (int j = 0; j < 16; ++j) {
                const int64_t v = k0 + j;
                if (v < vend) bacc += bf2f(dy[v * Co + co0 + lane]);
            }
        }
    }
    // D[m=co][n=ci]: lane holds rows kq..kq+3 (co), column r (ci)
#pragma unroll
    for (int i = 0; i < 4; ++i)
        atomicAdd(&dw[((int64_t)(co0 + kq + i) * 27 + tap) * Ci + ci0 + r], acc[i]);
    if (tap == 13 && cib == 0 && lane < 16 && db) atomicAdd(&db[co0 + lane], bacc);
}

// g = (a > 0) ? g : 0  (backward of a ReLU whose output is a).
__global__ void k_relu_mask(uint16_t *__restrict__ g, const uint16_t *__restrict__ a, int64_t n)
{
    for (int64_t i = bmx_gtid(); i < n; i += bmx_gstride())
        if (!(bf2f(a[i]) > 0.0f)) g[i] = 0;
}

// [Co][27][Ci] -> [Ci][27][Co]
__global__ void k_transpose_w(const uint16_t *__restrict__ w, int Co, int Ci, uint16_t *__restrict__ wt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)Co * 27 * Ci) return;
    const int ci = (int)(i % Ci);
    const int tap = (int)((i / Ci) % 27);
    const int co = (int)(i / ((int64_t)Ci * 27));
    wt[((int64_t)ci * 27 + tap) * Co + co] = w[i];
}

// 2x2x2 average pooling (channels-last), and its backward.
__global__ void k_pool2(const uint16_t *__restrict__ x, Shape s, int C, uint16_t *__restrict__ y)
{
    const int Do = s.D / 2, Ho = s.H / 2, Wo = s.W / 2;
    const int64_t n_out = (int64_t)s.N * Do * Ho * Wo * C;
    for (int64_t i = bmx_gtid(); i < n_out; i += bmx_gstride()) {
        const int c = (int)(i % C);
        int64_t t = i / C;
        const int xo = (int)(t % Wo); t /= Wo;
        const int yo = (int)(t % Ho); t /= Ho;
        const int zo = (int)(t % Do);
        const int n = (int)(t / Do);
        float acc = 0.0f;
        for (int k = 0; k < 8; ++k) {
            const int z = 2 * zo + (k >> 2), yy = 2 * yo + ((k >> 1) & 1), xx = 2 * xo + (k & 1);
            acc += bf2f(x[((((int64_t)n * s.D + z) * s.H + yy) * s.W + xx) * C + c]);
        }
        y[i] = f2bf(acc * 0.125f);
    }
}

__global__ void k_pool2_bwd(const uint16_t *__restrict__ dy, Shape s, int C, uint16_t *__restrict__ dx)
{
    const int64_t n_in = (int64_t)s.N * s.D * s.H * s.W * C;
    for (int64_t i = bmx_gtid(); i < n_in; i += bmx_gstride()) {
        const int c = (int)(i % C);
        int64_t t = i / C;
        const int xx = (int)(t % s.W); t /= s.W;
        const int yy = (int)(t % s.H); t /= s.H;
        const int z = (int)(t % s.D);
        const int n = (int)(t / s.D);
        const int64_t o =
            ((((int64_t)n * (s.D / 2) + z / 2) * (s.H / 2) + yy / 2) * (s.W / 2) + xx / 2) * C + c;
        dx[i] = f2bf(bf2f(dy[o]) * 0.125f);
    }
}

// Nearest 2x upsampling of x (shape s = the small grid) into y, and its
// backward (sum over each 2x2x2 block).
__global__ void k_up2(const uint16_t *__restrict__ x, Shape s, int C, uint16_t *__restrict__ y)
{
    const int D2 = s.D * 2, H2 = s.H * 2, W2 = s.W * 2;
    const int64_t n_out = (int64_t)s.N * D2 * H2 * W2 * C;
    for (int64_t i = bmx_gtid(); i < n_out; i += bmx_gstride()) {
        const int c = (int)(i % C);
        int64_t t = i / C;
        const int xx = (int)(t % W2); t /= W2;
        const int yy = (int)(t % H2); t /= H2;
        const int z = (int)(t % D2);
        const int n = (int)(t / D2);
        y[i] = x[((((int64_t)n * s.D + z / 2) * s.H + yy / 2) * s.W + xx / 2) * C + c];
    }
}

__global__ void k_up2_bwd(const uint16_t *__restrict__ dy, Shape s, int C, uint16_t *__restrict__ dx)
{
    const int64_t n_small = (int64_t)s.N * s.D * s.H * s.W * C;
    for (int64_t i = bmx_gtid(); i < n_small; i += bmx_gstride()) {
        const int c = (int)(i % C);
        int64_t t = i / C;
        const int xo = (int)(t % s.W); t /= s.W;
        const int yo = (int)(t % s.H); t /= s.H;
        const int zo = (int)(t % s.D);
        const int n = (int)(t / s.D);
        float acc = 0.0f;
        for (int k = 0; k < 8; ++k) {
            const int z = 2 * zo + (k >> 2), yy = 2 * yo + ((k >> 1) & 1), xx = 2 * xo + (k & 1);
            acc += bf2f(dy[((((int64_t)n * 2 * s.D + z) * 2 * s.H + yy) * 2 * s.W + xx) * C + c]);
        }
        dx[i] = f2bf(acc);
    }
}

// y[v] = concat(a[v][0..Ca), b[v][0..Cb)); split is its backward.
__global__ void k_concat(const uint16_t *__restrict__ a, int Ca, const uint16_t *__restrict__ b, int Cb,
                         int64_t nvox, uint16_t *__restrict__ y)
{
    const int C = Ca + Cb;
    for (int64_t i = bmx_gtid(); i < nvox * C; i += bmx_gstride()) {
        const int64_t v = i / C;
        const int c = (int)(i - v * C);
        y[i] = c < Ca ? a[v * Ca + c] : b[v * Cb + (c - Ca)];
    }
}

__global__ void k_split(const uint16_t *__restrict__ y, int64_t nvox, uint16_t *__restrict__ a, int Ca,
                        uint16_t *__restrict__ b, int Cb, int accumulate_a)
{
    const int C = Ca + Cb;
    for (int64_t i = bmx_gtid(); i < nvox * C; i += bmx_gstride()) {
        const int64_t v = i / C;
        const int c = (int)(i - v * C);
        if (c < Ca) {
            float val = bf2f(y[i]);
            if (accumulate_a) val += bf2f(a[v * Ca + c]);
            a[v * Ca + c] = f2bf(val);
        } else {
            b[v * Cb + (c - Ca)] = y[i];
        }
    }
}

// 1x1x1 head: logits = W x + b over C=16 input channels, L outputs;
// prob = softmax(logits).  With a target (soft labels, float [v][L]) it also
// writes dx (bf16, scale/nvox * W^T (prob - target)), accumulates dW, db
// (float atomics per wave) and the summed cross-entropy into *loss.
__global__ __launch_bounds__(256) void k_head(const uint16_t *__restrict__ x, int64_t nvox, int L,
                                              const float *__restrict__ W, const float *__restrict__ b,
                                              float *__restrict__ prob, const float *__restrict__ target,
                                              const uint8_t *__restrict__ tidx,
                                              const uint8_t *__restrict__ vmask, float scale,
                                              uint16_t *__restrict__ dx, float *__restrict__ dW,
                                              float *__restrict__ db, float *__restrict__ loss)
{
    constexpr int C = 16, LM = 32;
    const bool train = target || tidx;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = v < nvox;
    float xin[C];
    float lg[LM];
    float gl[LM];
    float ls = 0.0f;
    for (int c = 0; c < C; ++c) xin[c] = act ? bf2f(x[v * C + c]) : 0.0f;
    float mx = -1e30f;
    for (int l = 0; l < LM; ++l) {
        float s = 0.0f;
        if (l < L) {
            s = b[l];
            for (int c = 0; c < C; ++c) s += W[l * C + c] * xin[c];
            mx = fmaxf(mx, s);
        }
        lg[l] = s;
    }
    float den = 0.0f;
    for (int l = 0; l < LM; ++l)
        if (l < L) {
            lg[l] = __expf(lg[l] - mx);
            den += lg[l];
        }
    for (int l = 0; l < LM; ++l) {
        gl[l] = 0.0f;
        if (l < L) {
            const float p = lg[l] / den;
            if (act && prob) prob[v * L + l] = p;
            if (train) {
                const float w = (act && (!vmask || vmask[v])) ? 1.0f : 0.0f;
                const float t = !act ? 0.0f : (target ? target[v * L + l] : (tidx[v] == l ? 1.0f : 0.0f));
                gl[l] = w * scale * (p - t);
                if (w > 0.0f && t > 0.0f) ls -= t * __logf(fmaxf(p, 1e-30f));
            }
        }
    }
    if (!train) return;
    if (act) {
        for (int c = 0; c < C; ++c) {
            float s = 0.0f;
            for (int l = 0; l < L; ++l) s += W[l * C + c] * gl[l];
            dx[v * C + c] = f2bf(s);
        }
    }
    // wave reductions of dW, db, loss
    for (int l = 0; l < L; ++l) {
        float g = gl[l];
        for (int off = 32; off > 0; off >>= 1) g += __shfl_xor(g, off, 64);
        if ((threadIdx.x & 63) == 0) atomicAdd(&db[l], g);
        for (int c = 0; c < C; ++c) {
            float gw = gl[l] * xin[c];
            for (int off = 32; off > 0; off >>= 1) gw += __shfl_xor(gw, off, 64);
            if ((threadIdx.x & 63) == 0) atomicAdd(&dW[l * C + c], gw);
        }
    }
    for (int off = 32; off > 0; off >>= 1) ls += __shfl_xor(ls, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, ls);
}

// Adam step on FP32 master weights; refreshes the bf16 compute copy.
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, uint16_t *__restrict__ pbf, int64_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2, float wd)
{
    for (int64_t i = bmx_gtid(); i < n; i += bmx_gstride()) {
        const float gi = g[i] + wd * p[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        p[i] = pi;
        if (pbf) pbf[i] = f2bf(pi);
    }
}

__global__ void k_add_bf16(uint16_t *__restrict__ y, const uint16_t *__restrict__ x, int64_t n)
{
    for (int64_t i = bmx_gtid(); i < n; i += bmx_gstride()) y[i] = f2bf(bf2f(y[i]) + bf2f(x[i]));
}

__global__ void k_f2bf(const float *__restrict__ a, uint16_t *__restrict__ b, int64_t n)
{
    for (int64_t i = bmx_gtid(); i < n; i += bmx_gstride()) b[i] = f2bf(a[i]);
}

template <int NB>
int launch_fwd(bmx_ctx *ctx, const uint16_t *x, Shape s, int Ci, const uint16_t *w, const float *b, uint16_t *y,
               int Co, int relu, const uint16_t *xb = nullptr, int Ca = 0)
{
    // 16-voxel row blocks per wave: RB * NB independent accumulator chains in
    // flight (the MFMA latency, not its throughput, bounds these layers); the
    // counts are the fastest measured at 512^3 (two row blocks: 44 / 21 / 13 ms
    // for NB = 1 / 2 / 4 in the decoder; these: 38 / 15.5 / 8.4 ms)
    constexpr int RB = NB == 2 ? 8 : 4;
    const int64_t nvox = (int64_t)s.N * s.D * s.H * s.W;
    dim3 grid((unsigned)((nvox + 64 * RB - 1) / (64 * RB)), (unsigned)(Co / (16 * NB)));
    if (xb)
        k_conv_fwd<NB, true, RB><<<grid, 256, 0, ctx->stream>>>(x, s, Ci, w, b, y, Co, relu, xb, Ca);
    else
        k_conv_fwd<NB, false, RB><<<grid, 256, 0, ctx->stream>>>(x, s, Ci, w, b, y, Co, relu, nullptr, Ci);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

// The 16-output-channel forward: staged in LDS (k_conv_fwd_zm) for a plain
// input of at most 48 channels, where it measured 1.4x faster at 512^3 (first
// encoder layer: 27.9 against 39.6 ms); k_conv_fwd otherwise (staging the
// decoder's concatenated input measured slower: 50.1 against 44.4 ms).
int launch_fwd16(bmx_ctx *ctx, const uint16_t *x, Shape s, int Ci, const uint16_t *w, const float *b, uint16_t *y,
                 int relu, const uint16_t *xb = nullptr, int Ca = 0)
{
    if (Ci > kZmMaxCi || xb) return launch_fwd<1>(ctx, x, s, Ci, w, b, y, 16, relu, xb, Ca);
    const int nzc = (s.D + kZmZC - 1) / kZmZC;
    BMX_CHECK((int64_t)s.N * nzc < 65536, "conv3d: too many z chunks for one dispatch");
    dim3 grid((unsigned)((s.W + kZmTX - 1) / kZmTX), (unsigned)((s.H + kZmTY - 1) / kZmTY), (unsigned)(s.N * nzc));
    const size_t lds = (size_t)3 * kZmPlane * zm_pitch(Ci) * sizeof(uint16_t);
    k_conv_fwd_zm<<<grid, 256, lds, ctx->stream>>>(x, s, Ci, w, b, y, relu);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

template <int NB>
int launch_dgrad(bmx_ctx *ctx, const uint16_t *dy, Shape s, int Co, const uint16_t *wt, const uint16_t *xact,
                 uint16_t *dx, int Ci)
{
    const int64_t nvox = (int64_t)s.N * s.D * s.H * s.W;
    dim3 grid((unsigned)((nvox + 127) / 128), (unsigned)(Ci / (16 * NB)));
    k_conv_dgrad<NB><<<grid, 256, 0, ctx->stream>>>(dy, s, Co, wt, xact, dx, Ci);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int check_conv_shape(int N, int D, int H, int W, int Ci, int Co)
{
    BMX_CHECK(N >= 1 && D >= 1 && H >= 1 && W >= 1, "conv3d: bad grid shape");
    BMX_CHECK(Ci >= 16 && Ci % 16 == 0 && Co >= 16 && Co % 16 == 0, "conv3d: channels must be multiples of 16");
    BMX_CHECK((int64_t)N * D * H * W * (Ci > Co ? Ci : Co) < ((int64_t)1 << 40), "conv3d: tensor too large");
    // two work-items per voxel in the forward/dgrad grids; one dispatch holds < 2^32
    BMX_CHECK((int64_t)N * D * H * W < ((int64_t)1 << 31) - 128, "conv3d: grid exceeds one dispatch");
    return BMX_OK;
}

}  // namespace

extern "C" {

int bmx_conv3d_fwd(bmx_ctx *ctx, const uint16_t *x, int N, int D, int H, int W, int Ci, const uint16_t *w,
                   const float *b, uint16_t *y, int Co, int relu)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    rc = check_conv_shape(N, D, H, W, Ci, Co);
    if (rc) return rc;
    BMX_CHECK(x && w && y, "bmx_conv3d_fwd: null pointer");
    Shape s{N, D, H, W};
    if (Co % 64 == 0) return launch_fwd<4>(ctx, x, s, Ci, w, b, y, Co, relu);
    if (Co % 32 == 0) return launch_fwd<2>(ctx, x, s, Ci, w, b, y, Co, relu);
    return launch_fwd16(ctx, x, s, Ci, w, b, y, relu);
}

int bmx_conv3d_fwd_cat(bmx_ctx *ctx, const uint16_t *a, int N, int D, int H, int W, int Ca, const uint16_t *b,
                       int Cb, const uint16_t *w, const float *bias, uint16_t *y, int Co, int relu)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    rc = check_conv_shape(N, D, H, W, Ca + Cb, Co);
    if (rc) return rc;
    BMX_CHECK(a && b && w && y, "bmx_conv3d_fwd_cat: null pointer");
    BMX_CHECK(Ca >= 16 && Ca % 16 == 0 && Cb >= 16 && Cb % 16 == 0, "bmx_conv3d_fwd_cat: channels");
    BMX_CHECK(D % 2 == 0 && H % 2 == 0 && W % 2 == 0, "bmx_conv3d_fwd_cat: grid must be even");
    Shape s{N, D, H, W};
    const int Ci = Ca + Cb;
    if (Co % 64 == 0) return launch_fwd<4>(ctx, a, s, Ci, w, bias, y, Co, relu, b, Ca);
    if (Co % 32 == 0) return launch_fwd<2>(ctx, a, s, Ci, w, bias, y, Co, relu, b, Ca);
    return launch_fwd16(ctx, a, s, Ci, w, bias, y, relu, b, Ca);
}

int bmx_conv3d_dgrad(bmx_ctx *ctx, const uint16_t *dy, int N, int D, int H, int W, int Co, const uint16_t *w,
                     const uint16_t *xact, uint16_t *dx, int Ci, uint16_t *wt_scratch)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    rc = check_conv_shape(N, D, H, W, Ci, Co);
    if (rc) return rc;
    BMX_CHECK(dy && w && dx && wt_scratch, "bmx_conv3d_dgrad: null pointer");
    const int64_t nw = (int64_t)Co * 27 * Ci;
    k_transpose_w<<<bmx_blocks(nw, 256), 256, 0, ctx->stream>>>(w, Co, Ci, wt_scratch);
    BMX_LAUNCH_CHECK();
    Shape s{N, D, H, W};
    if (Ci % 64 == 0) return launch_dgrad<4>(ctx, dy, s, Co, wt_scratch, xact, dx, Ci);
    if (Ci % 32 == 0) return launch_dgrad<2>(ctx, dy, s, Co, wt_scratch, xact, dx, Ci);
    return launch_dgrad<1>(ctx, dy, s, Co, wt_scratch, xact, dx, Ci);
}

int bmx_conv3d_wgrad(bmx_ctx *ctx, const uint16_t *x, int N, int D, int H, int W, int Ci, const uint16_t *dy,
                     int Co, float *dw, float *db)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    rc = check_conv_shape(N, D, H, W, Ci, Co);
    if (rc) return rc;
    BMX_CHECK(x && dy && dw, "bmx_conv3d_wgrad: null pointer");
    Shape s{N, D, H, W};
    const int64_t nvox = (int64_t)N * D * H * W;
    int64_t chunk = 4096;
    dim3 grid((unsigned)((nvox + chunk - 1) / chunk), (unsigned)(27 * (Ci / 16)), (unsigned)(Co / 16));
    k_conv_wgrad<<<grid, 64, 0, ctx->stream>>>(x, s, Ci, dy, Co, chunk, dw, db);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_pool2(bmx_ctx *ctx, const uint16_t *x, int N, int D, int H, int W, int C, uint16_t *y)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(x && y && D % 2 == 0 && H % 2 == 0 && W % 2 == 0, "bmx_pool2: bad arguments");
    const int64_t n = (int64_t)N * (D / 2) * (H / 2) * (W / 2) * C;
    k_pool2<<<bmx_stride_blocks(n, 256), 256, 0, ctx->stream>>>(x, Shape{N, D, H, W}, C, y);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_pool2_bwd(bmx_ctx *ctx, const uint16_t *dy, int N, int D, int H, int W, int C, uint16_t *dx)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(dy && dx && D % 2 == 0 && H % 2 == 0 && W % 2 == 0, "bmx_pool2_bwd: bad arguments");
    const int64_t n = (int64_t)N * D * H * W * C;
    k_pool2_bwd<<<bmx_stride_blocks(n, 256), 256, 0, ctx->stream>>>(dy, Shape{N, D, H, W}, C, dx);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_up2(bmx_ctx *ctx, const uint16_t *x, int N, int D, int H, int W, int C, uint16_t *y)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(x && y, "bmx_up2: null pointer");
    const int64_t n = (int64_t)N * D * H * W * 8 * C;
    k_up2<<<bmx_stride_blocks(n, 256), 256, 0, ctx->stream>>>(x, Shape{N, D, H, W}, C, y);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_up2_bwd(bmx_ctx *ctx, const uint16_t *dy, int N, int D, int H, int W, int C, uint16_t *dx)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(dy && dx, "bmx_up2_bwd: null pointer");
    const int64_t n = (int64_t)N * D * H * W * C;
    k_up2_bwd<<<bmx_stride_blocks(n, 256), 256, 0, ctx->stream>>>(dy, Shape{N, D, H, W}, C, dx);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_concat(bmx_ctx *ctx, const uint16_t *a, int Ca, const uint16_t *b, int Cb, int64_t nvox, uint16_t *y)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(a && b && y, "bmx_concat: null pointer");
    k_concat<<<bmx_stride_blocks(nvox * (Ca + Cb), 256), 256, 0, ctx->stream>>>(a, Ca, b, Cb, nvox, y);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_split(bmx_ctx *ctx, const uint16_t *y, int64_t nvox, uint16_t *a, int Ca, uint16_t *b, int Cb,
              int accumulate_a)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(y && a && b, "bmx_split: null pointer");
    k_split<<<bmx_stride_blocks(nvox * (Ca + Cb), 256), 256, 0, ctx->stream>>>(y, nvox, a, Ca, b, Cb, accumulate_a);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_head(bmx_ctx *ctx, const uint16_t *x, int64_t nvox, int L, const float *W, const float *b, float *prob,
             const float *target, const uint8_t *target_idx, const uint8_t *vmask, float scale, uint16_t *dx,
             float *dW, float *db, float *loss)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(x && W && b && L >= 1 && L <= 32, "bmx_head: bad arguments (L <= 32)");
    BMX_CHECK(!(target && target_idx), "bmx_head: give soft targets or target indices, not both");
    BMX_CHECK(!(target || target_idx) || (dx && dW && db && loss), "bmx_head: training needs dx, dW, db, loss");
    BMX_CHECK(nvox >= 0 && nvox < ((int64_t)1 << 32) - 256, "bmx_head: nvox %lld exceeds one dispatch",
              (long long)nvox);
    k_head<<<bmx_blocks(nvox, 256), 256, 0, ctx->stream>>>(x, nvox, L, W, b, prob, target, target_idx, vmask,
                                                          scale, dx, dW, db, loss);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_adam(bmx_ctx *ctx, float *p, const float *g, float *m, float *v, uint16_t *pbf, int64_t n, float lr,
             float beta1, float beta2, float eps, int step, float weight_decay)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(p && g && m && v && step >= 1, "bmx_adam: bad arguments");
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    k_adam<<<bmx_stride_blocks(n, 256), 256, 0, ctx->stream>>>(p, g, m, v, pbf, n, lr, beta1, beta2, eps, c1, c2,
                                                       weight_decay);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_relu_mask(bmx_ctx *ctx, uint16_t *g, const uint16_t *a, int64_t n)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(g && a, "bmx_relu_mask: null pointer");
    k_relu_mask<<<bmx_stride_blocks(n, 256), 256, 0, ctx->stream>>>(g, a, n);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_add_bf16(bmx_ctx *ctx, uint16_t *y, const uint16_t *x, int64_t n)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(y && x, "bmx_add_bf16: null pointer");
    k_add_bf16<<<bmx_stride_blocks(n, 256), 256, 0, ctx->stream>>>(y, x, n);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_f32_to_bf16(bmx_ctx *ctx, const float *a, uint16_t *b, int64_t n)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(a && b, "bmx_f32_to_bf16: null pointer");
    k_f2bf<<<bmx_stride_blocks(n, 256), 256, 0, ctx->stream>>>(a, b, n);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

}  // extern "C"
