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
// The forward and the input gradient run on k_conv_mm (gfx950's
// v_mfma_f32_16x16x32_bf16, planes streamed through LDS, weights in
// registers; described at the kernel), the weight gradient on
// k_conv_wgrad_mm (the same instruction, with voxels as the reduction).
#include "bmx_internal.h"

#include <algorithm>

namespace {

typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// float -> bf16, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t f2bf(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}


__device__ __forceinline__ void unpack8(uint4 q, float *f)
{
    const uint32_t w4[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        f[2 * j] = bf2f((uint16_t)(w4[j] & 0xFFFFu));
        f[2 * j + 1] = bf2f((uint16_t)(w4[j] >> 16));
    }
}
__device__ __forceinline__ uint4 pack8(const float *f)
{
    return make_uint4((uint32_t)f2bf(f[0]) | ((uint32_t)f2bf(f[1]) << 16), (uint32_t)f2bf(f[2]) | ((uint32_t)f2bf(f[3]) << 16),
                      (uint32_t)f2bf(f[4]) | ((uint32_t)f2bf(f[5]) << 16), (uint32_t)f2bf(f[6]) | ((uint32_t)f2bf(f[7]) << 16));
}

struct Shape {
    int N, D, H, W;
};

// ---------------------------------------------------------------------------
// The forward on the gfx950 MFMA v_mfma_f32_16x16x32_bf16 (K = 32 per
// instruction): lane l supplies A[l%16][8(l/16)..+7] and B[8(l/16)..+7][l%16]
// and receives D[4(l/16)+i][l%16], i = 0..3.
//
// A block (4 waves) computes 16 output channels of a 32 x 8 (x, y) tile of
// output rows over a chunk of kMmZC output planes of one volume, and streams
// the input planes z' = zb-1 .. ze through two LDS buffers: the next plane is
// loaded straight from global memory into one while the other is used.  A
// plane holds the tile plus a 1-voxel halo, channels padded with zeros to CP
// = 32 CH, in 16-byte slots: CP/8 per voxel and one unused slot, so a voxel's
// pitch is CP + 8 bf16 and 16 consecutive voxels at one channel offset fall in
// distinct banks.  Each wave keeps the 27 x CH B fragments of the block's
// output channels in registers, so no other vector-memory load runs during
// the compute and nothing waits on the plane in flight.  Wave w owns output
// rows y0 + 2w, y0 + 2w + 1: four row blocks of 16 voxels.  For every (dy, dx)
// tap and 32-channel chunk of plane z' it reads each row block's A fragment
// once and applies it to the three output planes z'+1, z', z'-1 (taps dz =
// -1, 0, +1), whose accumulators it keeps: plane z'-1 is complete after z' and
// is written out.  K runs in the order (dz, dy, dx, chunk) for every output
// channel, whatever the channel count, the tile or a concatenated input, so
// results do not depend on them.  CAT: the input is concat(x, up2(xb)) with Ca
// channels from x, read in place.  MASK: the epilogue zeroes outputs where
// xact <= 0 (the input gradient through a ReLU) instead of adding a bias.
// SPLIT: a launch convolves input channels coff .. coff+Ci-1 of Cf, and wider
// inputs take several launches of at most 64 channels: SPLIT = 1 writes its
// float32 sums to `part`, 2 adds them to `part`, 3 adds `part` before the
// epilogue (0: a single launch).
constexpr int kMmTX = 32, kMmTY = 8, kMmZC = 32, kMmRB = 4;
constexpr int kMmPX = kMmTX + 2, kMmPlane = kMmPX * (kMmTY + 2);

typedef short short8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) void lds_void_t;

__device__ __forceinline__ float4_t mfma32(short8_t a, short8_t b, float4_t c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                   c, 0, 0, 0);
}

__host__ __device__ constexpr int mm_pitch(int ch) { return 32 * ch + 8; }
__host__ __device__ constexpr size_t mm_lds_bytes(int ch)
{
    return (size_t)2 * kMmPlane * mm_pitch(ch) * sizeof(uint16_t);
}

template <int CH, bool CAT, bool MASK, int SPLIT>
__global__ __launch_bounds__(256) void k_conv_mm(const uint16_t *__restrict__ x, Shape s, int Ci,
                                                 const uint16_t *__restrict__ w, const float *__restrict__ bias,
                                                 uint16_t *__restrict__ y, int Co, int relu,
                                                 const uint16_t *__restrict__ xb, int Ca,
                                                 const uint16_t *__restrict__ xact, int Cf, int coff,
                                                 float *__restrict__ part)
{
    constexpr int CP = 32 * CH, CPG = CP / 8, SPV = CPG + 1, P = mm_pitch(CH);
    constexpr int SLOTS = kMmPlane * SPV, NIT = (SLOTS + 255) / 256;
    constexpr int NSTEP = 9 * CH;   // (dy, dx, chunk) steps of a plane
    extern __shared__ __align__(16) uint16_t mm_lds[];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r = lane & 15, kg = lane >> 4;
    const int x0 = blockIdx.x * kMmTX, y0 = blockIdx.y * kMmTY;
    const int nzc = (s.D + kMmZC - 1) / kMmZC;
    const int ncob = Co / 16;
    const int co0 = (int)(blockIdx.z % (unsigned)ncob) * 16, zn = (int)(blockIdx.z / (unsigned)ncob);
    const int n = zn / nzc;
    const int zb = (zn % nzc) * kMmZC, ze = min(s.D, zb + kMmZC);
    // this launch's Ci input channels are channels coff .. coff + Ci - 1 of the Cf of the input
    const int Cx = CAT ? Ca : Cf, Cb = Cf - Ca;
    const int Dh = s.D / 2, Hh = s.H / 2, Wh = s.W / 2;

    // The slots this lane loads (slot wv*64 + it*256 + lane of every plane, so
    // that a wave instruction fills 64 consecutive slots) as element offsets
    // within a plane of x (>= 0), of xb (encoded -2 - offset) or -1 for a slot
    // no load writes (padded channels, voxels off the grid in y or x, the
    // unused slot of a voxel).
    int soff[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int u = wv * 64 + it * 256 + lane;
        const int p = u / SPV, q = u - p * SPV, c = 8 * q, cf = coff + c;
        const int ly = p / kMmPX, lx = p - ly * kMmPX;
        const int yy = y0 - 1 + ly, xx = x0 - 1 + lx;
        int o = -1;
        if (u < SLOTS && q < CPG && c < Ci && (unsigned)yy < (unsigned)s.H && (unsigned)xx < (unsigned)s.W) {
            if (!CAT || cf < Ca) o = (yy * s.W + xx) * Cx + cf;
            else o = -2 - (((yy / 2) * Wh + xx / 2) * Cb + (cf - Ca));
        }
        soff[it] = o;
    }
    // output element offsets within a plane of y (-1: off the grid)
    int yoff[kMmRB][4];
#pragma unroll
    for (int rb = 0; rb < kMmRB; ++rb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int yrow = y0 + 2 * wv + rb / 2, xx = x0 + 16 * (rb % 2) + 4 * kg + i;
            yoff[rb][i] = (yrow < s.H && xx < s.W) ? (yrow * s.W + xx) * Co + co0 + r : -1;
        }

    short8_t bw[27][CH];
#pragma unroll
    for (int t = 0; t < 27; ++t)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int k0 = 32 * c + 8 * kg;
            bw[t][c] = k0 < Ci ? *(const short8_t *)(w + ((int64_t)(co0 + r) * 27 + t) * Cf + coff + k0)
                               : short8_t{0, 0, 0, 0, 0, 0, 0, 0};
        }

    // plane zp into LDS buffer `buf`; a plane off the grid is cleared
    auto load_plane = [&](int zp, int buf) {
        uint16_t *dst = mm_lds + buf * kMmPlane * P;
        if ((unsigned)zp >= (unsigned)s.D) {
            for (int u = tid; u < SLOTS; u += 256) *(uint4 *)(dst + u * 8) = make_uint4(0u, 0u, 0u, 0u);
            return;
        }
        const uint16_t *xp = x + (((int64_t)n * s.D + zp) * s.H) * s.W * Cx;
        const uint16_t *bp = CAT ? xb + (((int64_t)n * Dh + zp / 2) * Hh) * Wh * Cb : nullptr;
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
            const int o = soff[it];
            const int s0 = wv * 64 + it * 256;
            if (o != -1)
                __builtin_amdgcn_global_load_lds((CAT && o < -1) ? bp + (-2 - o) : xp + o,
                                                 (lds_void_t *)(dst + s0 * 8), 16, 0, 0);
        }
    };
    // slots no load writes are zero in both buffers
    for (int u = tid; u < 2 * SLOTS; u += 256) {
        const int uu = u % SLOTS;
        const int p = uu / SPV, q = uu - p * SPV;
        const int ly = p / kMmPX, lx = p - ly * kMmPX;
        const int yy = y0 - 1 + ly, xx = x0 - 1 + lx;
        if (q < CPG && (8 * q >= Ci || (unsigned)yy >= (unsigned)s.H || (unsigned)xx >= (unsigned)s.W))
            *(uint4 *)(mm_lds + u * 8) = make_uint4(0u, 0u, 0u, 0u);
    }
    __syncthreads();

    float4_t acc[3][kMmRB];   // output planes z'-1, z', z'+1
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int rb = 0; rb < kMmRB; ++rb) acc[q][rb] = float4_t{0.f, 0.f, 0.f, 0.f};

    const float bb = (!MASK && (SPLIT == 0 || SPLIT == 3) && bias) ? bias[co0 + r] : 0.0f;
    // the partial sums and ReLU outputs the epilogue of output plane z reads,
    // fetched before the plane's MFMAs so that their latency is hidden
    float pv[kMmRB][4];
    uint16_t mv[kMmRB][4];
    auto prefetch = [&](int z) {
        const int64_t ybase = (((int64_t)n * s.D + z) * s.H) * s.W * Co;
#pragma unroll
        for (int rb = 0; rb < kMmRB; ++rb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t v = ybase + (yoff[rb][i] < 0 ? 0 : yoff[rb][i]);
                if (SPLIT >= 2) pv[rb][i] = yoff[rb][i] < 0 ? 0.0f : part[v];
                if (MASK) mv[rb][i] = (yoff[rb][i] < 0 || !xact) ? (uint16_t)0x3F80u : xact[v];
            }
    };
    auto store = [&](int z) {
        const int64_t ybase = (((int64_t)n * s.D + z) * s.H) * s.W * Co;
#pragma unroll
        for (int rb = 0; rb < kMmRB; ++rb)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (yoff[rb][i] < 0) continue;
                const int64_t v = ybase + yoff[rb][i];
                float val = acc[0][rb][i];
                if (SPLIT == 1 || SPLIT == 2) {
                    part[v] = SPLIT == 1 ? val : pv[rb][i] + val;
                    continue;
                }
                if (SPLIT == 3) val += pv[rb][i];
                if (MASK) {
                    if (!(bf2f(mv[rb][i]) > 0.0f)) val = 0.0f;
                } else {
                    val += bb;
                    if (relu) val = fmaxf(val, 0.0f);
                }
                y[v] = f2bf(val);
            }
    };

    // A fragments of step k (dy, dx, chunk) of the current plane
    auto read_a = [&](const uint16_t *pl, int k, short8_t *a) {
        const int dy = k / (3 * CH), dx = (k / CH) % 3, c = k % CH;
#pragma unroll
        for (int rb = 0; rb < kMmRB; ++rb) {
            const int row = 2 * wv + rb / 2, col = 16 * (rb % 2) + r;
            a[rb] = *(const short8_t *)(pl + ((row + dy) * kMmPX + col + dx) * P + 32 * c + 8 * kg);
        }
    };

    load_plane(zb - 1, 0);
    __syncthreads();
    int buf = 0;
    for (int zp = zb - 1; zp <= ze; ++zp) {
        if (zp + 1 <= ze) load_plane(zp + 1, buf ^ 1);
        if ((SPLIT >= 2 || MASK) && zp - 1 >= zb) prefetch(zp - 1);
        const uint16_t *pl = mm_lds + buf * kMmPlane * P;
        short8_t a[2][kMmRB];
        read_a(pl, 0, a[0]);
#pragma unroll
        for (int k = 0; k < NSTEP; ++k) {
            if (k + 1 < NSTEP) read_a(pl, k + 1, a[(k + 1) & 1]);
            const int dy = k / (3 * CH), dx = (k / CH) % 3, c = k % CH;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                // slot q holds output plane z' - 1 + q, reached through tap dz = 1 - q (the
                // first and last planes of a chunk also feed planes outside it, which are dropped)
                const int t = (2 - q) * 9 + dy * 3 + dx;
#pragma unroll
                for (int rb = 0; rb < kMmRB; ++rb) acc[q][rb] = mfma32(a[k & 1][rb], bw[t][c], acc[q][rb]);
            }
        }
        if (zp - 1 >= zb) store(zp - 1);
#pragma unroll
        for (int rb = 0; rb < kMmRB; ++rb) {
            acc[0][rb] = acc[1][rb];
            acc[1][rb] = acc[2][rb];
            acc[2][rb] = float4_t{0.f, 0.f, 0.f, 0.f};
        }
        __syncthreads();   // the next plane is in, and every wave is done with this one
        buf ^= 1;
    }
}

// Weight gradient on v_mfma_f32_16x16x32_bf16:
//   dw[co][tap][ci] += sum_v dy[v][co] * x[v + off(tap)][ci],
// a product with M = output channels, N = input channels and K = voxels.
// K runs along x: one K step is the 32 voxels x0 .. x0+31 of one row, and
// lane group g (lanes 16g .. 16g+15) supplies voxels x0+8g .. x0+8g+7, so a
// lane's 8 K values are 8 consecutive voxels of one channel.  A block stages
// its planes channel-major in LDS (the channels-last tensors are transposed
// on the way in), and a fragment is one 16-byte LDS read:
//  - dy plane z: [16*CO_T channels][kWgTY rows][32 voxels], channel pitch
//    kWgTY*32 + 8 (16 channels at one voxel offset fall in distinct banks);
//  - x planes z-1, z, z+1 (a ring of three): [16 channels][kWgTY+2 rows][48],
//    element i of a row holding x = x0 - 8 + i (the halo x0-1 .. x0+32 at
//    7 .. 40).  Tap offset dx = 0 reads at 8g + 8 (aligned); dx = -1 and +1
//    start one element before or after, and their fragment is rebuilt from two
//    aligned reads with v_alignbit (one 16-bit shift per dword).
// A block has three waves, wave w the taps of dz = w - 1: 9 taps x CO_T
// output-channel tiles of 16 x 16 accumulators, over the block's voxels (a
// 32 x kWgTY (x, y) tile and a chunk of zc planes of one volume), then
// added to dw with float atomics.  Each A fragment (dy) serves 9 taps and
// each B fragment (x) CO_T tiles.  lane l supplies A[l%16][8(l/16)..+7] =
// dy[voxel 8(l/16)+j][co l%16] and B[8(l/16)..+7][l%16] = x[that voxel
// shifted by the tap][ci l%16], and receives D[4(l/16)+i][l%16] = dw[co
// 4(l/16)+i][ci l%16].  Positions off the grid read as zero (the convolution's
// zero padding; dy is zero beyond the grid).
constexpr int kWgTX = 32, kWgTY = 4;
constexpr int kWgXR = 48;                                   // x row length (elements)
constexpr int kWgXC = (kWgTY + 2) * kWgXR + 8;              // x channel pitch
constexpr int kWgXS = 16 * kWgXC;                           // x plane (slot) size
constexpr int kWgDC = kWgTY * kWgTX + 8;                    // dy channel pitch

__host__ __device__ constexpr size_t wg_lds_bytes(int cot)
{
    return ((size_t)3 * kWgXS + (size_t)16 * cot * kWgDC) * sizeof(uint16_t);
}

template <int CO_T>
__global__ __launch_bounds__(192) void k_conv_wgrad_mm(const uint16_t *__restrict__ x, Shape s, int Ci,
                                                       const uint16_t *__restrict__ dy, int Co, int zc,
                                                       float *__restrict__ dw)
{
    extern __shared__ __align__(16) uint16_t wg_lds[];
    uint16_t *xs = wg_lds;                 // [3][16][kWgTY+2][kWgXR] (+ pads)
    uint16_t *ds = wg_lds + 3 * kWgXS;     // [16*CO_T][kWgTY][32] (+ pads)
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int l16 = lane & 15, kg = lane >> 4;
    const int x0 = blockIdx.x * kWgTX, y0 = blockIdx.y * kWgTY;
    const int nzc = (s.D + zc - 1) / zc;
    const int ncib = Ci / 16, ncob = Co / (16 * CO_T);
    int q = (int)blockIdx.z;
    const int cib = q % ncib; q /= ncib;
    const int cob = q % ncob; q /= ncob;
    const int n = q / nzc, zb = (q % nzc) * zc, ze = min(s.D, zb + zc);
    const int ci0 = cib * 16, co0 = cob * 16 * CO_T;

    // x plane zz (channels ci0..ci0+15) into ring slot `slot`, zero off the grid
    auto load_x = [&](int zz, int slot) {
        uint16_t *dst = xs + slot * kWgXS;
        const bool zin = (unsigned)zz < (unsigned)s.D;
        for (int it = tid; it < (kWgTY + 2) * (kWgTX + 2) * 2; it += 192) {
            const int g = it & 1, p = it >> 1, ly = p / (kWgTX + 2), lx = p - ly * (kWgTX + 2);
            const int yy = y0 - 1 + ly, xx = x0 - 1 + lx;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (zin && (unsigned)yy < (unsigned)s.H && (unsigned)xx < (unsigned)s.W)
                v = *(const uint4 *)(x + ((((int64_t)n * s.D + zz) * s.H + yy) * s.W + xx) * Ci + ci0 + 8 * g);
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
            uint16_t *o = dst + (8 * g) * kWgXC + ly * kWgXR + lx + 7;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                o[(2 * j) * kWgXC] = (uint16_t)(w4[j] & 0xFFFFu);
                o[(2 * j + 1) * kWgXC] = (uint16_t)(w4[j] >> 16);
            }
        }
    };
    // dy plane zz (channels co0 .. co0 + 16*CO_T - 1), zero off the grid
    auto load_dy = [&](int zz) {
        constexpr int G = 2 * CO_T;   // 8-channel groups
        for (int it = tid; it < kWgTY * kWgTX * G; it += 192) {
            const int g = it % G, p = it / G, ly = p / kWgTX, lx = p - ly * kWgTX;
            const int yy = y0 + ly, xx = x0 + lx;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if ((unsigned)yy < (unsigned)s.H && (unsigned)xx < (unsigned)s.W)
                v = *(const uint4 *)(dy + ((((int64_t)n * s.D + zz) * s.H + yy) * s.W + xx) * Co + co0 + 8 * g);
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
            uint16_t *o = ds + (8 * g) * kWgDC + ly * kWgTX + lx;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                o[(2 * j) * kWgDC] = (uint16_t)(w4[j] & 0xFFFFu);
                o[(2 * j + 1) * kWgDC] = (uint16_t)(w4[j] >> 16);
            }
        }
    };

    float4_t acc[9][CO_T];
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int c = 0; c < CO_T; ++c) acc[t][c] = float4_t{0.f, 0.f, 0.f, 0.f};

    load_x(zb - 1, (zb + 2) % 3);
    load_x(zb, zb % 3);
    for (int z = zb; z < ze; ++z) {
        load_x(z + 1, (z + 1) % 3);
        load_dy(z);
        __syncthreads();
        // this wave's x plane z + dz, dz = wv - 1
        const uint16_t *xp = xs + ((z + wv + 2) % 3) * kWgXS + l16 * kWgXC;
        const uint16_t *dp = ds + l16 * kWgDC + 8 * kg;
#pragma unroll
        for (int r = 0; r < kWgTY; ++r) {
            short8_t a[CO_T];
#pragma unroll
            for (int c = 0; c < CO_T; ++c) a[c] = *(const short8_t *)(dp + 16 * c * kWgDC + r * kWgTX);
#pragma unroll
            for (int ty = 0; ty < 3; ++ty) {
                const uint16_t *row = xp + (r + ty) * kWgXR + 8 * kg + 8;   // element 8g + 8: x0 + 8g
                const uint4 mid = *(const uint4 *)row;
                const uint32_t lo = *(const uint32_t *)(row - 2);          // elements 8g + 6, 8g + 7
                const uint32_t hi = *(const uint32_t *)(row + 8);          // elements 8g + 16, 8g + 17
                short8_t b[3];
                // dx = -1: elements 8g+7 .. 8g+14
                b[0] = __builtin_bit_cast(short8_t, make_uint4(__builtin_amdgcn_alignbit(mid.x, lo, 16),
                                                               __builtin_amdgcn_alignbit(mid.y, mid.x, 16),
                                                               __builtin_amdgcn_alignbit(mid.z, mid.y, 16),
                                                               __builtin_amdgcn_alignbit(mid.w, mid.z, 16)));
                b[1] = __builtin_bit_cast(short8_t, mid);
                // dx = +1: elements 8g+9 .. 8g+16
                b[2] = __builtin_bit_cast(short8_t, make_uint4(__builtin_amdgcn_alignbit(mid.y, mid.x, 16),
                                                               __builtin_amdgcn_alignbit(mid.z, mid.y, 16),
                                                               __builtin_amdgcn_alignbit(mid.w, mid.z, 16),
                                                               __builtin_amdgcn_alignbit(hi, mid.w, 16)));
#pragma unroll
                for (int tx = 0; tx < 3; ++tx)
#pragma unroll
                    for (int c = 0; c < CO_T; ++c) acc[ty * 3 + tx][c] = mfma32(a[c], b[tx], acc[ty * 3 + tx][c]);
            }
        }
        __syncthreads();   // the ring slot of plane z - 1 and the dy plane are reused
    }
    // D[4kg + i][l16]: dw[co0 + 16c + 4kg + i][tap][ci0 + l16], tap = (dz+1)*9 + (dy+1)*3 + (dx+1)
#pragma unroll
    for (int t = 0; t < 9; ++t)
#pragma unroll
        for (int c = 0; c < CO_T; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                atomicAdd(&dw[((int64_t)(co0 + 16 * c + 4 * kg + i) * 27 + wv * 9 + t) * Ci + ci0 + l16], acc[t][c][i]);
}

// Bias gradient db[co] += sum_v dy[v][co] (channels-last dy, Co <= 256 and a
// multiple of 16): each block sums a stride of voxels per channel in
// registers, then in LDS, and adds once per channel.
__global__ __launch_bounds__(256) void k_bias_grad(const uint16_t *__restrict__ dy, int64_t nvox, int Co,
                                                   float *__restrict__ db)
{
    __shared__ float part[256];
    const int G = Co / 8;                       // 16-byte groups per voxel
    const int per = 256 / G;                    // voxels a block reads at once
    const int g = threadIdx.x % G, vo = threadIdx.x / G;
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (vo < per)
        for (int64_t v = (int64_t)blockIdx.x * per + vo; v < nvox; v += (int64_t)gridDim.x * per) {
            const uint4 q = *(const uint4 *)(dy + v * Co + 8 * g);
            const uint32_t w4[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[2 * j] += bf2f((uint16_t)(w4[j] & 0xFFFFu));
                acc[2 * j + 1] += bf2f((uint16_t)(w4[j] >> 16));
            }
        }
    for (int c = threadIdx.x; c < Co; c += 256) part[c] = 0.0f;
    __syncthreads();
    if (vo < per)
#pragma unroll
        for (int j = 0; j < 8; ++j) atomicAdd(&part[8 * g + j], acc[j]);
    __syncthreads();
    for (int c = threadIdx.x; c < Co; c += 256) atomicAdd(&db[c], part[c]);
}

// g = (a > 0) ? g : 0  (backward of a ReLU whose output is a).
__global__ void k_relu_mask(uint16_t *__restrict__ g, const uint16_t *__restrict__ a, int64_t n)
{
    for (int64_t i = bmx_gtid(); i < n; i += bmx_gstride())
        if (!(bf2f(a[i]) > 0.0f)) g[i] = 0;
}

// [Co][27][Ci] -> [Ci][27][Co]
// wt[ci][26 - tap][co] = w[co][tap][ci]: the input gradient dx[v][ci] =
// sum_tap sum_co dy[v - off(tap)][co] w[co][tap][ci] is the forward
// convolution of dy with wt (off(26 - tap) = -off(tap)).
__global__ void k_transpose_w(const uint16_t *__restrict__ w, int Co, int Ci, uint16_t *__restrict__ wt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)Co * 27 * Ci) return;
    const int ci = (int)(i % Ci);
    const int tap = (int)((i / Ci) % 27);
    const int co = (int)(i / ((int64_t)Ci * 27));
    wt[((int64_t)ci * 27 + (26 - tap)) * Co + co] = w[i];
}

// 2x2x2 average pooling (channels-last), and its backward.
// One work-item per output voxel and 8-channel group (16-byte loads and stores;
// C is a multiple of 16); each channel sums its 8 inputs in (z, y, x) order.
__global__ void k_pool2(const uint16_t *__restrict__ x, Shape s, int C, uint16_t *__restrict__ y)
{
    const int Do = s.D / 2, Ho = s.H / 2, Wo = s.W / 2, G = C / 8;
    const int64_t n_out = (int64_t)s.N * Do * Ho * Wo * G;
    for (int64_t i = bmx_gtid(); i < n_out; i += bmx_gstride()) {
        const int g = (int)(i % G);
        int64_t t = i / G;
        const int xo = (int)(t % Wo); t /= Wo;
        const int yo = (int)(t % Ho); t /= Ho;
        const int zo = (int)(t % Do);
        const int n = (int)(t / Do);
        float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < 8; ++k) {
            const int z = 2 * zo + (k >> 2), yy = 2 * yo + ((k >> 1) & 1), xx = 2 * xo + (k & 1);
            const uint4 v = *(const uint4 *)(x + ((((int64_t)n * s.D + z) * s.H + yy) * s.W + xx) * C + 8 * g);
            const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[2 * j] += bf2f((uint16_t)(w4[j] & 0xFFFFu));
                acc[2 * j + 1] += bf2f((uint16_t)(w4[j] >> 16));
            }
        }
        uint4 o;
        o.x = (uint32_t)f2bf(acc[0] * 0.125f) | ((uint32_t)f2bf(acc[1] * 0.125f) << 16);
        o.y = (uint32_t)f2bf(acc[2] * 0.125f) | ((uint32_t)f2bf(acc[3] * 0.125f) << 16);
        o.z = (uint32_t)f2bf(acc[4] * 0.125f) | ((uint32_t)f2bf(acc[5] * 0.125f) << 16);
        o.w = (uint32_t)f2bf(acc[6] * 0.125f) | ((uint32_t)f2bf(acc[7] * 0.125f) << 16);
        *(uint4 *)(y + i * 8) = o;   // element i * 8 = output voxel * C + 8 g
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
// prob = softmax(logits).  With a target (soft labels, float [v][L], or label
// indices) it also writes dx (bf16, scale/nvox * W^T (prob - target)) and
// accumulates dW, db and the summed cross-entropy into *loss: each lane sums
// its share of dW/db over the voxels its wave visits (staged in LDS), then per
// block in LDS, then one float atomic per entry and block.
__global__ __launch_bounds__(256) void k_head(const uint16_t *__restrict__ x, int64_t nvox, int L,
                                              const float *__restrict__ W, const float *__restrict__ b,
                                              float *__restrict__ prob, const float *__restrict__ target,
                                              const uint8_t *__restrict__ tidx,
                                              const uint8_t *__restrict__ vmask, float scale,
                                              uint16_t *__restrict__ dx, float *__restrict__ dW,
                                              float *__restrict__ db, float *__restrict__ loss)
{
    constexpr int C = 16, LM = 32;
    // padded rows of the wave's staged gradients (odd: lanes writing their rows hit distinct banks) and
    // inputs (16-byte aligned: a lane reads its 8 channels with two 16-byte reads)
    constexpr int GP = LM + 1, XP = C + 4;
    // per wave: the logit gradients [64][GP] and inputs [64][XP] of its 64 voxels
    __shared__ float s_g[4][64 * GP], s_x[4][64 * XP];
    __shared__ float s_dw[LM * C], s_db[LM], s_loss;
    // the head's parameters, rows L .. 31 zero, read by every lane at the same address
    __shared__ __align__(16) float s_w[LM * C];
    __shared__ float s_b[LM];
    const bool train = target || tidx;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < LM * C; i += blockDim.x) {
        s_w[i] = i < L * C ? W[i] : 0.0f;
        s_dw[i] = 0.0f;
    }
    if (threadIdx.x < LM) {
        s_b[threadIdx.x] = (int)threadIdx.x < L ? b[threadIdx.x] : 0.0f;
        s_db[threadIdx.x] = 0.0f;
    }
    if (threadIdx.x == 0) s_loss = 0.0f;
    __syncthreads();
    // this lane's share of the weight gradient: label lane % 32, channels
    // 8 (lane / 32) .. +7, summed over every voxel the wave visits
    const int ml = lane & 31, mh = lane >> 5;
    float wacc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float bacc = 0.0f, ls = 0.0f;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // the loop bound is uniform per block
    for (int64_t v0 = (int64_t)blockIdx.x * blockDim.x; v0 < nvox; v0 += stride) {
        const int64_t v = v0 + threadIdx.x;
        const bool act = v < nvox;
        float xin[C];
        if (act) {
            unpack8(*(const uint4 *)(x + v * C), xin);
            unpack8(*(const uint4 *)(x + v * C + 8), xin + 8);
        } else {
#pragma unroll
            for (int c = 0; c < C; ++c) xin[c] = 0.0f;
        }
        float lg[LM];
        float mx = -1e30f;
#pragma unroll
        for (int l = 0; l < LM; ++l) {
            float s = s_b[l];
#pragma unroll
            for (int c = 0; c < C; ++c) s = __builtin_fmaf(s_w[l * C + c], xin[c], s);
            lg[l] = s;
            if (l < L) mx = fmaxf(mx, s);
        }
        float den = 0.0f;
#pragma unroll
        for (int l = 0; l < LM; ++l) {
            lg[l] = l < L ? __expf(lg[l] - mx) : 0.0f;
            den += lg[l];
        }
        const float rden = 1.0f / den;
        const bool wgt = act && (!vmask || vmask[v]);
        const int ti = (train && act && !target) ? (int)tidx[v] : -1;
#pragma unroll
        for (int l = 0; l < LM; ++l) {
            float gl = 0.0f;
            if (l < L) {
                const float p = lg[l] * rden;
                if (act && prob) prob[v * L + l] = p;
                if (train) {
                    const float t = !act ? 0.0f : (target ? target[v * L + l] : (ti == l ? 1.0f : 0.0f));
                    gl = wgt ? scale * (p - t) : 0.0f;
                    if (wgt && t > 0.0f) ls -= t * __logf(fmaxf(p, 1e-30f));
                }
            }
            lg[l] = gl;   // from here on: the logit gradient
        }
        if (!train) continue;
        if (act) {
            float d[C];
#pragma unroll
            for (int c = 0; c < C; ++c) {
                float s = 0.0f;
#pragma unroll
                for (int l = 0; l < LM; ++l) s = __builtin_fmaf(s_w[l * C + c], lg[l], s);
                d[c] = s;
            }
            *(uint4 *)(dx + v * C) = pack8(d);
            *(uint4 *)(dx + v * C + 8) = pack8(d + 8);
        }
        // dW[l][c] += sum over the wave's voxels of g[v][l] x[v][c], db[l] += g[v][l]:
        // each lane stages its voxel, then sums its entries over the 64 staged voxels
        float *sg = s_g[wv], *sx = s_x[wv];
#pragma unroll
        for (int l = 0; l < LM; ++l) sg[lane * GP + l] = lg[l];
#pragma unroll
        for (int c = 0; c < C; ++c) sx[lane * XP + c] = xin[c];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (int u = 0; u < 64; ++u) {
            const float g = sg[u * GP + ml];
#pragma unroll
            for (int j = 0; j < 8; ++j) wacc[j] = __builtin_fmaf(g, sx[u * XP + 8 * mh + j], wacc[j]);
            bacc += g;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    if (!train) return;
    for (int off = 32; off > 0; off >>= 1) ls += __shfl_xor(ls, off, 64);
    if (ml < L) {
#pragma unroll
        for (int j = 0; j < 8; ++j) atomicAdd(&s_dw[ml * C + 8 * mh + j], wacc[j]);
        if (mh == 0) atomicAdd(&s_db[ml], bacc);
    }
    if (lane == 0) atomicAdd(&s_loss, ls);
    __syncthreads();
    for (int i = threadIdx.x; i < L * C; i += blockDim.x) atomicAdd(&dW[i], s_dw[i]);
    if (threadIdx.x < L) atomicAdd(&db[threadIdx.x], s_db[threadIdx.x]);
    if (threadIdx.x == 0) atomicAdd(loss, s_loss);
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

// The same operations on 8 channels (16 bytes) per work-item, for tensors
// whose channel counts are multiples of 8 (the network's are multiples of 16)
// and whose pointers are 16-byte aligned; element for element the results are
// those of the kernels above.
__global__ void k_up2_v8(const uint16_t *__restrict__ x, Shape s, int C, uint16_t *__restrict__ y)
{
    const int D2 = s.D * 2, H2 = s.H * 2, W2 = s.W * 2, G = C / 8;
    const int64_t n_out = (int64_t)s.N * D2 * H2 * W2 * G;
    for (int64_t i = bmx_gtid(); i < n_out; i += bmx_gstride()) {
        const int g = (int)(i % G);
        int64_t t = i / G;
        const int xx = (int)(t % W2); t /= W2;
        const int yy = (int)(t % H2); t /= H2;
        const int z = (int)(t % D2);
        const int n = (int)(t / D2);
        *(uint4 *)(y + i * 8) = *(const uint4 *)(x + ((((int64_t)n * s.D + z / 2) * s.H + yy / 2) * s.W + xx / 2) * C + 8 * g);
    }
}

__global__ void k_up2_bwd_v8(const uint16_t *__restrict__ dy, Shape s, int C, uint16_t *__restrict__ dx)
{
    const int G = C / 8;
    const int64_t n_small = (int64_t)s.N * s.D * s.H * s.W * G;
    for (int64_t i = bmx_gtid(); i < n_small; i += bmx_gstride()) {
        const int g = (int)(i % G);
        int64_t t = i / G;
        const int xo = (int)(t % s.W); t /= s.W;
        const int yo = (int)(t % s.H); t /= s.H;
        const int zo = (int)(t % s.D);
        const int n = (int)(t / s.D);
        float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < 8; ++k) {
            const int z = 2 * zo + (k >> 2), yy = 2 * yo + ((k >> 1) & 1), xx = 2 * xo + (k & 1);
            float f[8];
            unpack8(*(const uint4 *)(dy + ((((int64_t)n * 2 * s.D + z) * 2 * s.H + yy) * 2 * s.W + xx) * C + 8 * g), f);
#pragma unroll
            for (int j = 0; j < 8; ++j) acc[j] += f[j];
        }
        *(uint4 *)(dx + i * 8) = pack8(acc);
    }
}

__global__ void k_pool2_bwd_v8(const uint16_t *__restrict__ dy, Shape s, int C, uint16_t *__restrict__ dx)
{
    const int G = C / 8;
    const int64_t n_in = (int64_t)s.N * s.D * s.H * s.W * G;
    for (int64_t i = bmx_gtid(); i < n_in; i += bmx_gstride()) {
        const int g = (int)(i % G);
        int64_t t = i / G;
        const int xx = (int)(t % s.W); t /= s.W;
        const int yy = (int)(t % s.H); t /= s.H;
        const int z = (int)(t % s.D);
        const int n = (int)(t / s.D);
        float f[8];
        unpack8(*(const uint4 *)(dy + ((((int64_t)n * (s.D / 2) + z / 2) * (s.H / 2) + yy / 2) * (s.W / 2) + xx / 2) * C +
                                 8 * g), f);
#pragma unroll
        for (int j = 0; j < 8; ++j) f[j] *= 0.125f;
        *(uint4 *)(dx + i * 8) = pack8(f);
    }
}

__global__ void k_concat_v8(const uint16_t *__restrict__ a, int Ca, const uint16_t *__restrict__ b, int Cb,
                            int64_t nvox, uint16_t *__restrict__ y)
{
    const int G = (Ca + Cb) / 8, Ga = Ca / 8;
    for (int64_t i = bmx_gtid(); i < nvox * G; i += bmx_gstride()) {
        const int64_t v = i / G;
        const int g = (int)(i - v * G);
        *(uint4 *)(y + i * 8) = g < Ga ? *(const uint4 *)(a + v * Ca + 8 * g) : *(const uint4 *)(b + v * Cb + 8 * (g - Ga));
    }
}

__global__ void k_split_v8(const uint16_t *__restrict__ y, int64_t nvox, uint16_t *__restrict__ a, int Ca,
                           uint16_t *__restrict__ b, int Cb, int accumulate_a)
{
    const int G = (Ca + Cb) / 8, Ga = Ca / 8;
    for (int64_t i = bmx_gtid(); i < nvox * G; i += bmx_gstride()) {
        const int64_t v = i / G;
        const int g = (int)(i - v * G);
        const uint4 q = *(const uint4 *)(y + i * 8);
        if (g < Ga) {
            uint16_t *pa = a + v * Ca + 8 * g;
            if (accumulate_a) {
                float f[8], e[8];
                unpack8(q, f);
                unpack8(*(const uint4 *)pa, e);
#pragma unroll
                for (int j = 0; j < 8; ++j) f[j] += e[j];
                *(uint4 *)pa = pack8(f);
            } else {
                *(uint4 *)pa = q;
            }
        } else {
            *(uint4 *)(b + v * Cb + 8 * (g - Ga)) = q;
        }
    }
}

__global__ void k_relu_mask_v8(uint16_t *__restrict__ g, const uint16_t *__restrict__ a, int64_t n8)
{
    for (int64_t i = bmx_gtid(); i < n8; i += bmx_gstride()) {
        float f[8];
        unpack8(*(const uint4 *)(a + i * 8), f);
        uint4 q = *(const uint4 *)(g + i * 8);
        uint32_t w4[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (!(f[2 * j] > 0.0f)) w4[j] &= 0xFFFF0000u;
            if (!(f[2 * j + 1] > 0.0f)) w4[j] &= 0x0000FFFFu;
        }
        *(uint4 *)(g + i * 8) = make_uint4(w4[0], w4[1], w4[2], w4[3]);
    }
}

__global__ void k_add_bf16_v8(uint16_t *__restrict__ y, const uint16_t *__restrict__ x, int64_t n8)
{
    for (int64_t i = bmx_gtid(); i < n8; i += bmx_gstride()) {
        float f[8], e[8];
        unpack8(*(const uint4 *)(y + i * 8), f);
        unpack8(*(const uint4 *)(x + i * 8), e);
#pragma unroll
        for (int j = 0; j < 8; ++j) f[j] += e[j];
        *(uint4 *)(y + i * 8) = pack8(f);
    }
}

static bool al16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

__global__ void k_f2bf(const float *__restrict__ a, uint16_t *__restrict__ b, int64_t n)
{
    for (int64_t i = bmx_gtid(); i < n; i += bmx_gstride()) b[i] = f2bf(a[i]);
}

template <int CH, bool CAT, bool MASK, int SPLIT>
int launch_mm(bmx_ctx *ctx, const uint16_t *x, Shape s, int Ci, const uint16_t *w, const float *b, uint16_t *y,
              int Co, int relu, const uint16_t *xb, int Ca, const uint16_t *xact, int Cf, int coff, float *part)
{
    const size_t lds = mm_lds_bytes(CH);
    const void *kern = (const void *)k_conv_mm<CH, CAT, MASK, SPLIT>;
    // an error here is left to the launch itself to report
    if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        (void)hipGetLastError();
    const int nzc = (s.D + kMmZC - 1) / kMmZC;
    const int64_t gz = (int64_t)s.N * nzc * (Co / 16);
    BMX_CHECK(gz < 65536, "conv3d: too many z chunks for one dispatch");
    BMX_CHECK((int64_t)s.H * s.W * (Cf > Co ? Cf : Co) < ((int64_t)1 << 31) - 2, "conv3d: plane too large");
    dim3 grid((unsigned)((s.W + kMmTX - 1) / kMmTX), (unsigned)((s.H + kMmTY - 1) / kMmTY), (unsigned)gz);
    k_conv_mm<CH, CAT, MASK, SPLIT><<<grid, 256, lds, ctx->stream>>>(x, s, Ci, w, b, y, Co, relu, xb, Ca, xact, Cf,
                                                                     coff, part);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

template <bool CAT, bool MASK, int SPLIT>
int launch_mm_ch(bmx_ctx *ctx, const uint16_t *x, Shape s, int Ci, const uint16_t *w, const float *b, uint16_t *y,
                 int Co, int relu, const uint16_t *xb, int Ca, const uint16_t *xact, int Cf, int coff, float *part)
{
    if (Ci <= 32) return launch_mm<1, CAT, MASK, SPLIT>(ctx, x, s, Ci, w, b, y, Co, relu, xb, Ca, xact, Cf, coff, part);
    return launch_mm<2, CAT, MASK, SPLIT>(ctx, x, s, Ci, w, b, y, Co, relu, xb, Ca, xact, Cf, coff, part);
}

// The MFMA (16x16x32) forward: up to 64 input channels in one launch (CH =
// ceil(Ci / 32) <= 2 chunks, the weights of a block's 16 output channels in
// registers), wider inputs in launches over 64-channel parts with float32
// partial sums in a context buffer.
template <bool CAT, bool MASK>
int conv_mm(bmx_ctx *ctx, const uint16_t *x, Shape s, int Ci, const uint16_t *w, const float *b, uint16_t *y, int Co,
            int relu, const uint16_t *xb, int Ca, const uint16_t *xact)
{
    if (Ci <= 64)
        return launch_mm_ch<CAT, MASK, 0>(ctx, x, s, Ci, w, b, y, Co, relu, xb, Ca, xact, Ci, 0, nullptr);
    const size_t bytes = (size_t)s.N * s.D * s.H * s.W * Co * sizeof(float);
    int rc = grow(ctx, &ctx->d_netpart, &ctx->netpart_cap, bytes);
    if (rc) return rc;
    float *part = (float *)ctx->d_netpart;
    for (int c0 = 0; c0 < Ci && rc == BMX_OK; c0 += 64) {
        const int cw = Ci - c0 < 64 ? Ci - c0 : 64;
        if (c0 == 0)
            rc = launch_mm_ch<CAT, MASK, 1>(ctx, x, s, cw, w, b, y, Co, relu, xb, Ca, xact, Ci, c0, part);
        else if (c0 + cw < Ci)
            rc = launch_mm_ch<CAT, MASK, 2>(ctx, x, s, cw, w, b, y, Co, relu, xb, Ca, xact, Ci, c0, part);
        else
            rc = launch_mm_ch<CAT, MASK, 3>(ctx, x, s, cw, w, b, y, Co, relu, xb, Ca, xact, Ci, c0, part);
    }
    return rc;
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
    return conv_mm<false, false>(ctx, x, s, Ci, w, b, y, Co, relu, nullptr, Ci, nullptr);
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
    return conv_mm<true, false>(ctx, a, s, Ci, w, bias, y, Co, relu, b, Ca, nullptr);
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
    return conv_mm<false, true>(ctx, dy, s, Co, wt_scratch, nullptr, dx, Ci, 0, nullptr, Co, xact);
}

int bmx_conv3d_wgrad(bmx_ctx *ctx, const uint16_t *x, int N, int D, int H, int W, int Ci, const uint16_t *dy,
                     int Co, float *dw, float *db)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    rc = check_conv_shape(N, D, H, W, Ci, Co);
    if (rc) return rc;
    BMX_CHECK(x && dy && dw, "bmx_conv3d_wgrad: null pointer");
    BMX_CHECK(Co <= 256, "bmx_conv3d_wgrad: at most 256 output channels");
    Shape s{N, D, H, W};
    const int64_t nvox = (int64_t)N * D * H * W;
    const int cot = Co % 64 == 0 ? 4 : (Co % 32 == 0 ? 2 : 1);
    // z chunk per block: about 1024 blocks in all, each adding its sums once
    const int64_t base = (int64_t)((W + kWgTX - 1) / kWgTX) * ((H + kWgTY - 1) / kWgTY) * N * (Co / (16 * cot)) *
                         (Ci / 16);
    const int64_t want = std::min<int64_t>(std::max<int64_t>((1024 + base - 1) / base, 1), D);
    const int zc = (int)((D + want - 1) / want);
    const int64_t nz = (int64_t)N * ((D + zc - 1) / zc) * (Co / (16 * cot)) * (Ci / 16);
    BMX_CHECK(nz < ((int64_t)1 << 31), "bmx_conv3d_wgrad: grid too large");
    dim3 grid((unsigned)((W + kWgTX - 1) / kWgTX), (unsigned)((H + kWgTY - 1) / kWgTY), (unsigned)nz);
    if (cot == 4)
        k_conv_wgrad_mm<4><<<grid, 192, wg_lds_bytes(4), ctx->stream>>>(x, s, Ci, dy, Co, zc, dw);
    else if (cot == 2)
        k_conv_wgrad_mm<2><<<grid, 192, wg_lds_bytes(2), ctx->stream>>>(x, s, Ci, dy, Co, zc, dw);
    else
        k_conv_wgrad_mm<1><<<grid, 192, wg_lds_bytes(1), ctx->stream>>>(x, s, Ci, dy, Co, zc, dw);
    BMX_LAUNCH_CHECK();
    if (db) {
        const int64_t per = 256 / (Co / 8);
        const int64_t blocks = std::min<int64_t>((nvox + per - 1) / per, 2048);
        k_bias_grad<<<(unsigned)blocks, 256, 0, ctx->stream>>>(dy, nvox, Co, db);
        BMX_LAUNCH_CHECK();
    }
    return BMX_OK;
}

int bmx_pool2(bmx_ctx *ctx, const uint16_t *x, int N, int D, int H, int W, int C, uint16_t *y)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(x && y && D % 2 == 0 && H % 2 == 0 && W % 2 == 0 && C >= 16 && C % 16 == 0,
              "bmx_pool2: bad arguments");
    const int64_t n = (int64_t)N * (D / 2) * (H / 2) * (W / 2) * (C / 8);
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
    if (C % 8 == 0 && al16(dy) && al16(dx))
        k_pool2_bwd_v8<<<bmx_stride_blocks(n / 8, 256), 256, 0, ctx->stream>>>(dy, Shape{N, D, H, W}, C, dx);
    else
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
    if (C % 8 == 0 && al16(x) && al16(y))
        k_up2_v8<<<bmx_stride_blocks(n / 8, 256), 256, 0, ctx->stream>>>(x, Shape{N, D, H, W}, C, y);
    else
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
    if (C % 8 == 0 && al16(dy) && al16(dx))
        k_up2_bwd_v8<<<bmx_stride_blocks(n / 8, 256), 256, 0, ctx->stream>>>(dy, Shape{N, D, H, W}, C, dx);
    else
        k_up2_bwd<<<bmx_stride_blocks(n, 256), 256, 0, ctx->stream>>>(dy, Shape{N, D, H, W}, C, dx);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_concat(bmx_ctx *ctx, const uint16_t *a, int Ca, const uint16_t *b, int Cb, int64_t nvox, uint16_t *y)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(a && b && y, "bmx_concat: null pointer");
    if (Ca % 8 == 0 && Cb % 8 == 0 && al16(a) && al16(b) && al16(y))
        k_concat_v8<<<bmx_stride_blocks(nvox * (Ca + Cb) / 8, 256), 256, 0, ctx->stream>>>(a, Ca, b, Cb, nvox, y);
    else
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
    if (Ca % 8 == 0 && Cb % 8 == 0 && al16(a) && al16(b) && al16(y))
        k_split_v8<<<bmx_stride_blocks(nvox * (Ca + Cb) / 8, 256), 256, 0, ctx->stream>>>(y, nvox, a, Ca, b, Cb,
                                                                                          accumulate_a);
    else
        k_split<<<bmx_stride_blocks(nvox * (Ca + Cb), 256), 256, 0, ctx->stream>>>(y, nvox, a, Ca, b, Cb,
                                                                                   accumulate_a);
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
    // training: a bounded grid, each block striding over voxels, so that the
    // parameter gradients take few global atomics
    const bool train = target || target_idx;
    const unsigned blocks = train ? (unsigned)std::min<int64_t>(bmx_blocks(nvox, 256), 2048) : bmx_blocks(nvox, 256);
    k_head<<<blocks > 0 ? blocks : 1, 256, 0, ctx->stream>>>(x, nvox, L, W, b, prob, target, target_idx, vmask,
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
    if (n % 8 == 0 && al16(g) && al16(a))
        k_relu_mask_v8<<<bmx_stride_blocks(n / 8, 256), 256, 0, ctx->stream>>>(g, a, n / 8);
    else
        k_relu_mask<<<bmx_stride_blocks(n, 256), 256, 0, ctx->stream>>>(g, a, n);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

int bmx_add_bf16(bmx_ctx *ctx, uint16_t *y, const uint16_t *x, int64_t n)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(y && x, "bmx_add_bf16: null pointer");
    if (n % 8 == 0 && al16(y) && al16(x))
        k_add_bf16_v8<<<bmx_stride_blocks(n / 8, 256), 256, 0, ctx->stream>>>(y, x, n / 8);
    else
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
