// bmx_smooth.hip -- curvature smoothing of the per-label hit volumes.
//
// One round of the reference is two launches (opencl_kernels.py:136-224):
// CURVATURE_SRC writes kappa(phi) for interior voxels, then UPDATE_SRC adds
// 0.01*kappa to phi.  kappa at round t only reads phi_t, so one fused stencil
// pass phi_{t+1} = phi_t + 0.01*kappa(phi_t) into a second buffer gives the
// same floats with half the traffic.  Interior expressions are written in
// the reference's operation order (the build disables FP contraction).
//
// Tiling: a 64x16 (x,y) tile per block (each thread owns 4 rows of a column)
// marches along a chunk of 32 z planes; the planes z-1, z, z+1 of the tile plus
// a 1-voxel halo rotate through three LDS slots and the next plane is fetched
// into registers during the computation, so each input voxel is read from
// memory about 1.15 times per pass.  k_smooth_pair does two rounds per pass
// (a 2-voxel halo, the middle round kept in LDS); an odd count ends with one
// k_smooth_step pass.
#include "bmx_internal.h"

#include <algorithm>
#include <climits>
#include <type_traits>
#include <vector>

namespace {

constexpr int TX = 64;            // x extent of a tile (one wave per row group)
constexpr int TYT = 4;            // thread rows per block
constexpr int RY = 4;             // output rows per thread (register blocking in y)
constexpr int TY = TYT * RY;      // y extent of a tile
constexpr int kSmoothThreads = TX * TYT;
constexpr int kZChunk = 32;
constexpr int LW = TX + 2, LH = TY + 2;   // LDS plane with a 1-voxel halo

template <typename T> __device__ __forceinline__ float ldf(const T *p, int64_t i);
template <> __device__ __forceinline__ float ldf<float>(const float *p, int64_t i) { return p[i]; }
template <> __device__ __forceinline__ float ldf<uint32_t>(const uint32_t *p, int64_t i)
{
    return (float)p[i];
}

// Box [z0,z1) x [y0,y1) x [x0,x1) of the volume that a launch updates; voxels
// outside it read as 0 (smooth_argmax restricts each label to the box of its
// non-zero hits grown by the iteration count, outside which phi stays 0).
struct Box {
    int z0, z1, y0, y1, x0, x1;
};

// Per-thread share of one tile plane plus halo: RY interior rows at column tx,
// and up to two halo values (rows 0 / LH-1 for thread rows 0 / 1, the left and
// right halo columns for thread row 2).
struct PlaneRegs {
    float v[RY];
    float h0, h1;
};

// Value of voxel (z, y, x) as the stencil sees it: clamped to the volume (the
// clamped values only feed boundary voxels, which are copied), and 0 outside
// the box.
template <typename Tin>
__device__ __forceinline__ float box_ld(const Tin *__restrict__ in, int z, int y, int x, int Y, int X,
                                        const Box &bx, int Z)
{
    z = min(max(z, 0), Z - 1);
    y = min(max(y, 0), Y - 1);
    x = min(max(x, 0), X - 1);
    const bool inb = z >= bx.z0 && z < bx.z1 && y >= bx.y0 && y < bx.y1 && x >= bx.x0 && x < bx.x1;
    return inb ? ldf(in, ((int64_t)z * Y + y) * X + x) : 0.0f;
}

template <typename Tin>
__device__ __forceinline__ void fetch_plane(const Tin *__restrict__ in, int zz, int Z, int Y, int X, int x0,
                                            int y0, int tx, int ty, const Box &bx, PlaneRegs &r)
{
#pragma unroll
    for (int k = 0; k < RY; ++k) r.v[k] = box_ld(in, zz, y0 + ty * RY + k, x0 + tx, Y, X, bx, Z);
    r.h0 = 0.0f;
    r.h1 = 0.0f;
    if (ty < 2) {
        // halo rows: ly = 0 (y0-1) for ty 0, ly = LH-1 (y0+TY) for ty 1
        r.h0 = box_ld(in, zz, ty == 0 ? y0 - 1 : y0 + TY, x0 + tx, Y, X, bx, Z);
    } else if (ty == 2 && tx < 2 * LH) {
        // halo columns: lx = 0 (x0-1) and lx = LW-1 (x0+TX), rows ly = 0..LH-1
        r.h0 = box_ld(in, zz, y0 + tx % LH - 1, tx < LH ? x0 - 1 : x0 + TX, Y, X, bx, Z);
    }
}

__device__ __forceinline__ void store_plane(float (*s)[LW], int tx, int ty, const PlaneRegs &r)
{
#pragma unroll
    for (int k = 0; k < RY; ++k) s[1 + ty * RY + k][1 + tx] = r.v[k];
    if (ty < 2) {
        s[ty == 0 ? 0 : LH - 1][1 + tx] = r.h0;
    } else if (ty == 2 && tx < 2 * LH) {
        s[tx % LH][tx < LH ? 0 : LW - 1] = r.h0;
    }
}

// phi + 0.01 kappa(phi) at an interior voxel from the planes z-1 (m), z (c) and
// z+1 (p) held in LDS rows of width W, in the reference's operation order
// (opencl_kernels.py:136-224).
template <int W>
__device__ __forceinline__ float curvature_update(const float (*m)[W], const float (*c)[W], const float (*p)[W],
                                                  int ly, int lx)
{
    const float eps = 1e-10f;
    const float pc = c[ly][lx];
    const float xm = c[ly][lx - 1], xp = c[ly][lx + 1];
    const float ym = c[ly - 1][lx], yp = c[ly + 1][lx];
    const float zm = m[ly][lx], zp = p[ly][lx];
    const float dx = (xm - xp) * 0.5f;
    const float dxx = (xm - 2.0f * pc + xp);
    const float dx2 = dx * dx;
    const float dy = (ym - yp) * 0.5f;
    const float dyy = (ym - 2.0f * pc + yp);
    const float dy2 = dy * dy;
    const float dz = (zm - zp) * 0.5f;
    const float dzz = (zm - 2.0f * pc + zp);
    const float dz2 = dz * dz;
    const float dxy = (c[ly - 1][lx - 1] + c[ly + 1][lx + 1] - c[ly - 1][lx + 1] - c[ly + 1][lx - 1]) * 0.25f;
    const float dxz = (p[ly][lx - 1] + m[ly][lx + 1] - p[ly][lx + 1] - m[ly][lx - 1]) * 0.25f;
    const float dyz = (p[ly - 1][lx] + m[ly + 1][lx] - p[ly + 1][lx] - m[ly - 1][lx]) * 0.25f;
    const float curv = (dxx * (dy2 + dz2)
                        + dyy * (dx2 + dz2)
                        + dzz * (dx2 + dy2)
                        - 2.0f * dx * dy * dxy
                        - 2.0f * dx * dz * dxz
                        - 2.0f * dy * dz * dyz)
                       / (dx2 + dy2 + dz2 + eps);
    return pc + 0.01f * curv;
}

// One round phi_{t+1} = phi_t + 0.01 kappa(phi_t) over a 64x16 (x,y) tile and
// a chunk of z planes.  Planes z-1, z, z+1 live in three LDS slots; the plane
// z+2 is fetched into registers while plane z is computed.  Each thread
// computes RY consecutive rows of one column.
template <typename Tin>
__global__ __launch_bounds__(kSmoothThreads) void k_smooth_step(const Tin *__restrict__ in,
                                                               float *__restrict__ out, int Z, int Y,
                                                               int X, Box bx)
{
    __shared__ float s[3][LH][LW];
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    const int x0 = bx.x0 + blockIdx.x * TX, y0 = bx.y0 + blockIdx.y * TY;
    const int x = x0 + tx;
    const int z_begin = bx.z0 + blockIdx.z * kZChunk;
    const int z_end = min(bx.z1, z_begin + kZChunk);
    const int64_t sy = X, sz = (int64_t)X * Y;

    PlaneRegs r;
    fetch_plane(in, z_begin - 1, Z, Y, X, x0, y0, tx, ty, bx, r);
    store_plane(s[0], tx, ty, r);
    fetch_plane(in, z_begin, Z, Y, X, x0, y0, tx, ty, bx, r);
    store_plane(s[1], tx, ty, r);
    fetch_plane(in, z_begin + 1, Z, Y, X, x0, y0, tx, ty, bx, r);
    int sm = 0, sc = 1, sp = 2;
    const bool inside_x = x > 0 && x < X - 1;
    const int lx = tx + 1;
    for (int z = z_begin; z < z_end; ++z) {
        store_plane(s[sp], tx, ty, r);
        __syncthreads();
        if (z + 1 < z_end) fetch_plane(in, z + 2, Z, Y, X, x0, y0, tx, ty, bx, r);
        const bool inside_z = z > 0 && z < Z - 1;
#pragma unroll
        for (int k = 0; k < RY; ++k) {
            const int ly = 1 + ty * RY + k;
            const int y = y0 + ty * RY + k;
            if (x < bx.x1 && y < bx.y1) {
                const float pc = s[sc][ly][lx];
                const float res = (inside_x && inside_z && y > 0 && y < Y - 1)
                                      ? curvature_update<LW>(s[sm], s[sc], s[sp], ly, lx)
                                      : pc;
                out[(int64_t)z * sz + (int64_t)y * sy + x] = res;
            }
        }
        __syncthreads();
        const int t = sm; sm = sc; sc = sp; sp = t;
    }
}

// Two rounds in one pass: phi_{t+2} from phi_t over a 64x16 (x,y) tile and a
// chunk of z planes.  Round t+1 is computed on the tile grown by one voxel in x
// and y (its planes z-1..z+1 rotate through three LDS slots, like the input's),
// so round t+2 reads only values of its own block; each input voxel is read
// about 1.5 times per two rounds.  Every voxel follows the single-round rules:
// 0 outside the box, copied on the volume boundary, the stencil elsewhere, so
// the results are those of two k_smooth_step passes bit for bit.
// Per-block boxes of non-zero outputs of the pair passes of one smoothing
// (6 ints per block of the launch grid: min z, max z, min y, max y, min x, max
// x; max < 0 = none): out is written by this pass, prev was written by the
// pass before (the input's), prev2 by the pass before that (the output
// buffer's current contents).  prev = null: no block is skipped; prev2 = null:
// the output buffer's contents are unknown; out = null: no boxes are kept.
// src (first pass, prev = null): the box of the input's non-zero voxels in
// the same format; blocks out of its reach are skipped as for prev.
struct PassRecords {
    const int *prev, *prev2;
    int *out;
    const int *src = nullptr;
};

constexpr int PW = TX + 4, PH = TY + 4;        // input plane with a 2-voxel halo
constexpr int kPairLoads = (PW * PH + kSmoothThreads - 1) / kSmoothThreads;
constexpr int kPairZChunk = 32;               // output planes per block

__device__ __forceinline__ int slot3(int z) { return (z + 3) % 3; }

template <typename Tin>
__global__ __launch_bounds__(kSmoothThreads) void k_smooth_pair(const Tin *__restrict__ in,
                                                               float *__restrict__ out, int Z, int Y,
                                                               int X, Box bx, PassRecords rec)
{
    __shared__ float L0[3][PH][PW];   // round t, planes z-1..z+1 (frame: y0-2.., x0-2..)
    __shared__ float L1[3][PH][PW];   // round t+1, rows 1..PH-2, columns 1..PW-2
    __shared__ int red[kSmoothThreads / 64][6];
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const int x0 = bx.x0 + blockIdx.x * TX, y0 = bx.y0 + blockIdx.y * TY;
    const int zb = bx.z0 + blockIdx.z * kPairZChunk;
    const int ze = min(bx.z1, zb + kPairZChunk);
    const int64_t sy = X, sz = (int64_t)X * Y;
    const int blk = ((int)blockIdx.z * (int)gridDim.y + (int)blockIdx.y) * (int)gridDim.x + (int)blockIdx.x;

    // Skip a block whose inputs within reach are all +0 (rec.prev: the boxes of
    // the non-zero outputs of the pass before, per block).  Its outputs are then
    // +0: the stencil of +0 inputs is +0, as are copied boundary voxels.  They
    // are written only where the buffer may hold something else (rec.prev2:
    // this buffer's boxes as the pass before last left them).
    // Planes [zc0, zc1) of the chunk are computed; the others are +0.
    int zc0 = zb, zc1 = ze;
    if (rec.prev || rec.src) {
        __shared__ int szr[2];
        if (tid == 0) {
            szr[0] = INT_MAX;
            szr[1] = INT_MIN;
        }
        __syncthreads();
        bool near = false;
        const int *q = nullptr;
        if (rec.prev) {
            if (tid < 27) {
                const int nx = (int)blockIdx.x + tid % 3 - 1, ny = (int)blockIdx.y + (tid / 3) % 3 - 1,
                          nz = (int)blockIdx.z + tid / 9 - 1;
                if (nx >= 0 && nx < (int)gridDim.x && ny >= 0 && ny < (int)gridDim.y && nz >= 0 &&
                    nz < (int)gridDim.z)
                    q = rec.prev + 6 * ((nz * (int)gridDim.y + ny) * (int)gridDim.x + nx);
            }
        } else if (tid == 0) {
            q = rec.src;
        }
        if (q) {
            const int xe = min(x0 + TX, bx.x1) - 1, ye = min(y0 + TY, bx.y1) - 1;
            near = q[1] >= 0 && q[0] <= ze - 1 + 2 && q[1] >= zb - 2 && q[2] <= ye + 2 && q[3] >= y0 - 2 &&
                   q[4] <= xe + 2 && q[5] >= x0 - 2;
            if (near) {
                atomicMin(&szr[0], q[0]);
                atomicMax(&szr[1], q[1]);
            }
        }
        const bool any = __syncthreads_or(near);
        // An output plane z reads input planes z-2..z+2 only; where the near
        // boxes hold no non-zero input plane in that reach, its outputs are +0.
        if (any) {
            zc0 = max(zb, szr[0] - 2);
            zc1 = min(ze, szr[1] + 3);
        } else {
            zc0 = zc1 = ze;
        }
        // +0 is written where the buffer may hold something else: all of the
        // block on the pass after the first, else the planes of this buffer's
        // box from the pass before last.
        int w0 = zb, w1 = ze;
        if (rec.prev2) {
            const int *q2 = rec.prev2 + 6 * blk;
            w0 = q2[1] >= 0 ? max(zb, q2[0]) : ze;
            w1 = q2[1] >= 0 ? min(ze, q2[1] + 1) : zb;
        }
        for (int z = w0; z < w1; ++z) {
            if (z >= zc0 && z < zc1) continue;
#pragma unroll
            for (int k = 0; k < RY; ++k) {
                const int y = y0 + ty * RY + k, x = x0 + tx;
                if (x < bx.x1 && y < bx.y1) out[(int64_t)z * sz + (int64_t)y * sy + x] = 0.0f;
            }
        }
        if (!any) {
            if (tid == 0) {
                int *o = rec.out + 6 * blk;
                o[0] = INT_MAX; o[1] = -1; o[2] = INT_MAX; o[3] = -1; o[4] = INT_MAX; o[5] = -1;
            }
            return;
        }
    }
    // box of this thread's non-zero outputs (float bits, so -0 counts as
    // non-zero): bit k of `rows` for row y0 + ty*RY + k of its column, and the
    // first and last plane (planes are computed in increasing order)
    uint32_t rows = 0u;
    int zlo = INT_MAX, zhi = -1;

    // Per-thread constants of the z-independent parts: the in-plane offset and
    // in-plane validity (inside the volume and the box) of the thread's input
    // voxels, and of the voxels it computes in each round.
    auto plane_ok = [&](int y, int x) {
        return y >= 0 && y < Y && x >= 0 && x < X && y >= bx.y0 && y < bx.y1 && x >= bx.x0 && x < bx.x1;
    };
    auto z_ok = [&](int z) { return z >= 0 && z < Z && z >= bx.z0 && z < bx.z1; };
    uint32_t off[kPairLoads];   // y * X + x < Y * X <= V < 2^32
    uint32_t okmask = 0;
#pragma unroll
    for (int k = 0; k < kPairLoads; ++k) {
        const int i = tid + k * kSmoothThreads;
        const int ly = i / PW, lx = i - (i / PW) * PW;
        const int y = y0 - 2 + ly, x = x0 - 2 + lx;
        const bool ok = i < PW * PH && plane_ok(y, x);
        off[k] = ok ? (uint32_t)y * (uint32_t)X + (uint32_t)x : 0u;
        okmask |= (ok ? 1u : 0u) << k;
    }
    // one round at (z, frame row, frame column) from three planes of the round
    // before; pin: in-plane validity, pint: in-plane interior (0 < y < Y-1, 0 < x < X-1)
    auto round_at = [&](const float (*m)[PW], const float (*c)[PW], const float (*p)[PW], bool zin, bool zint,
                        bool pin, bool pint, int ly, int lx) -> float {
        if (!(zin && pin)) return 0.0f;
        return (zint && pint) ? curvature_update<PW>(m, c, p, ly, lx) : c[ly][lx];
    };
    auto interior = [&](int y, int x) { return y > 0 && y < Y - 1 && x > 0 && x < X - 1; };
    // the thread's core voxels (rows 2 + ty*RY + k, column 2 + tx of the frame)
    const int xc = x0 + tx;
    bool core_in[RY], core_int[RY];
#pragma unroll
    for (int k = 0; k < RY; ++k) {
        const int y = y0 + ty * RY + k;
        core_in[k] = plane_ok(y, xc);
        core_int[k] = interior(y, xc);
    }
    // the thread's ring voxel of round t+1 (rows 1 and PH-2, columns 1 and PW-2)
    constexpr int kRing = 2 * (PW - 2) + 2 * TY;
    int rly = 0, rlx = 0;
    if (tid < PW - 2) {
        rly = 1; rlx = 1 + tid;
    } else if (tid < 2 * (PW - 2)) {
        rly = PH - 2; rlx = 1 + tid - (PW - 2);
    } else if (tid < 2 * (PW - 2) + TY) {
        rly = 2 + tid - 2 * (PW - 2); rlx = 1;
    } else if (tid < kRing) {
        rly = 2 + tid - 2 * (PW - 2) - TY; rlx = PW - 2;
    }
    const bool ring_in = tid < kRing && plane_ok(y0 - 2 + rly, x0 - 2 + rlx);
    const bool ring_int = interior(y0 - 2 + rly, x0 - 2 + rlx);

    float regs[kPairLoads];
    auto fetch = [&](int z) {
        const bool zv = z_ok(z);
        const Tin *plane = in + (int64_t)(zv ? z : 0) * sz;
#pragma unroll
        for (int k = 0; k < kPairLoads; ++k)
            regs[k] = (zv && ((okmask >> k) & 1u)) ? ldf(plane, (int64_t)off[k]) : 0.0f;
    };
    // One input plane q: stage it, round t+1 at plane q-1, round t+2 at plane
    // q-2.  The three LDS slots rotate with q (slot3(q) = q mod 3); S0 =
    // slot3(q) is a template constant, so every LDS address is a per-thread
    // base plus an immediate offset (the loop below dispatches on it).
    auto plane = [&](int q, auto phase) {
        constexpr int S0 = decltype(phase)::value, S1 = (S0 + 2) % 3, S2 = (S0 + 1) % 3;   // q, q-1, q-2
        float *l0 = &L0[S0][0][0];
#pragma unroll
        for (int k = 0; k < kPairLoads; ++k) {
            const int i = tid + k * kSmoothThreads;
            if (i < PW * PH) l0[i] = regs[k];
        }
        __syncthreads();
        if (q + 1 <= zc1 + 1) fetch(q + 1);
        // round t+1 at plane q-1 over frame rows 1..PH-2, columns 1..PW-2
        const int r1 = q - 1;
        if (r1 >= zc0 - 1) {
            const float (*m)[PW] = L0[S2], (*c)[PW] = L0[S1], (*p)[PW] = L0[S0];
            float (*o)[PW] = L1[S1];
            const bool zin = z_ok(r1), zint = r1 > 0 && r1 < Z - 1;
#pragma unroll
            for (int k = 0; k < RY; ++k) {
                const int ly = 2 + ty * RY + k, lx = 2 + tx;
                o[ly][lx] = round_at(m, c, p, zin, zint, core_in[k], core_int[k], ly, lx);
            }
            // the ring around the 64x16 core
            if (tid < kRing) o[rly][rlx] = round_at(m, c, p, zin, zint, ring_in, ring_int, rly, rlx);
        }
        __syncthreads();
        // round t+2 at plane q-2 over the 64x16 core, written out inside the box
        const int r2 = q - 2;
        if (r2 >= zc0 && r2 < zc1) {
            const float (*m)[PW] = L1[S0], (*c)[PW] = L1[S2], (*p)[PW] = L1[S1];   // q-3, q-2, q-1
            const bool zint = r2 > 0 && r2 < Z - 1;
            uint32_t nzrows = 0u;
#pragma unroll
            for (int k = 0; k < RY; ++k) {
                const int ly = 2 + ty * RY + k, lx = 2 + tx;
                const int y = y0 + ty * RY + k;
                if (core_in[k]) {
                    const float v = round_at(m, c, p, true, zint, true, core_int[k], ly, lx);
                    out[(int64_t)r2 * sz + (int64_t)y * sy + xc] = v;
                    nzrows |= __float_as_uint(v) != 0u ? 1u << k : 0u;
                }
            }
            if (nzrows) {
                zlo = min(zlo, r2);
                zhi = r2;
                rows |= nzrows;
            }
        }
    };
    fetch(zc0 - 2);
    int ph = slot3(zc0 - 2);
    for (int q = zc0 - 2; q <= zc1 + 1; ++q) {
        if (ph == 0)
            plane(q, std::integral_constant<int, 0>{});
        else if (ph == 1)
            plane(q, std::integral_constant<int, 1>{});
        else
            plane(q, std::integral_constant<int, 2>{});
        ph = ph == 2 ? 0 : ph + 1;
    }
    if (rec.out) {
        const int yr = y0 + ty * RY;
        int bb[6] = {zlo, zhi, rows ? yr + __builtin_ctz(rows) : INT_MAX, rows ? yr + 31 - __builtin_clz(rows) : -1,
                     rows ? xc : INT_MAX, rows ? xc : -1};
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
            for (int k = 0; k < 6; k += 2) {
                bb[k] = min(bb[k], __shfl_xor(bb[k], off, 64));
                bb[k + 1] = max(bb[k + 1], __shfl_xor(bb[k + 1], off, 64));
            }
        }
        if ((tid & 63) == 0)
            for (int k = 0; k < 6; ++k) red[tid >> 6][k] = bb[k];
        __syncthreads();
        if (tid < 6) {
            int v = red[0][tid];
            for (int w = 1; w < kSmoothThreads / 64; ++w) v = (tid & 1) ? max(v, red[w][tid]) : min(v, red[w][tid]);
            rec.out[6 * blk + tid] = v;
        }
    }
}

// Value of label `label` at voxel v for the argmax over labels: phi inside
// the label's box, 0 outside; the large-volume path sets the first label's
// negative values to 0 before the running maximum (pyopencl_large.py:464-469).
// A NaN becomes +inf for label 0 and -inf for the others (see below).
__device__ __forceinline__ float argmax_value(const float *__restrict__ phi, int64_t v, bool inb, int label,
                                              int clamp_first)
{
    float p = inb ? phi[v] : 0.0f;
    if (clamp_first && label == 0 && p < 0.0f) p = 0.0f;
    if (p != p) p = label == 0 ? INFINITY : -INFINITY;
    return p;
}

// The winner of a voxel is the label with the largest value, the lowest label
// on ties (the np.argmax of rw_small.py:137, and the "replace when strictly
// greater" pass over the labels in order of pyopencl_large.py:233-241), with a
// NaN of label 0 beating every value and a NaN of any other label losing to
// every value, as in that in-order pass.  This compares (value, -label) pairs,
// so labels may be applied in any order and partial results merged.
__device__ __forceinline__ bool argmax_beats(float p, int l, float q, int m)
{
    return p > q || (p == q && l < m);
}

__device__ __forceinline__ bool in_box(int64_t v, int Y, int X, const Box &bx)
{
    const int x = (int)(v % X), y = (int)((v / X) % Y), z = (int)(v / ((int64_t)X * Y));
    return z >= bx.z0 && z < bx.z1 && y >= bx.y0 && y < bx.y1 && x >= bx.x0 && x < bx.x1;
}

// One label's share of a running argmax (best, bidx); `first` marks the first
// label applied to it.
__global__ __launch_bounds__(256) void k_argmax_update(const float *__restrict__ phi,
                                                       float *__restrict__ best,
                                                       uint8_t *__restrict__ bidx, int label, int Z, int Y,
                                                       int X, Box bx, int clamp_first, int first)
{
    const int64_t V = (int64_t)Z * Y * X;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const float p = argmax_value(phi, v, in_box(v, Y, X, bx), label, clamp_first);
    if (first || argmax_beats(p, label, best[v], bidx[v])) {
        best[v] = p;
        bidx[v] = (uint8_t)label;
    }
}

struct LabelValues {
    uint8_t v[256];
};

// Bounding box of the non-zero hits of each label (label = blockIdx.y, hits
// of label l at hits + l * V): box[6l..6l+5] = min z, max z, min y, max y,
// min x, max x (min fields start at INT_MAX, max at -1).  16-byte loads when
// X is a multiple of 4 and `hits` is 16-byte aligned, else one wave per x row.
__global__ __launch_bounds__(256) void k_hits_box(const uint32_t *__restrict__ hits, int Z, int Y, int X,
                                                  int *__restrict__ box)
{
    const int64_t V = (int64_t)Z * Y * X;
    const uint32_t *h = hits + (int64_t)blockIdx.y * V;
    const int lane = threadIdx.x & 63;
    int b[6] = {INT_MAX, -1, INT_MAX, -1, INT_MAX, -1};
    // non-zeros at x offsets lo..hi from element e of the label (e < V < 2^32)
    auto note = [&](uint32_t e, int lo, int hi) {
        const uint32_t row = e / (uint32_t)X, zq = row / (uint32_t)Y;
        const int x = (int)(e - row * (uint32_t)X), z = (int)zq, y = (int)(row - zq * (uint32_t)Y);
        b[0] = min(b[0], z); b[1] = max(b[1], z);
        b[2] = min(b[2], y); b[3] = max(b[3], y);
        b[4] = min(b[4], x + lo); b[5] = max(b[5], x + hi);
    };
    if ((X & 3) == 0 && (reinterpret_cast<uintptr_t>(hits) & 15) == 0) {
        // 16-byte groups never cross a row; four loads in flight per thread
        // n4 < 2^30 and stride <= 2^18, so the indexes stay below 2^31
        const uint4 *h4 = reinterpret_cast<const uint4 *>(h);
        const uint32_t n4 = (uint32_t)(V / 4), stride = gridDim.x * blockDim.x;
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += 4 * stride) {
            uint4 q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) q[u] = i + u * stride < n4 ? h4[i + u * stride] : make_uint4(0, 0, 0, 0);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (q[u].x | q[u].y | q[u].z | q[u].w)
                    note(4u * (i + u * stride), q[u].x ? 0 : q[u].y ? 1 : q[u].z ? 2 : 3,
                         q[u].w ? 3 : q[u].z ? 2 : q[u].y ? 1 : 0);
        }
    } else {
        // one wave per x row
        const int64_t nrows = (int64_t)Z * Y;
        for (int64_t r = blockIdx.x * 4 + (threadIdx.x >> 6); r < nrows; r += gridDim.x * 4) {
            const uint32_t *row = h + r * X;
            int lo = INT_MAX, hi = -1;
            for (int x = lane; x < X; x += 64)
                if (row[x]) {
                    lo = min(lo, x);
                    hi = x;
                }
            if (hi >= 0) note((uint32_t)(r * X), lo, hi);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 6; k += 2) {
            b[k] = min(b[k], __shfl_xor(b[k], off, 64));
            b[k + 1] = max(b[k + 1], __shfl_xor(b[k + 1], off, 64));
        }
    }
    if (lane == 0 && b[1] >= 0) {
        int *bl = box + 6 * blockIdx.y;
        for (int k = 0; k < 6; k += 2) {
            atomicMin(&bl[k], b[k]);
            atomicMax(&bl[k + 1], b[k + 1]);
        }
    }
}

// The last label and up to two partial argmaxes (null when unused) merged,
// and the winner's label value written to out.
__global__ __launch_bounds__(256) void k_argmax_final(const float *__restrict__ phi, int label, Box bx,
                                                      int clamp_first, const float *__restrict__ best1,
                                                      const uint8_t *__restrict__ bidx1,
                                                      const float *__restrict__ best2,
                                                      const uint8_t *__restrict__ bidx2, LabelValues lv,
                                                      uint8_t *__restrict__ out, int Z, int Y, int X)
{
    const int64_t V = (int64_t)Z * Y * X;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    float p = argmax_value(phi, v, in_box(v, Y, X, bx), label, clamp_first);
    int l = label;
    if (best1 && argmax_beats(best1[v], bidx1[v], p, l)) {
        p = best1[v];
        l = bidx1[v];
    }
    if (best2 && argmax_beats(best2[v], bidx2[v], p, l)) l = bidx2[v];
    out[v] = lv.v[l];
}

__global__ void k_box_init(int *box, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) box[i] = (i & 1) ? -1 : INT_MAX;
}

__global__ __launch_bounds__(256) void k_map_values(const uint8_t *__restrict__ bidx, LabelValues lv,
                                                    uint8_t *__restrict__ out, int64_t V)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    out[v] = lv.v[bidx[v]];
}

template <typename Tin>
int launch_step(hipStream_t st, const Tin *in, float *out, int Z, int Y, int X, const Box &bx)
{
    dim3 grid((bx.x1 - bx.x0 + TX - 1) / TX, (bx.y1 - bx.y0 + TY - 1) / TY, (bx.z1 - bx.z0 + kZChunk - 1) / kZChunk);
    k_smooth_step<Tin><<<grid, kSmoothThreads, 0, st>>>(in, out, Z, Y, X, bx);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

template <typename Tin>
int launch_pair(hipStream_t st, const Tin *in, float *out, int Z, int Y, int X, const Box &bx,
                PassRecords rec = PassRecords{nullptr, nullptr, nullptr})
{
    dim3 grid((bx.x1 - bx.x0 + TX - 1) / TX, (bx.y1 - bx.y0 + TY - 1) / TY,
              (bx.z1 - bx.z0 + kPairZChunk - 1) / kPairZChunk);
    k_smooth_pair<Tin><<<grid, kSmoothThreads, 0, st>>>(in, out, Z, Y, X, bx, rec);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

// Buffer sets of smooth_argmax: labels smoothed at the same time.
constexpr int kSmoothSets = 3;

// Ints of the pass records of one pair pass over the whole volume (an upper
// bound for any box): 6 per block of the launch grid.
static int64_t pass_record_ints(int64_t Z, int64_t Y, int64_t X)
{
    return 6 * ((X + TX - 1) / TX) * ((Y + TY - 1) / TY) * ((Z + kPairZChunk - 1) / kPairZChunk);
}

// iters >= 1 rounds starting from `in`, two per pass and a single pass for an
// odd count; the result lands in *result (one of a / b).  With recs (three
// arrays of pass_record_ints each), pair passes after the first skip the
// blocks that only see +0 (PassRecords).
template <typename Tin>
int smooth_rounds(hipStream_t st, const Tin *in, float *a, float *b, int Z, int Y, int X, int iters,
                  const Box &bx, float **result, int *recs = nullptr, int64_t rec_ints = 0,
                  const int *src = nullptr)
{
    auto rec_of = [&](int pass) -> PassRecords {
        if (!recs) return PassRecords{nullptr, nullptr, nullptr};
        int *r[3] = {recs, recs + rec_ints, recs + 2 * rec_ints};
        return PassRecords{pass >= 1 ? r[(pass - 1) % 3] : nullptr, pass >= 2 ? r[(pass - 2) % 3] : nullptr,
                           r[pass % 3], pass == 0 ? src : nullptr};
    };
    int rc = iters >= 2 ? launch_pair<Tin>(st, in, a, Z, Y, X, bx, rec_of(0))
                        : launch_step<Tin>(st, in, a, Z, Y, X, bx);
    if (rc) return rc;
    float *cur = a, *nxt = b;
    int pass = 1;
    for (int it = iters >= 2 ? 2 : 1; it < iters; it += 2, ++pass) {
        rc = iters - it >= 2 ? launch_pair<float>(st, cur, nxt, Z, Y, X, bx, rec_of(pass))
                             : launch_step<float>(st, cur, nxt, Z, Y, X, bx);
        if (rc) return rc;
        float *t = cur; cur = nxt; nxt = t;
    }
    *result = cur;
    return BMX_OK;
}

}  // namespace

extern "C" {

int bmx_smooth(bmx_ctx *ctx, float *phi, float *tmp, int64_t Z, int64_t Y, int64_t X, int iters)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(phi && tmp, "bmx_smooth: null pointer");
    BMX_CHECK(Z >= 1 && Y >= 1 && X >= 1 && Z * Y * X < (int64_t)1 << 32, "bmx_smooth: bad dims");
    BMX_CHECK(iters >= 0, "bmx_smooth: negative iters");
    float *cur = phi, *nxt = tmp;
    const Box whole{0, (int)Z, 0, (int)Y, 0, (int)X};
    for (int it = 0; it < iters; it += 2) {
        rc = iters - it >= 2 ? launch_pair<float>(ctx->stream, cur, nxt, (int)Z, (int)Y, (int)X, whole)
                             : launch_step<float>(ctx->stream, cur, nxt, (int)Z, (int)Y, (int)X, whole);
        if (rc) return rc;
        float *t = cur; cur = nxt; nxt = t;
    }
    if (cur != phi)
        BMX_HIP(hipMemcpyAsync(phi, cur, (size_t)(Z * Y * X) * sizeof(float), hipMemcpyDeviceToDevice,
                               ctx->stream));
    return BMX_OK;
}

int bmx_smooth_argmax(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int64_t Z, int64_t Y, int64_t X,
                      int iters, const uint8_t *values, uint8_t *out)
{
    return bmx_smooth_argmax_ex(ctx, hits, nlabels, Z, Y, X, iters, values, out, 0);
}

int bmx_smooth_argmax_ex(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int64_t Z, int64_t Y, int64_t X,
                         int iters, const uint8_t *values, uint8_t *out, int clamp_first)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(hits && values && out, "bmx_smooth_argmax: null pointer");
    BMX_CHECK(nlabels >= 1 && nlabels <= 256, "bmx_smooth_argmax: nlabels %d not in [1,256]", nlabels);
    BMX_CHECK(Z >= 1 && Y >= 1 && X >= 1 && Z * Y * X < (int64_t)1 << 32, "bmx_smooth_argmax: bad dims");
    BMX_CHECK(iters >= 1, "bmx_smooth_argmax: iters must be >= 1");
    const int64_t V = Z * Y * X;
    const int64_t rec_ints = pass_record_ints(Z, Y, X);
    // Buffer sets of the labels being smoothed at the same time: two phi
    // volumes and the pass records, and for sets 1.. a partial argmax (set 0
    // has one only when the labels run one after another).
    float *a[kSmoothSets] = {}, *b[kSmoothSets] = {}, *pbest[kSmoothSets] = {};
    uint8_t *pidx[kSmoothSets] = {};
    int *recs[kSmoothSets] = {};
    auto alloc_set = [&](int k) {
        hipError_t e = hipMalloc((void **)&a[k], (size_t)V * sizeof(float));
        if (e == hipSuccess) e = hipMalloc((void **)&b[k], (size_t)V * sizeof(float));
        if (e == hipSuccess) e = hipMalloc((void **)&recs[k], (size_t)(3 * rec_ints) * sizeof(int));
        if (e == hipSuccess && k >= 1) e = hipMalloc((void **)&pbest[k], (size_t)V * sizeof(float));
        if (e == hipSuccess && k >= 1) e = hipMalloc((void **)&pidx[k], (size_t)V);
        return e;
    };
    auto free_set = [&](int k) {
        if (a[k]) (void)hipFree(a[k]);
        if (b[k]) (void)hipFree(b[k]);
        if (recs[k]) (void)hipFree(recs[k]);
        if (pbest[k]) (void)hipFree(pbest[k]);
        if (pidx[k]) (void)hipFree(pidx[k]);
        a[k] = b[k] = pbest[k] = nullptr;
        recs[k] = nullptr;
        pidx[k] = nullptr;
    };
    int *dbox = nullptr;
    hipError_t e = hipMalloc((void **)&dbox, (size_t)nlabels * 6 * sizeof(int));
    if (e == hipSuccess) e = alloc_set(0);
    int nset = 1;
    std::vector<hipStream_t> used;
    if (e == hipSuccess) {
        // Each label is smoothed only inside the box of its non-zero hits grown by
        // `iters`: where all 19 stencil inputs are 0 the update gives exactly 0,
        // so phi's support grows by at most one voxel per round and stays inside.
        std::vector<int> box((size_t)nlabels * 6);
        k_box_init<<<bmx_blocks(nlabels * 6, 256), 256, 0, ctx->stream>>>(dbox, nlabels * 6);
        k_hits_box<<<dim3((unsigned)std::min<int64_t>((Z * Y + 15) / 16, 1024), (unsigned)nlabels), 256, 0,
                     ctx->stream>>>(hits, (int)Z, (int)Y, (int)X, dbox);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(box.data(), dbox, box.size() * sizeof(int), hipMemcpyDeviceToHost, ctx->stream) !=
                hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = bmx_set_error(BMX_E_DEVICE, "bmx_smooth_argmax: hit boxes failed");
        std::vector<Box> bxs((size_t)nlabels, Box{0, 0, 0, 0, 0, 0});
        std::vector<int64_t> cost((size_t)nlabels, 0);
        int big = 0;
        for (int l = 0; l < nlabels; ++l) {
            const int *h = &box[(size_t)6 * l];
            if (h[1] < 0) continue;
            Box &bx = bxs[(size_t)l];
            bx = Box{max(h[0] - iters, 0), min(h[1] + 1 + iters, (int)Z), max(h[2] - iters, 0),
                     min(h[3] + 1 + iters, (int)Y), max(h[4] - iters, 0), min(h[5] + 1 + iters, (int)X)};
            cost[(size_t)l] = (int64_t)(bx.z1 - bx.z0) * (bx.y1 - bx.y0) * (bx.x1 - bx.x0);
            if (cost[(size_t)l] > cost[(size_t)big]) big = l;
        }
        // Labels are smoothed side by side (a small label's passes leave much
        // of the GPU idle): the label with the largest box on the context
        // stream in set 0, the others split over sets 1.. on streams of their
        // own, largest boxes first, each set keeping a partial argmax of its
        // labels.  A last kernel merges the partial argmaxes with the largest
        // box's label.  Without the memory or the streams for a second set,
        // the labels run one after another into set 0's argmax.
        if (rc == BMX_OK && nlabels >= 2 && bmx_aux_streams(ctx) == BMX_OK) {
            while (nset < std::min(kSmoothSets, nlabels)) {
                if (alloc_set(nset) != hipSuccess) {
                    (void)hipGetLastError();
                    free_set(nset);
                    break;
                }
                ++nset;
            }
        }
        (void)hipGetLastError();
        auto smooth_label = [&](int l, int k, hipStream_t st, float **res) -> int {
            *res = a[k];
            if (cost[(size_t)l] == 0) return BMX_OK;
            return smooth_rounds<uint32_t>(st, hits + (int64_t)l * V, a[k], b[k], (int)Z, (int)Y, (int)X, iters,
                                           bxs[(size_t)l], res, recs[k], rec_ints, dbox + 6 * l);
        };
        auto update = [&](int l, const float *res, float *best, uint8_t *bidx, int first, hipStream_t st) -> int {
            k_argmax_update<<<bmx_blocks(V, 256), 256, 0, st>>>(res, best, bidx, l, (int)Z, (int)Y, (int)X,
                                                                bxs[(size_t)l], clamp_first, first);
            return hipGetLastError() == hipSuccess ? BMX_OK
                                                   : bmx_set_error(BMX_E_DEVICE, "bmx_smooth_argmax: launch failed");
        };
        LabelValues lv;
        for (int i = 0; i < 256; ++i) lv.v[i] = i < nlabels ? values[i] : 0;
        if (rc == BMX_OK && nset == 1) {
            // one after another: set 0 keeps the argmax
            hipError_t e0 = hipMalloc((void **)&pbest[0], (size_t)V * sizeof(float));
            if (e0 == hipSuccess) e0 = hipMalloc((void **)&pidx[0], (size_t)V);
            if (e0 != hipSuccess) {
                (void)hipGetLastError();
                rc = bmx_set_error(e0 == hipErrorOutOfMemory ? BMX_E_OOM : BMX_E_DEVICE,
                                   "bmx_smooth_argmax: allocation failed: %s", hipGetErrorString(e0));
            }
        }
        if (rc == BMX_OK && nset == 1) {
            for (int l = 0; l < nlabels && rc == BMX_OK; ++l) {
                float *res = nullptr;
                rc = smooth_label(l, 0, ctx->stream, &res);
                if (rc == BMX_OK) rc = update(l, res, pbest[0], pidx[0], l == 0, ctx->stream);
            }
            if (rc == BMX_OK) {
                k_map_values<<<bmx_blocks(V, 256), 256, 0, ctx->stream>>>(pidx[0], lv, out, V);
                if (hipGetLastError() != hipSuccess)
                    rc = bmx_set_error(BMX_E_DEVICE, "bmx_smooth_argmax: launch failed");
            }
        } else if (rc == BMX_OK) {
            // rolling labels, largest box first, each to the set with the least work so far
            std::vector<int> rest;
            for (int l = 0; l < nlabels; ++l)
                if (l != big) rest.push_back(l);
            std::stable_sort(rest.begin(), rest.end(),
                             [&](int p, int q) { return cost[(size_t)p] > cost[(size_t)q]; });
            std::vector<int64_t> load((size_t)nset, 0);
            std::vector<int> nused((size_t)nset, 0);
            hipEvent_t start = ctx->aux_ev[0];
            if (hipEventRecord(start, ctx->stream) != hipSuccess)
                rc = bmx_set_error(BMX_E_DEVICE, "bmx_smooth_argmax: event failed");
            for (int k = 1; k < nset && rc == BMX_OK; ++k) {
                used.push_back(ctx->aux[k - 1]);
                if (hipStreamWaitEvent(ctx->aux[k - 1], start, 0) != hipSuccess)
                    rc = bmx_set_error(BMX_E_DEVICE, "bmx_smooth_argmax: event failed");
            }
            float *res_big = nullptr;
            if (rc == BMX_OK) rc = smooth_label(big, 0, ctx->stream, &res_big);
            for (size_t i = 0; i < rest.size() && rc == BMX_OK; ++i) {
                int k = 1;
                for (int j = 2; j < nset; ++j)
                    if (load[(size_t)j] < load[(size_t)k]) k = j;
                const int l = rest[i];
                load[(size_t)k] += cost[(size_t)l] + V / 64;   // V / 64: the update's full-volume pass
                float *res = nullptr;
                rc = smooth_label(l, k, ctx->aux[k - 1], &res);
                if (rc == BMX_OK)
                    rc = update(l, res, pbest[k], pidx[k], nused[(size_t)k] == 0, ctx->aux[k - 1]);
                ++nused[(size_t)k];
            }
            for (int k = 1; k < nset && rc == BMX_OK; ++k) {
                if (hipEventRecord(ctx->aux_ev[k], ctx->aux[k - 1]) != hipSuccess ||
                    hipStreamWaitEvent(ctx->stream, ctx->aux_ev[k], 0) != hipSuccess)
                    rc = bmx_set_error(BMX_E_DEVICE, "bmx_smooth_argmax: event failed");
            }
            if (rc == BMX_OK) {
                auto part = [&](int k, bool idx) -> const void * {
                    if (k >= nset || nused[(size_t)k] == 0) return nullptr;
                    return idx ? (const void *)pidx[k] : (const void *)pbest[k];
                };
                k_argmax_final<<<bmx_blocks(V, 256), 256, 0, ctx->stream>>>(
                    res_big, big, bxs[(size_t)big], clamp_first, (const float *)part(1, false),
                    (const uint8_t *)part(1, true), (const float *)part(2, false), (const uint8_t *)part(2, true),
                    lv, out, (int)Z, (int)Y, (int)X);
                if (hipGetLastError() != hipSuccess)
                    rc = bmx_set_error(BMX_E_DEVICE, "bmx_smooth_argmax: launch failed");
            }
        }
        used.push_back(ctx->stream);
        for (hipStream_t st : used) {
            hipError_t s = hipStreamSynchronize(st);
            if (s != hipSuccess && rc == BMX_OK)
                rc = bmx_set_error(BMX_E_DEVICE, "bmx_smooth_argmax: %s", hipGetErrorString(s));
        }
    } else {
        (void)hipGetLastError();
        rc = bmx_set_error(e == hipErrorOutOfMemory ? BMX_E_OOM : BMX_E_DEVICE,
                           "bmx_smooth_argmax: allocation failed: %s", hipGetErrorString(e));
    }
    for (int k = 0; k < kSmoothSets; ++k) free_set(k);
    if (dbox) (void)hipFree(dbox);
    return rc;
}

}  // extern "C"
