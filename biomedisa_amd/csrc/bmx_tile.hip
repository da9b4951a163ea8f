// bmx_tile.hip -- tile-owner walker: hits privatised in LDS.
//
// Same transition law and Philox stream as the other walkers (bmx_walk.hip);
// only the order of work and where the increments land differ, so the hit
// counts are identical.
//
// Why: a device-scope atomic on this GPU is performed past the XCD's L2 (the
// L2s of the 8 XCDs are not coherent with one another), so one global atomic
// per step caps the walker at the fabric's atomic rate whatever the locality.
// Here a workgroup owns one (core tile, label) bucket: a C^3 core tile plus a
// halo of H voxels, i.e. an R^3 region (R = C + 2H).  It stages the region's
// image (with one more voxel of halo for the neighbour reads) and R^3 hit
// counters in LDS and runs every walker of the bucket - walkers whose current
// voxel lies in the core - until the walker leaves the region, finishes its
// walk, or has made `cap` steps in this round.  Steps taken outside the region
// (inside the Philox block during which the walker left) read the image from
// global memory and add their hit with a global atomic.  At the end the
// workgroup adds each non-zero LDS counter to the hit volume with one global
// atomic.  Between rounds the walkers are counting-sorted by (core tile,
// label) again.  A walker pauses only at a multiple of 4 steps, so every
// Philox block is used whole within one round.
#include "bmx_walk_common.h"

namespace {

constexpr int kTileThreads = 1024;

template <typename T> struct TileCfg;
template <> struct TileCfg<uint8_t> {
    static constexpr int C = 16, H = 6;
};
template <> struct TileCfg<float> {
    static constexpr int C = 16, H = 4;
};

template <typename T> constexpr int tile_R() { return TileCfg<T>::C + 2 * TileCfg<T>::H; }
template <typename T> constexpr size_t tile_img_bytes()
{
    return ((size_t)(tile_R<T>() + 2) * (tile_R<T>() + 2) * (tile_R<T>() + 2) * sizeof(T) + 15) & ~(size_t)15;
}
template <typename T> constexpr size_t tile_lds_bytes()
{
    return tile_img_bytes<T>() +
           (size_t)(tile_R<T>() + 2) * (tile_R<T>() + 2) * (tile_R<T>() + 2) * sizeof(uint32_t);
}

struct TileGeom {
    uint32_t ntz, nty, ntx, nlabels;
};

template <int C>
__device__ __forceinline__ uint32_t tile_key(uint32_t pos, uint32_t label, const WalkGeom &g,
                                             const TileGeom &tg)
{
    const uint32_t z = pos / g.sz;
    const uint32_t r = pos - z * g.sz;
    const uint32_t y = r / g.sy;
    const uint32_t x = r - y * g.sy;
    return (((z / C) * tg.nty + y / C) * tg.ntx + x / C) * tg.nlabels + label;
}

// Walker state: {pos, seed index, walk | label << 16, step}.  With overrides a
// seed running fewer than nbrw walks gets its extra walkers created finished
// (step = sorw), so the first re-bucketing drops them.
__global__ __launch_bounds__(256) void k_tile_init(const bmx_seed *__restrict__ seeds, int64_t nseeds,
                                                   int nbrw, int sorw, int overrides, uint4 *__restrict__ st)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseeds * nbrw) return;
    const int64_t s = i / nbrw;
    const bmx_seed sd = seeds[s];
    const uint32_t w = (uint32_t)(i - s * nbrw);
    const uint32_t walks = overrides ? (sd.reserved & 0xFFFFu) : 0u;
    const uint32_t step0 = (walks != 0u && w >= walks) ? (uint32_t)sorw : 0u;
    st[i] = make_uint4(sd.pos, (uint32_t)s, w | (sd.label << 16), step0);
}

__global__ __launch_bounds__(256) void k_seed_walks(bmx_seed *__restrict__ seeds, int64_t n, uint32_t value)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) seeds[i].reserved = value;
}

template <int C>
__global__ __launch_bounds__(256) void k_tile_hist(const uint4 *__restrict__ st, int64_t n, WalkGeom g,
                                                   TileGeom tg, int sorw, uint32_t *counts)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    bool live = false;
    uint32_t key = 0;
    if (i < n) {
        const uint4 w4 = st[i];
        live = (int)w4.w < sorw;
        key = tile_key<C>(w4.x, w4.z >> 16, g, tg);
    }
    (void)agg_atomic_inc(counts, key, live);
}

template <int C>
__global__ __launch_bounds__(256) void k_tile_scatter(const uint4 *__restrict__ in, int64_t n,
                                                      uint4 *__restrict__ out, uint32_t *cursor,
                                                      WalkGeom g, TileGeom tg, int sorw)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    uint4 w4 = make_uint4(0, 0, 0, 0);
    bool live = false;
    uint32_t key = 0;
    if (i < n) {
        w4 = in[i];
        live = (int)w4.w < sorw;
        key = tile_key<C>(w4.x, w4.z >> 16, g, tg);
    }
    const uint32_t dst = agg_atomic_inc(cursor, key, live);
    if (live) out[dst] = w4;
}

// Non-empty buckets of the exclusive offsets off[0..nkeys] -> (key, offset, count).
__global__ __launch_bounds__(256) void k_bucket_list(const uint32_t *__restrict__ off, int64_t nkeys,
                                                     uint32_t *bkey, uint32_t *boff, uint32_t *bcnt,
                                                     uint32_t *nb)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t c = 0, o = 0;
    if (k < nkeys) {
        o = off[k];
        c = off[k + 1] - o;
    }
    const uint32_t idx = agg_atomic_inc(nb, 0u, c > 0);
    if (c > 0) {
        bkey[idx] = (uint32_t)k;
        boff[idx] = o;
        bcnt[idx] = c;
    }
}

// Weight exp(-(B-A)^2 * div1) of the reference (pyopencl_small.py:187-190).
__device__ __forceinline__ float weight_nonpos(float B, float A, float div1)
{
    return bmx_weight(B, A, div1);
}

// One Philox block (4 steps) of one walker.  Local coordinates (lz,ly,lx)
// address the LDS tile, whose voxel (1,1,1) is the region origin; the region
// proper is [1,R]^3 and the tile adds one voxel on each side for neighbour
// reads.  li is the walker's index in the tile; the global voxel index is
// rebuilt from the coordinates only on the rare steps that leave the tile.
// EDGE: the tile may reach the volume boundary, so moves are checked against
// [lmin, lmax] per axis (the local range of interior voxels); otherwise every
// move is interior.  A step that starts inside the region lands inside the
// tile, so its hit always goes to the LDS counters.
template <typename T, bool EDGE>
__device__ __forceinline__ void tile_block(const T *__restrict__ img, const WalkGeom &g, const T *simg,
                                           uint32_t *shits, uint32_t *__restrict__ hg, int sorw,
                                           const uint32_t rr[4], float B, float div1, int &step, int &li,
                                           int &lz, int &ly, int &lx, int bz, int by, int bx, int lmin0,
                                           int lmin1, int lmin2, int lmax0, int lmax1, int lmax2,
                                           uint32_t weight, uint32_t &acc)
{
    constexpr int R = TileCfg<T>::C + 2 * TileCfg<T>::H, RI = R + 2;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (step < sorw) {
            const bool in = max(max((unsigned)(lz - 1), (unsigned)(ly - 1)), (unsigned)(lx - 1)) <
                            (unsigned)R;
            float a0, a1, a2, a3, a4, a5;
            if (in) {
                a0 = (float)simg[li + RI * RI];
                a1 = (float)simg[li - RI * RI];
                a2 = (float)simg[li + RI];
                a3 = (float)simg[li - RI];
                a4 = (float)simg[li + 1];
                a5 = (float)simg[li - 1];
            } else {
                const int64_t pos = ((int64_t)(bz + lz) * g.sz + (int64_t)(by + ly) * g.sy) + (bx + lx);
                a0 = bmx_ld(img, pos + g.sz);
                a1 = bmx_ld(img, pos - g.sz);
                a2 = bmx_ld(img, pos + g.sy);
                a3 = bmx_ld(img, pos - g.sy);
                a4 = bmx_ld(img, pos + 1);
                a5 = bmx_ld(img, pos - 1);
            }
            const int d = choose(weight_nonpos(B, a0, div1), weight_nonpos(B, a1, div1),
                                 weight_nonpos(B, a2, div1), weight_nonpos(B, a3, div1),
                                 weight_nonpos(B, a4, div1), weight_nonpos(B, a5, div1), u01(rr[q]));
            const int axis = d >> 1;
            const int sgn = 1 - ((d & 1) << 1);
            bool ok = true;
            if (EDGE) {
                const int c = (axis == 0 ? lz : (axis == 1 ? ly : lx)) + sgn;
                const int lo = axis == 0 ? lmin0 : (axis == 1 ? lmin1 : lmin2);
                const int hi = axis == 0 ? lmax0 : (axis == 1 ? lmax1 : lmax2);
                ok = (c >= lo) & (c <= hi);
            }
            if (ok) {
                lz += axis == 0 ? sgn : 0;
                ly += axis == 1 ? sgn : 0;
                lx += axis == 2 ? sgn : 0;
                const int lmag = axis == 0 ? RI * RI : (axis == 1 ? RI : 1);
                li += sgn * lmag;
                ++acc;
                if (in || max(max((unsigned)lz, (unsigned)ly), (unsigned)lx) < (unsigned)RI) {
                    atomicAdd(&shits[li], weight);
                } else {
                    const int64_t pos =
                        ((int64_t)(bz + lz) * g.sz + (int64_t)(by + ly) * g.sy) + (bx + lx);
                    atomicAdd(&hg[pos], weight);
                }
            }
            ++step;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kTileThreads) void k_tile_walk(
    const T *__restrict__ img, WalkGeom g, TileGeom tg, const bmx_seed *__restrict__ seeds,
    uint4 *__restrict__ st, const uint32_t *__restrict__ bkey, const uint32_t *__restrict__ boff,
    const uint32_t *__restrict__ bcnt, int sorw, int cap, uint32_t weight0, int overrides, uint32_t k0,
    uint32_t k1, uint32_t *__restrict__ hits, unsigned long long *accepted)
{
    constexpr int C = TileCfg<T>::C, H = TileCfg<T>::H, R = C + 2 * H, RI = R + 2;
    constexpr int NT = RI * RI * RI;
    extern __shared__ __align__(16) unsigned char smem[];
    T *simg = (T *)smem;
    uint32_t *shits = (uint32_t *)(smem + tile_img_bytes<T>());
    __shared__ uint32_t next;

    const uint32_t key = bkey[blockIdx.x];
    const uint32_t off = boff[blockIdx.x], cnt = bcnt[blockIdx.x];
    const uint32_t label = key % tg.nlabels;
    const uint32_t tile = key / tg.nlabels;
    const int tx = (int)(tile % tg.ntx), ty = (int)((tile / tg.ntx) % tg.nty),
              tz = (int)(tile / (tg.ntx * tg.nty));
    // global coordinates of tile voxel (0,0,0)
    const int bz = tz * C - H - 1, by = ty * C - H - 1, bx = tx * C - H - 1;

    // Tile load: each thread issues a batch of image loads before storing them to
    // LDS, so one memory latency is paid per batch rather than per element; the
    // hit counters are cleared with 16-byte stores.
    {
        constexpr int BATCH = 8;
        for (int base = 0; base < NT; base += BATCH * kTileThreads) {
            T v[BATCH];
#pragma unroll
            for (int q = 0; q < BATCH; ++q) {
                const int k = base + threadIdx.x + q * kTileThreads;
                if (k < NT) {
                    const int lz = k / (RI * RI), ly = (k / RI) % RI, lx = k % RI;
                    const int gz = min(max(bz + lz, 0), g.Z - 1);
                    const int gy = min(max(by + ly, 0), g.Y - 1);
                    const int gx = min(max(bx + lx, 0), g.X - 1);
                    v[q] = img[(int64_t)gz * g.sz + (int64_t)gy * g.sy + gx];
                }
            }
#pragma unroll
            for (int q = 0; q < BATCH; ++q) {
                const int k = base + threadIdx.x + q * kTileThreads;
                if (k < NT) simg[k] = v[q];
            }
        }
        uint4 *sh4 = (uint4 *)shits;
        for (int k = threadIdx.x; k < NT / 4; k += kTileThreads) sh4[k] = make_uint4(0u, 0u, 0u, 0u);
        for (int k = (NT / 4) * 4 + threadIdx.x; k < NT; k += kTileThreads) shits[k] = 0u;
    }
    if (threadIdx.x == 0) next = 0u;
    __syncthreads();

    // Interior voxels are 1..dim-2; in tile coordinates 1-b..dim-2-b.  A walker
    // is never more than 3 steps outside the region, so a tile whose region
    // grown by 4 stays inside the interior needs no move checks.
    const int lmin0 = 1 - bz, lmin1 = 1 - by, lmin2 = 1 - bx;
    const int lmax0 = g.Z - 2 - bz, lmax1 = g.Y - 2 - by, lmax2 = g.X - 2 - bx;
    const bool edge = (lmin0 > -3) | (lmin1 > -3) | (lmin2 > -3) | (lmax0 < R + 4) |
                      (lmax1 < R + 4) | (lmax2 < R + 4);

    uint32_t *__restrict__ hg = hits + (int64_t)label * g.V;
    uint32_t acc = 0;  // accepted moves of this thread in this launch (< 2^32: the launch is bounded)
    bool have = false;
    uint32_t j = 0, p0 = 0, seed = 0, wl = 0, weight = weight0;
    int step = 0, lz = 0, ly = 0, lx = 0, li = 0, taken = 0;
    float B = 0.0f, div1 = 0.0f;
    while (true) {
        if (!have) {
            j = atomicAdd(&next, 1u);
            if (j >= cnt) break;
            const uint4 w4 = st[off + j];
            const uint32_t pos = w4.x;
            seed = w4.y;
            wl = w4.z;
            step = (int)w4.w;
            const bmx_seed sd = seeds[seed];
            p0 = sd.pos;
            B = bmx_ld(img, (int64_t)p0);
            div1 = 1.0f / (2.0f * sd.var);
            weight = weight0;
            if (overrides && (sd.reserved >> 16) != 0u) weight = weight0 * (sd.reserved >> 16);
            const int z = (int)(pos / g.sz);
            const uint32_t r = pos - (uint32_t)z * g.sz;
            const int y = (int)(r / g.sy);
            const int x = (int)(r - (uint32_t)y * g.sy);
            lz = z - bz;
            ly = y - by;
            lx = x - bx;
            li = (lz * RI + ly) * RI + lx;
            taken = 0;
            have = true;
        }
        const bmx_u4 r4 = bmx_philox(p0, 0u, wl & 0xFFFFu, (uint32_t)(step >> 2), k0, k1);
        const uint32_t rr[4] = {r4.x, r4.y, r4.z, r4.w};
        if (edge)
            tile_block<T, true>(img, g, simg, shits, hg, sorw, rr, B, div1, step, li, lz, ly, lx, bz, by, bx,
                                lmin0, lmin1, lmin2, lmax0, lmax1, lmax2, weight, acc);
        else
            tile_block<T, false>(img, g, simg, shits, hg, sorw, rr, B, div1, step, li, lz, ly, lx, bz, by, bx,
                                 lmin0, lmin1, lmin2, lmax0, lmax1, lmax2, weight, acc);
        taken += 4;
        const bool inside =
            max(max((unsigned)(lz - 1), (unsigned)(ly - 1)), (unsigned)(lx - 1)) < (unsigned)R;
        if (step >= sorw || !inside || taken >= cap) {
            const uint32_t pos = (uint32_t)(bz + lz) * g.sz + (uint32_t)(by + ly) * g.sy + (uint32_t)(bx + lx);
            st[off + j] = make_uint4(pos, seed, wl, (uint32_t)step);
            have = false;
        }
    }
    wave_add_u64(accepted, (uint64_t)acc);
    __syncthreads();
    for (int k = threadIdx.x; k < NT; k += kTileThreads) {
        const uint32_t c = shits[k];
        if (c) {
            const int lz2 = k / (RI * RI), ly2 = (k / RI) % RI, lx2 = k % RI;
            const int64_t gi = (int64_t)(bz + lz2) * g.sz + (int64_t)(by + ly2) * g.sy + (bx + lx2);
            atomicAdd(&hg[gi], c);
        }
    }
}

template <typename T>
int walk_tiled_t(bmx_ctx *ctx, const T *img, const WalkGeom &g, const bmx_seed *seeds, int64_t nseeds,
                 int nlabels, int nbrw, int sorw, int cap, uint32_t weight, int overrides, uint32_t k0,
                 uint32_t k1, uint32_t *hits, unsigned long long *acc, hipEvent_t *ev, int max_timed,
                 int *launches)
{
    constexpr int C = TileCfg<T>::C;
    const int64_t n0 = nseeds * (int64_t)nbrw;
    BMX_CHECK(n0 < ((int64_t)1 << 32), "bmx_walk: %lld walkers exceed 2^32", (long long)n0);
    BMX_CHECK(nbrw <= 65535, "bmx_walk: tiled walker needs nbrw <= 65535");
    TileGeom tg;
    tg.ntz = (uint32_t)((g.Z + C - 1) / C);
    tg.nty = (uint32_t)((g.Y + C - 1) / C);
    tg.ntx = (uint32_t)((g.X + C - 1) / C);
    tg.nlabels = (uint32_t)nlabels;
    const int64_t nkeys = (int64_t)tg.ntz * tg.nty * tg.ntx * nlabels;
    BMX_CHECK(nkeys < ((int64_t)1 << 31), "bmx_walk: too many tiles");
    int rc = grow(ctx, &ctx->d_state, &ctx->state_cap, (size_t)n0 * 2 * sizeof(uint4));
    if (rc) return rc;
    rc = grow(ctx, &ctx->d_keys, &ctx->keys_cap, (size_t)(nkeys + 1) * sizeof(uint32_t));
    if (rc) return rc;
    rc = grow(ctx, &ctx->d_buckets, &ctx->buckets_cap, (size_t)nkeys * 3 * sizeof(uint32_t));
    if (rc) return rc;
    const size_t lds = tile_lds_bytes<T>();
    // Raise the dynamic-LDS limit where the runtime wants it; an error here is
    // left to the launch itself to report.
    if (hipFuncSetAttribute((const void *)k_tile_walk<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
        (void)hipGetLastError();
    uint4 *A = (uint4 *)ctx->d_state, *Bst = A + n0;
    uint32_t *counts = (uint32_t *)ctx->d_keys;
    uint32_t *bkey = (uint32_t *)ctx->d_buckets, *boff = bkey + nkeys, *bcnt = boff + nkeys;
    uint32_t *d_nb = (uint32_t *)(ctx->d_counter + 1);
    k_tile_init<<<bmx_blocks(n0, 256), 256, 0, ctx->stream>>>(seeds, nseeds, nbrw, sorw, overrides, A);
    BMX_LAUNCH_CHECK();
    int64_t n = n0;
    int round = 0;
    // BMX_WALK_TRACE=1 prints one line per round (live walkers, buckets, kernel ms) to stderr.
    static const bool trace = getenv("BMX_WALK_TRACE") && getenv("BMX_WALK_TRACE")[0] == '1';
    while (n > 0) {
        BMX_HIP(hipMemsetAsync(counts, 0, (size_t)(nkeys + 1) * sizeof(uint32_t), ctx->stream));
        BMX_HIP(hipMemsetAsync(d_nb, 0, sizeof(uint32_t), ctx->stream));
        k_tile_hist<C><<<bmx_blocks(n, 256), 256, 0, ctx->stream>>>(A, n, g, tg, sorw, counts);
        BMX_LAUNCH_CHECK();
        rc = scan_u32(ctx, counts, nkeys + 1);
        if (rc) return rc;
        k_bucket_list<<<bmx_blocks(nkeys, 256), 256, 0, ctx->stream>>>(counts, nkeys, bkey, boff, bcnt,
                                                                       d_nb);
        BMX_LAUNCH_CHECK();
        k_tile_scatter<C><<<bmx_blocks(n, 256), 256, 0, ctx->stream>>>(A, n, Bst, counts, g, tg, sorw);
        BMX_LAUNCH_CHECK();
        uint32_t host[2] = {0, 0};
        BMX_HIP(hipMemcpyAsync(&host[0], counts + nkeys, sizeof(uint32_t), hipMemcpyDeviceToHost,
                               ctx->stream));
        BMX_HIP(hipMemcpyAsync(&host[1], d_nb, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        BMX_HIP(hipStreamSynchronize(ctx->stream));
        const int64_t live = host[0];
        const uint32_t nb = host[1];
        if (live == 0) break;
        if (round < max_timed) BMX_HIP(hipEventRecord(ev[2 * round], ctx->stream));
        k_tile_walk<T><<<nb, kTileThreads, lds, ctx->stream>>>(img, g, tg, seeds, Bst, bkey, boff, bcnt,
                                                              sorw, cap, weight, overrides, k0, k1, hits,
                                                              acc);
        BMX_LAUNCH_CHECK();
        if (round < max_timed) BMX_HIP(hipEventRecord(ev[2 * round + 1], ctx->stream));
        if (trace && round < max_timed) {
            float ms = 0.0f;
            BMX_HIP(hipEventSynchronize(ev[2 * round + 1]));
            BMX_HIP(hipEventElapsedTime(&ms, ev[2 * round], ev[2 * round + 1]));
            fprintf(stderr, "[bmx_walk] round %d live %lld buckets %u walkers/bucket %.0f kernel %.3f ms\n", round,
                    (long long)live, nb, (double)live / (nb ? nb : 1), ms);
        }
        uint4 *t = A; A = Bst; Bst = t;
        n = live;
        ++round;
        BMX_CHECK(round <= sorw / 4 + 2, "bmx_walk: tiled walker did not converge");
    }
    *launches = round < max_timed ? round : max_timed;
    return BMX_OK;
}

}  // namespace

extern "C" int bmx_seed_walks(bmx_ctx *ctx, bmx_seed *seeds, int64_t n, int walks, int weight)
{
    int rc = bmx_set_device(ctx);
    if (rc) return rc;
    BMX_CHECK(n == 0 || seeds, "bmx_seed_walks: null seeds");
    BMX_CHECK(n >= 0 && n < ((int64_t)1 << 32) - 256, "bmx_seed_walks: n out of range");
    BMX_CHECK(walks >= 0 && walks <= 65535 && weight >= 0 && weight <= 65535,
              "bmx_seed_walks: walks / weight not in [0, 65535]");
    if (n > 0) {
        k_seed_walks<<<bmx_blocks(n, 256), 256, 0, ctx->stream>>>(seeds, n,
                                                                 (uint32_t)walks | ((uint32_t)weight << 16));
        BMX_LAUNCH_CHECK();
    }
    return BMX_OK;
}

// Entry from bmx_walk_ex (bmx_walk.hip).
int bmx_walk_tiled(bmx_ctx *ctx, const void *img, int img_dtype, int Z, int Y, int X,
                   const bmx_seed *seeds, int64_t nseeds, int nlabels, int nbrw, int sorw, int cap,
                   uint32_t weight, int overrides, uint32_t k0, uint32_t k1, uint32_t *hits,
                   unsigned long long *acc, hipEvent_t *ev, int max_timed, int *launches)
{
    WalkGeom g{Z, Y, X, (uint32_t)X, (uint32_t)X * (uint32_t)Y, (int64_t)Z * Y * X};
    if (img_dtype == BMX_IMG_U8)
        return walk_tiled_t<uint8_t>(ctx, (const uint8_t *)img, g, seeds, nseeds, nlabels, nbrw, sorw, cap,
                                     weight, overrides, k0, k1, hits, acc, ev, max_timed, launches);
    return walk_tiled_t<float>(ctx, (const float *)img, g, seeds, nseeds, nlabels, nbrw, sorw, cap, weight,
                               overrides, k0, k1, hits, acc, ev, max_timed, launches);
}
