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
// image and hit counters, both with one more voxel of halo, in LDS and runs
// every walker of the bucket - walkers whose current voxel lies in the core -
// until the walker steps out of the region, finishes its walk, or has made
// `cap` steps in this round (or, once the bucket is handed out, its wave is
// down to half its lanes: the drain pause).  Every step starts in the region, so its
// neighbour reads and its hit stay in LDS.  At the end the workgroup adds
// each non-zero LDS counter to the hit volume with one global atomic.
// Between rounds the walkers are counting-sorted by (core tile, label) again.
// A walker may pause inside a Philox block; step s always uses word s & 3 of
// block s >> 2, wherever the step is taken.
#include "bmx_walk_common.h"

#include <algorithm>
#include <vector>

namespace {

// Tile shapes: core side C, halo H, workgroup size NTH, bytes per LDS hit
// counter HB and the waves per SIMD the kernel is built for (WPE).
// BMX_TILE_SHAPES of them are compiled; bmx_walk_opts.tile_shape picks one
// (the walk autotuner times them).
//  - HB = 4: one uint32 counter per tile voxel.  A u8 16/6 tile of this kind
//    takes 135 KB, so one workgroup per CU.
//  - HB = 2: two 16-bit counters per 32-bit LDS word (voxels 2k and 2k+1 of
//    the tile).  The u8 16/6 tile then takes 81 KB and two workgroups share a
//    CU: one bucket's last walkers, its tile load and its flush overlap the
//    other workgroup's walk.  Counter wraps are accounted for exactly (see
//    hit16_wrap); a walk whose per-hit weight can reach 2^16 takes an HB = 4
//    shape instead (Hit32Sibling, bmx_walk_tiled).
// Shape 0 is the default.  u8: 16/6 x 768 threads with 16-bit counters (the
// 512^3 plain walk kernel 4.49 s -> 4.08 s against shape 1, the guided walk
// 1.10 s -> 0.94 s).  float32 (4-byte image voxels): 16/4 x 1024 with 32-bit
// counters; the two-per-CU float tiles need a halo of 2-3 and were not faster
// (384^3, 12-bit image).  The smaller cores trade fewer voxels per round for
// more resident workgroups, which can pay on small volumes.
template <typename T, int S> struct TileShape;
template <> struct TileShape<uint8_t, 0> { static constexpr int C = 16, H = 6, NTH = 768, HB = 2, WPE = 6; };
template <> struct TileShape<uint8_t, 1> { static constexpr int C = 16, H = 6, NTH = 1024, HB = 4, WPE = 4; };
template <> struct TileShape<uint8_t, 2> { static constexpr int C = 16, H = 4, NTH = 512, HB = 2, WPE = 6; };
template <> struct TileShape<uint8_t, 3> { static constexpr int C = 12, H = 5, NTH = 1024, HB = 4, WPE = 4; };
template <> struct TileShape<uint8_t, 4> { static constexpr int C = 8, H = 4, NTH = 512, HB = 4, WPE = 2; };
template <> struct TileShape<uint8_t, 5> { static constexpr int C = 16, H = 6, NTH = 1024, HB = 2, WPE = 8; };
template <> struct TileShape<float, 0> { static constexpr int C = 16, H = 4, NTH = 1024, HB = 4, WPE = 4; };
template <> struct TileShape<float, 1> { static constexpr int C = 14, H = 3, NTH = 768, HB = 2, WPE = 6; };
template <> struct TileShape<float, 2> { static constexpr int C = 16, H = 2, NTH = 768, HB = 2, WPE = 6; };
template <> struct TileShape<float, 3> { static constexpr int C = 12, H = 3, NTH = 1024, HB = 4, WPE = 4; };
template <> struct TileShape<float, 4> { static constexpr int C = 8, H = 3, NTH = 512, HB = 4, WPE = 2; };
template <> struct TileShape<float, 5> { static constexpr int C = 12, H = 3, NTH = 512, HB = 4, WPE = 2; };
static_assert(BMX_TILE_SHAPES == 6, "one TileShape per shape index");

template <typename T, int S> constexpr int tile_RI() { return TileShape<T, S>::C + 2 * TileShape<T, S>::H + 2; }
template <typename T, int S> constexpr size_t tile_img_bytes()
{
    return ((size_t)tile_RI<T, S>() * tile_RI<T, S>() * tile_RI<T, S>() * sizeof(T) + 15) & ~(size_t)15;
}
template <typename T, int S> constexpr size_t tile_hit_bytes()
{
    return (size_t)tile_RI<T, S>() * tile_RI<T, S>() * tile_RI<T, S>() * TileShape<T, S>::HB;
}
template <typename T, int S> constexpr size_t tile_lds_bytes() { return tile_img_bytes<T, S>() + tile_hit_bytes<T, S>(); }

// A walk whose hits add at most this much per accepted move may use 16-bit
// LDS counters: one add then wraps a counter at most once.
constexpr uint32_t kHit16MaxWeight = 0xFFFFu;

// An HB = 4 shape to take instead of shape S when the walk's hit weight may
// reach 2^16 (S itself when S has HB = 4).
template <typename T, int S> struct Hit32Sibling {
    static constexpr int value = TileShape<T, S>::HB == 4 ? S : (sizeof(T) == 1 ? 1 : 0);
};
static_assert(TileShape<uint8_t, 1>::HB == 4 && TileShape<float, 0>::HB == 4, "the 32-bit counter fallbacks");

// One volume of a walk (a batch walks many at once): voxels [base, base + V)
// of the concatenated image, z-major with x fastest; tiles [tile_base,
// tile_base + ntz*nty*ntx) of the batch's tile numbering; seeds [seed_base,
// next volume's seed_base) of the seed array, with volume-local positions; its
// hit block [labels][V] at element offset hits_base.  Walker positions are
// batch-global (base + local position).
struct VolGeom {
    uint32_t base;
    int Z, Y, X;
    uint32_t sy, sz, V;
    uint32_t tile_base, nty, ntx;
    uint32_t seed_base, pad;
    uint64_t hits_base;
    uint32_t mz, my;     // multipliers of the exact divisions by sz and sy (div_u32)
    uint32_t shz, shy;   // their shifts
};

// Exact quotient n / d for every 32-bit n, d >= 2, from the multiplier m and
// shift sh of magic_u32 (Granlund and Montgomery, "Division by invariant
// integers using multiplication", PLDI 1994, Figure 4.1: the multiplier is
// 2^32 + m, applied as hi + ((n - hi) >> 1) so that no sum passes 32 bits).
__device__ __forceinline__ uint32_t div_u32(uint32_t n, uint32_t m, uint32_t sh)
{
    const uint32_t hi = __umulhi(n, m);
    return (hi + ((n - hi) >> 1)) >> sh;
}
static void magic_u32(uint32_t d, uint32_t *m, uint32_t *sh)
{
    int l = 0;
    while (((uint64_t)1 << l) < d) ++l;   // l = ceil(log2 d), 1 <= l <= 32 for d >= 2
    *m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << l) - d)) / d + 1);
    *sh = (uint32_t)(l - 1);
}

// Per-seed constants of a walk, built once per walk (k_walk_seeds): the seed's
// volume-local position (the key of its Philox streams), its image value B,
// div1 = 1 / (2 var) and the weight its walkers add per hit.
struct WalkSeed {
    uint32_t pos;
    float B, div1;
    uint32_t weight;
};

struct TileGeom {
    const VolGeom *vols;
    int nvol;
    uint32_t nlabels;   // label slots per tile key (the largest label count of the batch)
};

// Last volume whose first voxel (resp. first tile) is <= pos (resp. tile).
__device__ __forceinline__ int vol_of_pos(const VolGeom *__restrict__ vols, int nvol, uint32_t pos)
{
    int lo = 0, hi = nvol - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (vols[mid].base <= pos) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ int vol_of_seed(const VolGeom *__restrict__ vols, int nvol, uint32_t seed)
{
    int lo = 0, hi = nvol - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (vols[mid].seed_base <= seed) lo = mid; else hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ int vol_of_tile(const VolGeom *__restrict__ vols, int nvol, uint32_t tile)
{
    int lo = 0, hi = nvol - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (vols[mid].tile_base <= tile) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <int C>
__device__ __forceinline__ uint32_t tile_key(uint32_t pos, uint32_t label, const TileGeom &tg)
{
    const VolGeom &g = tg.vols[vol_of_pos(tg.vols, tg.nvol, pos)];
    const uint32_t r0 = pos - g.base;
    const uint32_t z = div_u32(r0, g.mz, g.shz);
    const uint32_t r = r0 - z * g.sz;
    const uint32_t y = div_u32(r, g.my, g.shy);
    const uint32_t x = r - y * g.sy;
    return (g.tile_base + ((z / C) * g.nty + y / C) * g.ntx + x / C) * tg.nlabels + label;
}

// Walker state: {batch-global position, seed index, walk | label << 16, step},
// one per (seed, walk), seed-major.  With overrides a seed running fewer than
// nbrw walks gets its extra walkers created finished (step = sorw), so the
// first re-bucketing drops them.
__global__ __launch_bounds__(256) void k_tile_init(const bmx_seed *__restrict__ seeds, int64_t nseeds,
                                                   int nbrw, int sorw, int overrides, TileGeom tg,
                                                   uint4 *__restrict__ st)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseeds * nbrw) return;
    const int64_t s = i / nbrw;
    const bmx_seed sd = seeds[s];
    const uint32_t base = tg.vols[vol_of_seed(tg.vols, tg.nvol, (uint32_t)s)].base;
    const uint32_t w = (uint32_t)(i - s * nbrw);
    const uint32_t walks = overrides ? (sd.reserved & 0xFFFFu) : 0u;
    const uint32_t step0 = (walks != 0u && w >= walks) ? (uint32_t)sorw : 0u;
    st[i] = make_uint4(base + sd.pos, (uint32_t)s, w | (sd.label << 16), step0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_walk_seeds(const T *__restrict__ img_all, TileGeom tg,
                                                    const bmx_seed *__restrict__ seeds, int64_t nseeds,
                                                    uint32_t weight0, int overrides, WalkSeed *__restrict__ ws)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseeds) return;
    const bmx_seed sd = seeds[s];
    const uint32_t base = tg.vols[vol_of_seed(tg.vols, tg.nvol, (uint32_t)s)].base;
    uint32_t weight = weight0;
    if (overrides && (sd.reserved >> 16) != 0u) weight = weight0 * (sd.reserved >> 16);
    ws[s] = WalkSeed{sd.pos, bmx_ld(img_all + base, (int64_t)sd.pos), 1.0f / (2.0f * sd.var), weight};
}

__global__ __launch_bounds__(256) void k_seed_walks(bmx_seed *__restrict__ seeds, int64_t n, uint32_t value)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) seeds[i].reserved = value;
}

// Round bookkeeping in device memory, so that a round needs no host values
// and a sequence of rounds can be captured in one graph: n = walkers in the
// current state array, nnext = live walkers among them (the key-count total),
// nb = non-empty buckets, bnext = next bucket to walk.  The bucket list holds
// the nheavy buckets of at least one walker per thread at its front and the
// nlight others at its back (nb = nheavy + nlight), so a round starts its
// longest buckets first and ends on short ones.  With a tiny threshold,
// buckets below it go to a list of their own (ntiny of them, taken from
// tnext), walked without the LDS tile.
struct RoundCtl {
    uint32_t nb, bnext, n, done, nheavy, nlight, ntiny, tnext, nnext;
    // persistent walk kernels: first start and last end of the current round
    // (device wall clock), and the summed spans of the rounds so far
    unsigned long long t0, t1, ticks;
};

// Where a walker goes in the next round's state array, written by whatever
// moved it last (k_keys_first before the first round, the walk kernel after):
//  {kDestDead, 0}: its walk has ended, it is dropped;
//  {kDestDirect | key, slot}: entry `slot` of key's range;
//  {bi, rank}: the rank-th walker of round bucket bi to stop in the bucket's
//    neighbour tile nb (0..26, kept in bits 24-28 of the walker's walk word);
//    the bucket reserved that neighbour's entries at bres[27 bi + nb] =
//    {key, first entry within the key's range}.
// The key counts are added by the same atomics, so no pass over the walkers
// counts keys and the placement needs no atomics.
constexpr uint32_t kDestDead = 0xFFFFFFFFu, kDestDirect = 0x80000000u;

// Slot in the bucket arrays of the round's bucket i (heavy ones from the front,
// light ones from the back).
__device__ __forceinline__ uint32_t bucket_slot(uint32_t i, uint32_t nheavy, uint32_t nbcap)
{
    return i < nheavy ? i : nbcap - 1u - (i - nheavy);
}

// Start of a walk: n walkers in the state array, nothing listed yet.
__global__ void k_set_count(RoundCtl *__restrict__ ctl, uint32_t n)
{
    if (threadIdx.x == 0) {
        ctl->n = n;
        ctl->nnext = n;
        ctl->nb = 0u;
        ctl->bnext = 0u;
        ctl->nheavy = 0u;
        ctl->nlight = 0u;
        ctl->ntiny = 0u;
        ctl->tnext = 0u;
        ctl->done = 0u;
        ctl->t0 = ~0ull;
        ctl->t1 = 0ull;
        ctl->ticks = 0ull;
    }
}

__global__ __launch_bounds__(256) void k_clear_counts(uint32_t *__restrict__ counts, int64_t ncounts)
{
    for (int64_t i = bmx_gtid(); i < ncounts; i += bmx_gstride()) counts[i] = 0u;
}

// Keys and destinations of the walkers as created (before the first round).
// The loop bound is uniform per block, as agg_atomic_inc needs every lane of a
// wave.
template <int C>
__global__ __launch_bounds__(256) void k_keys_first(const uint4 *__restrict__ st, const RoundCtl *__restrict__ ctl,
                                                    TileGeom tg, int sorw, uint32_t *counts,
                                                    uint2 *__restrict__ dest)
{
    const int64_t n = ctl->n;
    for (int64_t b0 = (int64_t)blockIdx.x * 256; b0 < n; b0 += (int64_t)gridDim.x * 256) {
        const int64_t i = b0 + threadIdx.x;
        bool live = false;
        uint32_t key = 0;
        if (i < n) {
            const uint4 w4 = st[i];
            live = (int)w4.w < sorw;
            key = tile_key<C>(w4.x, (w4.z >> 16) & 0xFFu, tg);
        }
        const uint32_t slot = agg_atomic_inc(counts, key, live);
        if (i < n) dest[i] = live ? make_uint2(kDestDirect | key, slot) : make_uint2(kDestDead, 0u);
    }
}

// After the scan of the key counts: the live total, and the bucket counters
// cleared for k_bucket_list.
__global__ void k_round_mid(const uint32_t *__restrict__ off, int64_t nkeys, RoundCtl *__restrict__ ctl)
{
    if (threadIdx.x == 0) {
        ctl->nnext = off[nkeys];
        ctl->nheavy = 0u;
        ctl->nlight = 0u;
        ctl->ntiny = 0u;
    }
}

// Each live walker of `in` (ctl->n entries) to its entry of `out`: the
// exclusive offset of its key plus its place in the key's range (dest).
__global__ __launch_bounds__(256) void k_tile_place(const uint4 *__restrict__ in, const uint2 *__restrict__ dest,
                                                    const RoundCtl *__restrict__ ctl, const uint32_t *__restrict__ off,
                                                    const uint2 *__restrict__ bres, uint4 *__restrict__ out)
{
    const int64_t n = ctl->n;
    for (int64_t i = bmx_gtid(); i < n; i += bmx_gstride()) {
        const uint2 d = dest[i];
        if (d.x == kDestDead) continue;
        uint4 w4 = in[i];
        uint32_t at;
        if (d.x & kDestDirect) {
            at = off[d.x & ~kDestDirect] + d.y;
        } else {
            const uint2 r = bres[(int64_t)d.x * 27 + ((w4.z >> 24) & 31u)];
            at = off[r.x] + r.y + d.y;
            w4.z &= 0x00FFFFFFu;
        }
        out[at] = w4;
    }
}

// Start of a round's walk: the placed walkers are the state array, the counts
// are cleared for the walk to fill, and the bucket counters reset.
__global__ __launch_bounds__(256) void k_round_start(uint32_t *__restrict__ counts, int64_t ncounts,
                                                     RoundCtl *__restrict__ ctl)
{
    for (int64_t i = bmx_gtid(); i < ncounts; i += bmx_gstride()) counts[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->n = ctl->nnext;
        ctl->nb = ctl->nheavy + ctl->nlight;
        ctl->bnext = 0u;
        ctl->tnext = 0u;
    }
}

// A lane's `want` consecutive entries of the counter's range, for every lane
// of the wave at once (one atomic per wave): the first one's index.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t *ctr, uint32_t want)
{
    const int lane = (int)(threadIdx.x & 63u);
    uint32_t incl = want;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t v = __shfl_up(incl, (unsigned)o, 64);
        if (lane >= o) incl += v;
    }
    const uint32_t total = __shfl(incl, 63, 64);
    uint32_t base = 0;
    if (lane == 63 && total) base = atomicAdd(ctr, total);
    return __shfl(base, 63, 64) + incl - want;
}

// Keys whose walkers a round splits over several buckets: a bucket holds at
// most max(kSplitMin, live / (2 * grid)) walkers, so no bucket holds more than
// about half a resident workgroup's share of the round.  A few tiles can hold
// most of a sparse walk's walkers (the repair walk of the guided method); one
// workgroup per (tile, label) would then walk them alone while the others
// wait.  The buckets of a key stage the same tile and add their own hits, so
// the sums are unchanged; each has its own destination reservations.
constexpr uint32_t kSplitMin = 1024;

// Extra bucket-list entries the splits can need beyond one per key: a key of
// c walkers takes ceil(c / S) <= c / S + 1 entries, and sum(c) / S <= 2 grid.
__host__ __device__ constexpr int64_t split_extra(int64_t grid) { return 2 * grid + 1; }

// Non-empty buckets of the exclusive offsets off[0..nkeys] -> (key, offset,
// count): buckets of at least `heavy` walkers from the front of the arrays,
// the others from the back of their nbcap entries (nheavy + nlight <= nbcap,
// so the two never meet); buckets of fewer than `tiny` walkers (tiny = 0:
// none) into the tiny arrays tkey/toff/tcnt instead.  A key of more than S
// walkers (see kSplitMin; `grid` resident workgroups) is listed as
// ceil(count / S) heavy buckets of at most S walkers each.
__global__ __launch_bounds__(256) void k_bucket_list(const uint32_t *__restrict__ off, int64_t nkeys,
                                                     uint32_t nbcap, uint32_t *bkey, uint32_t *boff,
                                                     uint32_t *bcnt, uint32_t *tkey, uint32_t *toff,
                                                     uint32_t *tcnt, RoundCtl *ctl, uint32_t heavy, uint32_t tiny,
                                                     uint32_t grid)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    uint32_t c = 0, o = 0;
    if (k < nkeys) {
        o = off[k];
        c = off[k + 1] - o;
    }
    const uint32_t live = ctl->nnext;
    const uint32_t share = (uint32_t)(((uint64_t)live + 2ull * grid - 1ull) / (2ull * grid));
    const uint32_t S = max(kSplitMin, share);
    const bool is_tiny = c < tiny;
    const bool is_heavy = c >= heavy;
    const uint32_t m = c > 0 && is_heavy ? (c + S - 1u) / S : 0u;   // heavy buckets of this key
    const uint32_t ih = wave_reserve(&ctl->nheavy, m);
    const uint32_t il = agg_atomic_inc(&ctl->nlight, 0u, c > 0 && !is_heavy && !is_tiny);
    const uint32_t it = agg_atomic_inc(&ctl->ntiny, 0u, c > 0 && is_tiny);
    if (c > 0 && is_tiny) {
        tkey[it] = (uint32_t)k;
        toff[it] = o;
        tcnt[it] = c;
    } else if (c > 0 && is_heavy) {
        for (uint32_t i = 0; i < m; ++i) {
            bkey[ih + i] = (uint32_t)k;
            boff[ih + i] = o + i * S;
            bcnt[ih + i] = min(S, c - i * S);
        }
    } else if (c > 0) {
        const uint32_t idx = nbcap - 1u - il;
        bkey[idx] = (uint32_t)k;
        boff[idx] = o;
        bcnt[idx] = c;
    }
}

// Buckets of fewer walkers than this are walked without the LDS tile in the
// graph-driven rounds: a wave per bucket, a lane per walker, image reads and
// hits in global memory.  Staging a whole tile for a few dozen walkers leaves
// the workgroup waiting on one wave's dependent steps; walking from global
// memory costs a device-scope atomic per step, whose rate is the device's
// (a wave that took several tiny buckets at once was slower, not faster), so
// the threshold stays low.
constexpr uint32_t kTinyWalkers = 32;

// Weight exp(-(B-A)^2 * div1) of the reference (pyopencl_small.py:187-190).
__device__ __forceinline__ float weight_nonpos(float B, float A, float div1)
{
    return bmx_weight(B, A, div1);
}

// A 16-bit LDS hit counter wrapped.  LDS word k holds the counters of tile
// voxels 2k (bits 0-15, "low") and 2k+1 (bits 16-31, "high"), and every
// change to a word is an atomic add of w (low) or w << 16 (high), w < 2^16,
// so the word always equals low_total + 2^16 * high_total mod 2^32.  The
// add that wraps a field is known from the value the atomic returns:
//  - a low add wraps when old_low + w >= 2^16: the low total has gained
//    2^16, and its carry has added 1 to the high field, which itself wraps
//    out of the word when old_high == 0xFFFF;
//  - a high add wraps when old_high + w >= 2^16 (the high total gains 2^16).
// Each wrap is seen by exactly one add, whatever the order of the atomics, so
// the lane that sees it settles it in the global hits at once: +2^16 for the
// field that wrapped, and for a low wrap -1 on the high voxel (its carry) plus
// 2^16 when that carry wrapped the high field.  The flush then adds the two
// fields as they stand.  Global counts are sums mod 2^32, so the order of
// these adds does not matter.  A word changes only when a walker lands on
// one of its voxels, an interior voxel, and voxels 2k and 2k+1 lie in the
// same row of the tile (its side is even): both are voxels of the volume.
// gi: global index of the voxel the walker landed on.
__device__ __noinline__ void hit16_wrap(uint32_t *hg, int64_t gi, bool high, uint32_t old)
{
    atomicAdd(&hg[gi], 0x10000u);
    if (!high) atomicAdd(&hg[gi + 1], (old >> 16) == 0xFFFFu ? 0xFFFFu : 0xFFFFFFFFu);
}

// Tile coordinates packed in one word, P = lz << 20 | ly << 10 | lx: 10-bit
// fields, every tile coordinate below 32, so a field can be tested with a
// carry into (or a borrow out of) its bit 5 without disturbing its neighbours.
constexpr uint32_t kP1 = (1u << 20) | (1u << 10) | 1u;   // 1 in every field
constexpr uint32_t kPBit5 = kP1 << 5;                    // bit 5 of every field
__device__ __forceinline__ uint32_t ppack(int z, int y, int x)
{
    return ((uint32_t)z << 20) | ((uint32_t)y << 10) | (uint32_t)x;
}
__device__ __forceinline__ int pz(uint32_t P) { return (int)(P >> 20); }
__device__ __forceinline__ int py(uint32_t P) { return (int)((P >> 10) & 1023u); }
__device__ __forceinline__ int px(uint32_t P) { return (int)(P & 1023u); }
// Every field of P is in [1, R] (fields in [0, R + 1] <= 31): a field f = R + 1
// sets bit 5 of f + 31 - R, and f = 0 borrows in P - kP1, which sets its bit
// 5 (the borrow may flag the next field too, which changes nothing).
template <int R>
__device__ __forceinline__ bool p_inside(uint32_t P)
{
    static_assert(R + 1 <= 31, "packed tile coordinates are below 32");
    return (((P + (uint32_t)(31 - R) * kP1) | (P - kP1)) & kPBit5) == 0u;
}
// Every field of P (in [0, 31]) is in [lo, hi] of LO and HI (fields in
// [0, 31]): a field below lo borrows in P - LO, one above hi borrows in HI - P.
__device__ __forceinline__ bool p_within(uint32_t P, uint32_t LO, uint32_t HI)
{
    return (((P - LO) | (HI - P)) & kPBit5) == 0u;
}

// The steps of one Philox block (4 steps) that one walker takes in this
// tile.  Local coordinates (lz,ly,lx), packed in P, address the LDS tile, whose voxel
// (1,1,1) is the region origin; the region proper is [1,R]^3 and the tile adds
// one voxel on each side for neighbour reads.  lq is the walker's index in the
// tile minus RI^2 (the index of its -z neighbour), so that its six neighbours
// and its counter are at non-negative constant offsets from lq (LDS
// instructions take unsigned offsets).  Step s uses word s & 3 of block s >> 2, so a walker whose visit began
// or ended inside a block resumes at the right word.  A step is taken only
// from a voxel of the region: its neighbours and its landing voxel are in the
// tile, so every read and hit stays in LDS; a walker that steps out of the
// region stops, and the next round continues it from the tile that owns its
// voxel.  EDGE: the tile may reach the volume boundary, so moves are checked
// against the local range of interior voxels (LO, HI: packed per axis,
// clamped to [0, 31]; the walker's other coordinates are interior already);
// otherwise every move is interior.
template <typename T, int S, bool EDGE>
__device__ __forceinline__ void tile_block(const T *simg, uint32_t *shits, int sorw, const uint32_t rr[4],
                                           float B, float div1, int &step, int &lq, uint32_t &P, uint32_t LO,
                                           uint32_t HI, uint32_t weight, uint32_t &acc, uint32_t *hg, int64_t g0,
                                           uint32_t gsz, uint32_t gsy)
{
    constexpr int R = TileShape<T, S>::C + 2 * TileShape<T, S>::H, RI = R + 2;
    static_assert(TileShape<T, S>::HB == 4 || RI % 2 == 0, "16-bit counters pair voxels of one tile row");
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (p_inside<R>(P) && step < sorw && (step & 3) == q) {
            const float a0 = (float)simg[lq + 2 * RI * RI];
            const float a1 = (float)simg[lq];
            const float a2 = (float)simg[lq + RI * RI + RI];
            const float a3 = (float)simg[lq + RI * RI - RI];
            const float a4 = (float)simg[lq + RI * RI + 1];
            const float a5 = (float)simg[lq + RI * RI - 1];
            const int d = choose(weight_nonpos(B, a0, div1), weight_nonpos(B, a1, div1),
                                 weight_nonpos(B, a2, div1), weight_nonpos(B, a3, div1),
                                 weight_nonpos(B, a4, div1), weight_nonpos(B, a5, div1), u01(rr[q]));
            const int axis = d >> 1;
            const int sgn = 1 - ((d & 1) << 1);
            const uint32_t pmag = axis == 0 ? (1u << 20) : (axis == 1 ? (1u << 10) : 1u);
            const uint32_t nP = P + (uint32_t)(sgn * (int)pmag);
            if (!EDGE || p_within(nP, LO, HI)) {
                P = nP;
                const int lmag = axis == 0 ? RI * RI : (axis == 1 ? RI : 1);
                lq += sgn * lmag;
                ++acc;
                if (TileShape<T, S>::HB == 4) {
                    atomicAdd(&shits[lq + RI * RI], weight);
                } else {
                    // counter index lq + RI^2 (RI^2 is even: same parity as lq;
                    // lq may be negative in the tile's rim, >> floors)
                    const uint32_t sh = ((uint32_t)lq & 1u) << 4;
                    const uint32_t old = atomicAdd(&shits[(lq >> 1) + RI * RI / 2], weight << sh);
                    if (__builtin_amdgcn_ubfe(old, sh, 16u) > 0xFFFFu - weight)
                        hit16_wrap(hg, g0 + (int64_t)pz(P) * gsz + (int64_t)py(P) * gsy + px(P), sh != 0u, old);
                }
            }
            ++step;
        }
    }
}

// One walker of a tiny bucket, walked from global memory for up to `cap` steps
// (or to the end of its walk): the transition law, Philox stream, interior
// rule and hit weights of tile_block, so the hits it adds are the ones the tile
// walker would add.  w4: its state, written back by the caller.
template <typename T>
__device__ __forceinline__ void tiny_walker(const T *__restrict__ img, const WalkGeom &g, const VolGeom &gv,
                                            const WalkSeed *__restrict__ wseeds, uint4 &w4, int sorw, int cap,
                                            uint32_t k0, uint32_t k1, uint32_t *__restrict__ hg, uint32_t &acc)
{
    const uint32_t gbase = gv.base;
    uint32_t pos = w4.x - gbase;
    const uint32_t seed = w4.y, wl = w4.z;
    int step = (int)w4.w;
    const WalkSeed ws = wseeds[seed];
    const uint32_t p0 = ws.pos, weight = ws.weight;
    const float B = ws.B, div1 = ws.div1;
    int z = (int)div_u32(pos, gv.mz, gv.shz);
    const uint32_t rr0 = pos - (uint32_t)z * g.sz;
    int y = (int)div_u32(rr0, gv.my, gv.shy);
    int x = (int)(rr0 - (uint32_t)y * g.sy);
    // interior bounds and index strides as scalars (a select over struct fields
    // would be lowered to an indexed scratch load)
    const int zmax = g.Z - 2, ymax = g.Y - 2, xmax = g.X - 2;
    const uint32_t sz = g.sz, sy = g.sy;
    const int step0 = step;
    while (step - step0 < cap && step < sorw) {
        const bmx_u4 r4 = bmx_philox(p0, 0u, wl & 0xFFFFu, (uint32_t)(step >> 2), k0, k1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (step < sorw && (step & 3) == q) {
                const uint32_t word = q == 0 ? r4.x : (q == 1 ? r4.y : (q == 2 ? r4.z : r4.w));
                const int64_t i = pos;
                const int d = choose(weight_nonpos(B, bmx_ld(img, i + sz), div1),
                                     weight_nonpos(B, bmx_ld(img, i - sz), div1),
                                     weight_nonpos(B, bmx_ld(img, i + sy), div1),
                                     weight_nonpos(B, bmx_ld(img, i - sy), div1),
                                     weight_nonpos(B, bmx_ld(img, i + 1), div1),
                                     weight_nonpos(B, bmx_ld(img, i - 1), div1), u01(word));
                const int axis = d >> 1;
                const int sgn = 1 - ((d & 1) << 1);
                const int c = (axis == 0 ? z : (axis == 1 ? y : x)) + sgn;
                const int hi = axis == 0 ? zmax : (axis == 1 ? ymax : xmax);
                if (c >= 1 && c <= hi) {
                    z += axis == 0 ? sgn : 0;
                    y += axis == 1 ? sgn : 0;
                    x += axis == 2 ? sgn : 0;
                    pos += (uint32_t)sgn * (axis == 0 ? sz : (axis == 1 ? sy : 1u));
                    atomicAdd(&hg[pos], weight);
                    ++acc;
                }
                ++step;
            }
        }
    }
    w4 = make_uint4(gbase + pos, seed, wl, (uint32_t)step);
}

// Walks the round's buckets as described at the top of this file.  PERSIST =
// false: workgroup b walks bucket b (the grid is ctl->nb).  PERSIST = true: a
// grid of resident workgroups takes buckets from ctl->bnext until all ctl->nb
// are taken (a fixed grid, as a captured graph needs).  The loop keeps more
// values live across buckets (about 80 VGPRs against 44), which costs nothing
// where LDS already limits a CU to one workgroup.
template <typename T, int S, bool PERSIST>
__global__ __launch_bounds__((TileShape<T, S>::NTH), (TileShape<T, S>::WPE)) void k_tile_walk(
    const T *__restrict__ img_all, TileGeom tg, const WalkSeed *__restrict__ wseeds,
    uint4 *__restrict__ st, const uint32_t *__restrict__ bkey, const uint32_t *__restrict__ boff,
    const uint32_t *__restrict__ bcnt, const uint32_t *__restrict__ tkey, const uint32_t *__restrict__ toff,
    const uint32_t *__restrict__ tcnt, RoundCtl *__restrict__ ctl, uint32_t nbcap, int sorw, int cap,
    uint32_t k0, uint32_t k1, uint32_t *__restrict__ hits, unsigned long long *accepted,
    uint2 *__restrict__ dest, uint2 *__restrict__ bres, uint32_t *__restrict__ counts)
{
    using Sh = TileShape<T, S>;
    constexpr int C = Sh::C, H = Sh::H, R = C + 2 * H, RI = R + 2, NTH = Sh::NTH;
    constexpr int NT = RI * RI * RI;
    static_assert(H + 1 <= C, "a walker stops in its tile or a neighbouring one");
    extern __shared__ __align__(16) unsigned char smem[];
    T *simg = (T *)smem;
    uint32_t *shits = (uint32_t *)(smem + tile_img_bytes<T, S>());
    __shared__ uint32_t next, bucket;
    __shared__ uint32_t snb[27];   // walkers of the bucket stopping in each neighbour tile

    const uint32_t nb = ctl->nb, nheavy = ctl->nheavy;
    if (PERSIST && threadIdx.x == 0) atomicMin(&ctl->t0, (unsigned long long)__builtin_amdgcn_s_memrealtime());
    while (true) {
        if (threadIdx.x == 0) {
            bucket = PERSIST ? atomicAdd(&ctl->bnext, 1u) : blockIdx.x;
            next = 0u;
        }
        __syncthreads();
        // the same value in every thread: made wave-uniform so that the bucket's
        // geometry lives in scalar registers
        const uint32_t bi = PERSIST ? __builtin_amdgcn_readfirstlane(bucket) : blockIdx.x;
        if (bi >= nb) break;
        const uint32_t slot = bucket_slot(bi, nheavy, nbcap);
        const uint32_t key = bkey[slot];
        const uint32_t off = boff[slot], cnt = bcnt[slot];
        const uint32_t label = key % tg.nlabels;
        const uint32_t tile = key / tg.nlabels;
        // the bucket's volume: geometry, image and hit block in volume-local indices
        const VolGeom gv = tg.vols[vol_of_tile(tg.vols, tg.nvol, tile)];
        const WalkGeom g{gv.Z, gv.Y, gv.X, gv.sy, gv.sz, (int64_t)gv.V};
        const T *__restrict__ img = img_all + gv.base;
        const uint32_t lt = tile - gv.tile_base;
        const int tx = (int)(lt % gv.ntx), ty = (int)((lt / gv.ntx) % gv.nty),
                  tz = (int)(lt / (gv.ntx * gv.nty));
        // volume coordinates of tile voxel (0,0,0)
        const int bz = tz * C - H - 1, by = ty * C - H - 1, bx = tx * C - H - 1;

        // Tile load: each thread issues a batch of image loads before storing them
        // to LDS, so one memory latency is paid per batch rather than per element;
        // the hit counters are cleared with 16-byte stores.
        // tid: threadIdx.x made opaque to the compiler, so that the per-thread tile
        // addresses below are recomputed per bucket rather than hoisted out of
        // the bucket loop, where they would hold registers through the walk
        int tid;
        asm volatile("" : "=v"(tid) : "0"((int)threadIdx.x));
        {
            // a thread's loads in two batches (two memory latencies per tile,
            // and few enough registers); tile element k = tid + q * NTH at
            // (lz, ly, lx), advanced by NTH elements per entry; image indices
            // fit 32 bits (a volume has fewer than 2^32 voxels) and are
            // clamped only for a tile that reaches past the volume
            constexpr int BATCH = (NT + 2 * NTH - 1) / (2 * NTH);
            constexpr int DZ = NTH / (RI * RI), DY = (NTH / RI) % RI, DX = NTH % RI;
            int lz = tid / (RI * RI), ly = (tid / RI) % RI, lx = tid % RI;
            const bool inner = bz >= 0 && by >= 0 && bx >= 0 && bz + RI <= g.Z && by + RI <= g.Y && bx + RI <= g.X;
            for (int base = 0; base < NT; base += BATCH * NTH) {
                T v[BATCH];
#pragma unroll
                for (int q = 0; q < BATCH; ++q) {
                    if (base + tid + q * NTH < NT) {
                        int gz = bz + lz, gy = by + ly, gx = bx + lx;
                        if (!inner) {
                            gz = min(max(gz, 0), g.Z - 1);
                            gy = min(max(gy, 0), g.Y - 1);
                            gx = min(max(gx, 0), g.X - 1);
                        }
                        v[q] = img[(uint32_t)gz * g.sz + (uint32_t)gy * g.sy + (uint32_t)gx];
                    }
                    lx += DX;
                    if (lx >= RI) { lx -= RI; ++ly; }
                    ly += DY;
                    if (ly >= RI) { ly -= RI; ++lz; }
                    lz += DZ;
                }
#pragma unroll
                for (int q = 0; q < BATCH; ++q)
                    if (base + tid + q * NTH < NT) simg[base + tid + q * NTH] = v[q];
            }
            constexpr int NW = (int)(tile_hit_bytes<T, S>() / 4);   // 32-bit counter words
            if (tid < 27) snb[tid] = 0u;
            uint4 *sh4 = (uint4 *)shits;
            for (int k = tid; k < NW / 4; k += NTH) sh4[k] = make_uint4(0u, 0u, 0u, 0u);
            for (int k = (NW / 4) * 4 + tid; k < NW; k += NTH) shits[k] = 0u;
        }
        __syncthreads();

        // Interior voxels are 1..dim-2; in tile coordinates 1-b..dim-2-b.  A
        // walker is never more than 3 steps outside the region, so a tile whose
        // region grown by 4 stays inside the interior needs no move checks.
        const int lmin0 = 1 - bz, lmin1 = 1 - by, lmin2 = 1 - bx;
        const int lmax0 = g.Z - 2 - bz, lmax1 = g.Y - 2 - by, lmax2 = g.X - 2 - bx;
        const bool edge = (lmin0 > -3) | (lmin1 > -3) | (lmin2 > -3) | (lmax0 < R + 4) |
                          (lmax1 < R + 4) | (lmax2 < R + 4);
        const auto c31 = [](int v) { return min(max(v, 0), 31); };
        const uint32_t LO = ppack(c31(lmin0), c31(lmin1), c31(lmin2));
        const uint32_t HI = ppack(c31(lmax0), c31(lmax1), c31(lmax2));

        uint32_t *__restrict__ hg = hits + gv.hits_base + (int64_t)label * gv.V;
        // global index of tile voxel (0,0,0) (negative at the volume's low faces)
        const int64_t g0 = (int64_t)bz * g.sz + (int64_t)by * g.sy + bx;
        uint32_t acc = 0;  // accepted moves of this thread in this bucket (< cnt * cap < 2^32)
        bool have = false;
        uint32_t j = 0, p0 = 0, seed = 0, wl = 0, weight = 0;
        int step = 0, step0 = 0, lq = 0;
        uint32_t P = 0;   // packed tile coordinates
        float B = 0.0f, div1 = 0.0f;
        while (true) {
            if (!have) {
                j = atomicAdd(&next, 1u);
                if (j >= cnt) break;
                const uint4 w4 = st[off + j];
                const uint32_t pos = w4.x - gv.base;
                seed = w4.y;
                wl = w4.z;
                step = (int)w4.w;
                const WalkSeed ws = wseeds[seed];
                p0 = ws.pos;   // volume-local: a seed's Philox stream does not depend on the batch
                B = ws.B;
                div1 = ws.div1;
                weight = ws.weight;
                const int z = (int)div_u32(pos, gv.mz, gv.shz);
                const uint32_t r = pos - (uint32_t)z * g.sz;
                const int y = (int)div_u32(r, gv.my, gv.shy);
                const int x = (int)(r - (uint32_t)y * g.sy);
                const int lz = z - bz, ly = y - by, lx = x - bx;
                P = ppack(lz, ly, lx);
                lq = (lz * RI + ly) * RI + lx - RI * RI;
                step0 = step;
                have = true;
            }
            const bmx_u4 r4 = bmx_philox(p0, 0u, wl & 0xFFFFu, (uint32_t)(step >> 2), k0, k1);
            const uint32_t rr[4] = {r4.x, r4.y, r4.z, r4.w};
            if (edge)
                tile_block<T, S, true>(simg, shits, sorw, rr, B, div1, step, lq, P, LO, HI, weight, acc, hg, g0,
                                       g.sz, g.sy);
            else
                tile_block<T, S, false>(simg, shits, sorw, rr, B, div1, step, lq, P, LO, HI, weight, acc, hg, g0,
                                        g.sz, g.sy);
            const bool inside = p_inside<R>(P);
            // Drain pause: once the bucket has no walkers left to hand out, a
            // wave with at most half its lanes busy stops them like a cap stop
            // (they resume next round, grouped with other walkers, instead of
            // keeping the wave's 64 lanes for a few long walks).  A walker's
            // path is a function of its seed, walk and step only and the hit
            // counts are sums, so where it pauses does not change the result.
            // A walker pauses after 4 steps of its visit at the earliest, so a
            // visit that does not end the walk takes at least min(4, H + 1)
            // steps, as without pauses (a visit starts in the core, H + 1
            // steps from the outside of the region; see min_visit_steps).
            bool pause = false;
            if (*(volatile uint32_t *)&next >= cnt) pause = __popcll(__ballot(true)) <= 32 && step - step0 >= 4;
            if (step >= sorw || !inside || step - step0 >= cap || pause) {
                const int lz = pz(P), ly = py(P), lx = px(P);
                const uint32_t pos =
                    (uint32_t)(bz + lz) * g.sz + (uint32_t)(by + ly) * g.sy + (uint32_t)(bx + lx);
                uint2 d = make_uint2(kDestDead, 0u);
                uint32_t wl2 = wl;
                if (step < sorw) {
                    // the tile it stopped in, relative to this one: local core [1+H, 1+H+C)
                    const auto rel = [](int l) { return l < 1 + H ? 0 : (l < 1 + H + C ? 1 : 2); };
                    const uint32_t nbr = (uint32_t)(9 * rel(lz) + 3 * rel(ly) + rel(lx));
                    d = make_uint2(bi, atomicAdd(&snb[nbr], 1u));
                    wl2 = (wl & 0x00FFFFFFu) | (nbr << 24);
                }
                st[off + j] = make_uint4(gv.base + pos, seed, wl2, (uint32_t)step);
                dest[off + j] = d;
                have = false;
            }
        }
        wave_add_u64(accepted, (uint64_t)acc);
        __syncthreads();
        // reserve the next round's entries of the neighbour keys this bucket's
        // walkers stopped in
        if (tid < 27) {
            const uint32_t c = snb[tid];
            if (c) {
                const int nz = tz + tid / 9 - 1, ny = ty + (tid / 3) % 3 - 1, nx = tx + tid % 3 - 1;
                const uint32_t key_nb =
                    (gv.tile_base + ((uint32_t)nz * gv.nty + (uint32_t)ny) * gv.ntx + (uint32_t)nx) * tg.nlabels +
                    label;
                bres[(int64_t)bi * 27 + tid] = make_uint2(key_nb, atomicAdd(&counts[key_nb], c));
            }
        }
        // one counter per lane, consecutive lanes on consecutive voxels; a 16-bit
        // counter k is half k of the LDS words (little-endian: voxel 2w is the
        // low half of word w), and it is not 0 only for a voxel of the volume
        // (hit16_wrap).  Counter k = tid + q * NTH at tile (lz, ly, lx) is
        // voxel gi of the volume, both advanced by NTH counters per q; gi is
        // kept modulo 2^32 (g0 may be negative, a voxel index is below 2^32).
        {
            constexpr int DZ = NTH / (RI * RI), DY = (NTH / RI) % RI, DX = NTH % RI;
            int fz = tid / (RI * RI), fy = (tid / RI) % RI, fx = tid % RI;
            const uint32_t fsy = g.sy, fsz = g.sz;
            uint32_t gi = (uint32_t)g0 + (uint32_t)fz * fsz + (uint32_t)fy * fsy + (uint32_t)fx;
            const uint32_t dgi = (uint32_t)DZ * fsz + (uint32_t)DY * fsy + (uint32_t)DX;
#pragma unroll 4
            for (int q = 0; q < (NT + NTH - 1) / NTH; ++q) {
                const int k = tid + q * NTH;
                if (k < NT) {
                    const uint32_t c = Sh::HB == 4 ? shits[k] : (uint32_t)((const uint16_t *)shits)[k];
                    if (c) atomicAdd(&hg[gi], c);
                }
                gi += dgi;
                fx += DX;
                if (fx >= RI) { fx -= RI; ++fy; gi += fsy - (uint32_t)RI; }
                fy += DY;
                if (fy >= RI) { fy -= RI; ++fz; gi += fsz - (uint32_t)RI * fsy; }
                fz += DZ;
            }
        }
        if (!PERSIST) break;
        __syncthreads();   // the LDS tile and `bucket` are reused by the next bucket
    }
    // Tiny buckets (graph-driven rounds): each wave takes one, each lane walks one
    // of its walkers from global memory.  Workgroups come here as they run out of
    // tile buckets, so this work fills the end of the round.
    if (PERSIST) {
        const uint32_t ntiny = ctl->ntiny;
        const int lane = (int)(threadIdx.x & 63u);
        uint32_t tacc = 0;
        while (true) {
            uint32_t tb = 0;
            if (lane == 0) tb = atomicAdd(&ctl->tnext, 1u);
            tb = __builtin_amdgcn_readfirstlane(__shfl(tb, 0, 64));
            if (tb >= ntiny) break;
            const uint32_t key = tkey[tb], off = toff[tb], cnt = tcnt[tb];
            const uint32_t label = key % tg.nlabels, tile = key / tg.nlabels;
            const VolGeom gv = tg.vols[vol_of_tile(tg.vols, tg.nvol, tile)];
            const WalkGeom g{gv.Z, gv.Y, gv.X, gv.sy, gv.sz, (int64_t)gv.V};
            uint32_t *__restrict__ hg = hits + gv.hits_base + (int64_t)label * gv.V;
            for (uint32_t w = (uint32_t)lane; w < cnt; w += 64u) {
                uint4 w4 = st[off + w];
                tiny_walker<T>(img_all + gv.base, g, gv, wseeds, w4, sorw, cap, k0, k1, hg, tacc);
                uint2 d = make_uint2(kDestDead, 0u);
                if ((int)w4.w < sorw) {
                    const uint32_t key2 = tile_key<C>(w4.x, label, tg);
                    d = make_uint2(kDestDirect | key2, atomicAdd(&counts[key2], 1u));
                }
                st[off + w] = w4;
                dest[off + w] = d;
            }
        }
        wave_add_u64(accepted, (uint64_t)tacc);
    }
    // The span of a persistent launch, first workgroup start to last workgroup
    // end, is added to ctl->ticks by the workgroup that ends last, which then
    // clears the round's fields for the next launch.
    if (PERSIST && threadIdx.x == 0) {
        atomicMax(&ctl->t1, (unsigned long long)__builtin_amdgcn_s_memrealtime());
        __threadfence();
        if (atomicAdd(&ctl->done, 1u) == gridDim.x - 1) {
            __threadfence();
            const unsigned long long t0 = atomicAdd(&ctl->t0, 0ull), t1 = atomicAdd(&ctl->t1, 0ull);
            ctl->ticks += t1 > t0 ? t1 - t0 : 0ull;
            ctl->t0 = ~0ull;
            ctl->t1 = 0ull;
            ctl->done = 0u;
        }
    }
}

// The fewest steps a bucket visit takes unless it ends the walk: a visit
// starts in the core, H + 1 steps from the outside of the region; cap stops
// and drain pauses come after 4 steps at the earliest (the cap is a positive
// multiple of 4); a tiny walker walks the cap.  Bounds the rounds of a walk.
template <typename T, int S>
constexpr int min_visit_steps()
{
    return TileShape<T, S>::H + 1 < 4 ? TileShape<T, S>::H + 1 : 4;
}

// Resident workgroups of the persistent k_tile_walk<T, S> on the context's
// device: the grid of graph-driven rounds.
template <typename T, int S>
int walk_grid(bmx_ctx *ctx, int *grid)
{
    const size_t lds = tile_lds_bytes<T, S>();
    // Raise the dynamic-LDS limit where the runtime wants it; an error here is
    // left to the launch itself to report.
    if (hipFuncSetAttribute((const void *)k_tile_walk<T, S, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
        (void)hipGetLastError();
    if (hipFuncSetAttribute((const void *)k_tile_walk<T, S, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds) != hipSuccess)
        (void)hipGetLastError();
    int per_cu = 0, cus = 0;
    BMX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_tile_walk<T, S, true>,
                                                         TileShape<T, S>::NTH, lds));
    BMX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    *grid = (per_cu < 1 ? 1 : per_cu) * (cus < 1 ? 1 : cus);
    return BMX_OK;
}

// Everything a round launches, fixed for the whole walk.
template <typename T>
struct RoundArgs {
    const T *img;
    TileGeom tg;
    const WalkSeed *wseeds;
    uint32_t *counts, *bkey, *boff, *bcnt, *tkey, *toff, *tcnt;
    RoundCtl *ctl;
    int64_t nkeys;
    uint32_t nbcap;   // bucket-list entries: nkeys + split_extra(split_grid)
    uint32_t split_grid;
    int sorw, cap;
    uint32_t k0, k1;
    uint32_t *hits;
    unsigned long long *acc;
    uint2 *dest;   // per entry of the walked state array (see kDestDead)
    uint2 *bres;   // 27 per round bucket
};

// Re-bucketing of one round: the key counts left by the walk before (or by
// k_keys_first) scanned into offsets, listed as buckets, and the live walkers
// of `in` (ctl->n entries) placed into `out` by their destination records;
// ctl->n becomes the live count and the counts are cleared for the walk.
// `grid` blocks run the per-walker kernel.
template <typename T, int S>
int round_bucket(bmx_ctx *ctx, const RoundArgs<T> &a, const uint4 *in, uint4 *out, unsigned grid, uint32_t tiny)
{
    int rc = scan_u32(ctx, a.counts, a.nkeys + 1);
    if (rc) return rc;
    k_round_mid<<<1, 64, 0, ctx->stream>>>(a.counts, a.nkeys, a.ctl);
    BMX_LAUNCH_CHECK();
    k_bucket_list<<<bmx_blocks(a.nkeys, 256), 256, 0, ctx->stream>>>(
        a.counts, a.nkeys, a.nbcap, a.bkey, a.boff, a.bcnt, a.tkey, a.toff, a.tcnt, a.ctl,
        (uint32_t)TileShape<T, S>::NTH, tiny, a.split_grid);
    BMX_LAUNCH_CHECK();
    k_tile_place<<<grid, 256, 0, ctx->stream>>>(in, a.dest, a.ctl, a.counts, a.bres, out);
    BMX_LAUNCH_CHECK();
    k_round_start<<<bmx_stride_blocks(a.nkeys + 1, 256), 256, 0, ctx->stream>>>(a.counts, a.nkeys + 1, a.ctl);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

template <typename T, int S, bool PERSIST>
int round_walk(bmx_ctx *ctx, const RoundArgs<T> &a, uint4 *st, unsigned grid)
{
    k_tile_walk<T, S, PERSIST><<<grid, TileShape<T, S>::NTH, tile_lds_bytes<T, S>(), ctx->stream>>>(
        a.img, a.tg, a.wseeds, st, a.bkey, a.boff, a.bcnt, a.tkey, a.toff, a.tcnt, a.ctl, a.nbcap,
        a.sorw, a.cap, a.k0, a.k1, a.hits, a.acc, a.dest, a.bres, a.counts);
    BMX_LAUNCH_CHECK();
    return BMX_OK;
}

// Captures 2 * pairs rounds (state A -> B -> A ...) as one HIP graph.
template <typename T, int S>
int capture_rounds(bmx_ctx *ctx, const RoundArgs<T> &a, uint4 *A, uint4 *Bst, unsigned grid, unsigned wgrid,
                   int pairs, hipGraphExec_t *exec)
{
    hipGraph_t graph = nullptr;
    BMX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = BMX_OK;
    for (int r = 0; r < 2 * pairs && rc == BMX_OK; ++r) {
        uint4 *in = (r & 1) ? Bst : A, *out = (r & 1) ? A : Bst;
        rc = round_bucket<T, S>(ctx, a, in, out, grid, kTinyWalkers);
        if (rc == BMX_OK) rc = round_walk<T, S, true>(ctx, a, out, wgrid);
    }
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc == BMX_OK && e != hipSuccess)
        rc = bmx_set_error(BMX_E_DEVICE, "bmx_walk: graph capture failed: %s", hipGetErrorString(e));
    if (rc == BMX_OK) {
        const hipError_t ei = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
        if (ei != hipSuccess)
            rc = bmx_set_error(BMX_E_DEVICE, "bmx_walk: graph instantiation failed: %s", hipGetErrorString(ei));
    }
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
}

// Rounds driven from the device: 2 * pairs rounds captured once in a HIP graph
// (no host value enters a round; a round with no live walker does no work),
// replayed until no walker is live, the host reading one count per replay.
// An even round count returns the state to A.  *kernel_ms receives the summed
// spans of the walk launches, which the persistent kernel measures itself on
// the device wall clock (events cannot bracket launches inside a graph).
template <typename T, int S>
int rounds_graph(bmx_ctx *ctx, const RoundArgs<T> &a, uint4 *A, uint4 *Bst, unsigned grid, unsigned wgrid,
                 int pairs, int *rounds, float *kernel_ms)
{
    int khz = 0;
    BMX_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
    hipGraphExec_t exec = nullptr;
    int rc = capture_rounds<T, S>(ctx, a, A, Bst, grid, wgrid, pairs, &exec);
    if (rc) return rc;
    const int max_rounds = a.sorw / min_visit_steps<T, S>() + 2 + 2 * pairs;
    RoundCtl host{};
    while (rc == BMX_OK) {
        hipError_t el = hipGraphLaunch(exec, ctx->stream);
        if (el == hipSuccess)
            el = hipMemcpyAsync(&host, a.ctl, sizeof(host), hipMemcpyDeviceToHost, ctx->stream);
        if (el == hipSuccess) el = hipStreamSynchronize(ctx->stream);
        if (el != hipSuccess) {
            rc = bmx_set_error(BMX_E_DEVICE, "bmx_walk: graph replay failed: %s", hipGetErrorString(el));
            break;
        }
        *rounds += 2 * pairs;
        if (host.n == 0) break;
        if (*rounds > max_rounds) rc = bmx_set_error(BMX_E_INVALID, "bmx_walk: tiled walker did not converge");
    }
    (void)hipGraphExecDestroy(exec);
    *kernel_ms = khz > 0 ? (float)((double)host.ticks / (double)khz) : -2.0f;
    return rc;
}

template <typename T, int S>
int walk_tiled_t(bmx_ctx *ctx, const T *img, const bmx_volume *vols, int nvol, const bmx_seed *seeds,
                 int64_t nseeds, int nlabels, int nbrw, int sorw, int cap, uint32_t weight, int overrides,
                 uint32_t k0, uint32_t k1, uint32_t *hits, unsigned long long *acc, int graph_pairs,
                 hipEvent_t *ev, int max_timed, int *launches, float *kernel_ms)
{
    constexpr int C = TileShape<T, S>::C;
    const int64_t n0 = nseeds * (int64_t)nbrw;
    BMX_CHECK(n0 < ((int64_t)1 << 32), "bmx_walk: %lld walkers exceed 2^32", (long long)n0);
    BMX_CHECK(nbrw <= 65535, "bmx_walk: tiled walker needs nbrw <= 65535");
    BMX_CHECK(nvol >= 1, "bmx_walk: no volumes");
    std::vector<VolGeom> vg((size_t)nvol);
    int64_t ntiles = 0;
    for (int v = 0; v < nvol; ++v) {
        const bmx_volume &b = vols[v];
        VolGeom &g = vg[(size_t)v];
        g.base = (uint32_t)b.offset;
        g.Z = b.Z;
        g.Y = b.Y;
        g.X = b.X;
        g.sy = (uint32_t)b.X;
        g.sz = (uint32_t)b.X * (uint32_t)b.Y;
        g.V = (uint32_t)((int64_t)b.Z * b.Y * b.X);
        g.nty = (uint32_t)((b.Y + C - 1) / C);
        g.ntx = (uint32_t)((b.X + C - 1) / C);
        g.tile_base = (uint32_t)ntiles;
        g.hits_base = b.hits_offset;
        g.seed_base = (uint32_t)b.seed_offset;
        g.pad = 0;
        BMX_CHECK(b.X >= 2, "bmx_walk: volume %d has X = %d < 2", v, b.X);
        magic_u32(g.sz, &g.mz, &g.shz);
        magic_u32(g.sy, &g.my, &g.shy);
        ntiles += (int64_t)((b.Z + C - 1) / C) * g.nty * g.ntx;
    }
    const int64_t nkeys = ntiles * nlabels;
    BMX_CHECK(nkeys < ((int64_t)1 << 31), "bmx_walk: too many tiles");
    int rc = grow(ctx, &ctx->d_vols, &ctx->vols_cap, (size_t)nvol * sizeof(VolGeom));
    if (rc) return rc;
    BMX_HIP(hipMemcpyAsync(ctx->d_vols, vg.data(), (size_t)nvol * sizeof(VolGeom), hipMemcpyHostToDevice,
                           ctx->stream));
    // two walker-state arrays and the destination records of one of them
    rc = grow(ctx, &ctx->d_state, &ctx->state_cap, (size_t)n0 * (2 * sizeof(uint4) + sizeof(uint2)));
    if (rc) return rc;
    rc = grow(ctx, &ctx->d_wseeds, &ctx->wseeds_cap, (size_t)(nseeds > 0 ? nseeds : 1) * sizeof(WalkSeed));
    if (rc) return rc;
    rc = grow(ctx, &ctx->d_keys, &ctx->keys_cap, (size_t)(nkeys + 1) * sizeof(uint32_t));
    if (rc) return rc;
    int wgrid_max = 0;
    rc = walk_grid<T, S>(ctx, &wgrid_max);
    if (rc) return rc;
    // three uint32 lists of nbcap round buckets (keys split over several
    // buckets need up to split_extra entries more than the keys), three of
    // nkeys tiny buckets, then 27 neighbour reservations per round bucket at
    // an 8-byte boundary
    const int64_t nbcap = nkeys + split_extra(wgrid_max);
    const int64_t list_words = (3 * nbcap + 3 * nkeys + 1) & ~(int64_t)1;
    rc = grow(ctx, &ctx->d_buckets, &ctx->buckets_cap,
              (size_t)list_words * sizeof(uint32_t) + (size_t)nbcap * 27 * sizeof(uint2));
    if (rc) return rc;
    // the scan's block partials, allocated before any capture
    rc = grow(ctx, &ctx->d_scan, &ctx->scan_cap,
              (size_t)((nkeys + 1 + kScanTile - 1) / kScanTile) * sizeof(uint32_t));
    if (rc) return rc;

    RoundArgs<T> a;
    a.img = img;
    a.tg = TileGeom{(const VolGeom *)ctx->d_vols, nvol, (uint32_t)nlabels};
    a.wseeds = (const WalkSeed *)ctx->d_wseeds;
    a.counts = (uint32_t *)ctx->d_keys;
    a.bkey = (uint32_t *)ctx->d_buckets;
    a.boff = a.bkey + nbcap;
    a.bcnt = a.boff + nbcap;
    a.tkey = a.bcnt + nbcap;
    a.toff = a.tkey + nkeys;
    a.tcnt = a.toff + nkeys;
    a.nbcap = (uint32_t)nbcap;
    a.split_grid = (uint32_t)wgrid_max;
    a.ctl = (RoundCtl *)(ctx->d_counter + 8);
    static_assert(sizeof(RoundCtl) <= 8 * sizeof(uint64_t), "RoundCtl fits the context's scratch counters");
    a.nkeys = nkeys;
    a.sorw = sorw;
    a.cap = cap;
    a.k0 = k0;
    a.k1 = k1;
    a.hits = hits;
    a.acc = acc;
    uint4 *A = (uint4 *)ctx->d_state, *Bst = A + n0;
    a.dest = (uint2 *)(Bst + n0);
    a.bres = (uint2 *)(a.bkey + list_words);
    if (nseeds > 0) {
        k_walk_seeds<T><<<bmx_blocks(nseeds, 256), 256, 0, ctx->stream>>>(img, a.tg, seeds, nseeds, weight, overrides,
                                                                         (WalkSeed *)ctx->d_wseeds);
        BMX_LAUNCH_CHECK();
    }
    k_tile_init<<<bmx_blocks(n0, 256), 256, 0, ctx->stream>>>(seeds, nseeds, nbrw, sorw, overrides, a.tg, A);
    BMX_LAUNCH_CHECK();
    k_set_count<<<1, 64, 0, ctx->stream>>>(a.ctl, (uint32_t)n0);
    BMX_LAUNCH_CHECK();
    constexpr int64_t kMaxWalkerBlocks = (int64_t)1 << 16;
    const auto walker_grid = [](int64_t n) {
        const int64_t b = (n + 255) / 256;
        return (unsigned)(b < 1 ? 1 : (b < kMaxWalkerBlocks ? b : kMaxWalkerBlocks));
    };
    // the first round's key counts and destinations, from the walkers as created
    k_clear_counts<<<bmx_stride_blocks(nkeys + 1, 256), 256, 0, ctx->stream>>>(a.counts, nkeys + 1);
    BMX_LAUNCH_CHECK();
    k_keys_first<C><<<walker_grid(n0), 256, 0, ctx->stream>>>(A, a.ctl, a.tg, sorw, a.counts, a.dest);
    BMX_LAUNCH_CHECK();
    int round = 0;
    *launches = 0;
    *kernel_ms = -1.0f;
    if (graph_pairs > 0) {
        rc = rounds_graph<T, S>(ctx, a, A, Bst, walker_grid(n0), (unsigned)wgrid_max, graph_pairs, &round,
                                kernel_ms);
        *launches = round;
        return rc;
    }

    // Host-driven rounds: the host reads the live and bucket counts after each
    // re-bucketing and sizes the next launches by them.
    int64_t n = n0;
    // BMX_WALK_TRACE=1 prints one line per round (live walkers, buckets, kernel ms) to stderr.
    static const bool trace = getenv("BMX_WALK_TRACE") && getenv("BMX_WALK_TRACE")[0] == '1';
    while (n > 0) {
        rc = round_bucket<T, S>(ctx, a, A, Bst, walker_grid(n), 0u);
        if (rc) return rc;
        RoundCtl host{};
        BMX_HIP(hipMemcpyAsync(&host, a.ctl, sizeof(host), hipMemcpyDeviceToHost, ctx->stream));
        BMX_HIP(hipStreamSynchronize(ctx->stream));
        const int64_t live = host.n;
        const uint32_t nb = host.nb;
        if (live == 0) break;
        if (round < max_timed) BMX_HIP(hipEventRecord(ev[2 * round], ctx->stream));
        rc = round_walk<T, S, false>(ctx, a, Bst, nb);
        if (rc) return rc;
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
        BMX_CHECK(round <= sorw / (min_visit_steps<T, S>()) + 2, "bmx_walk: tiled walker did not converge");
    }
    *launches = round < max_timed ? round : max_timed;
    return BMX_OK;
}

// Walker indices, the live count and the round buffers of one walk are 32-bit.
// A single volume with 2^32 walkers or more is walked in balanced chunks of
// whole seeds, one after the other, into the same hits and move counter.  A
// walker's stream is keyed by its seed's voxel, walk and step, never by its
// index, so the chunks' hits add up to those of one walk over every seed.
// The chunks also bound the round buffers (40 bytes per walker) at 2^32
// walkers.  BMX_WALK_CHUNK=n lowers the chunk size to n walkers, so that
// chunked and unchunked walks can be compared on small volumes.
template <typename T, int S>
int walk_tiled_chunked(bmx_ctx *ctx, const T *img, const bmx_volume *vols, int nvol, const bmx_seed *seeds,
                       int64_t nseeds, int nlabels, int nbrw, int sorw, int cap, uint32_t weight, int overrides,
                       uint32_t k0, uint32_t k1, uint32_t *hits, unsigned long long *acc, int graph_pairs,
                       hipEvent_t *ev, int max_timed, int *launches, float *kernel_ms)
{
    int64_t lim = ((int64_t)1 << 32) - 1;
    const char *env = getenv("BMX_WALK_CHUNK");
    if (env && *env) {
        const long long v = atoll(env);
        if (v > 0 && v < lim) lim = v;
    }
    const int64_t n0 = nseeds * (int64_t)nbrw;
    if (nvol != 1 || nbrw <= 0 || n0 <= lim)
        return walk_tiled_t<T, S>(ctx, img, vols, nvol, seeds, nseeds, nlabels, nbrw, sorw, cap, weight, overrides,
                                  k0, k1, hits, acc, graph_pairs, ev, max_timed, launches, kernel_ms);
    const int64_t per = std::max<int64_t>(1, lim / nbrw);
    const int64_t nchunks = (nseeds + per - 1) / per;
    const int64_t chunk = (nseeds + nchunks - 1) / nchunks;
    int total = 0;
    bool unknown = false, host = false;
    float ms_sum = 0.0f;
    for (int64_t s0 = 0; s0 < nseeds; s0 += chunk) {
        const int64_t cnt = std::min(chunk, nseeds - s0);
        bmx_volume v = vols[0];
        v.seed_offset = 0;
        v.nseeds = (uint64_t)cnt;
        const int used = std::min(total, max_timed);
        int l = 0;
        float ms = -1.0f;
        const int rc = walk_tiled_t<T, S>(ctx, img, &v, 1, seeds + s0, cnt, nlabels, nbrw, sorw, cap, weight,
                                          overrides, k0, k1, hits, acc, graph_pairs, ev + 2 * used,
                                          max_timed - used, &l, &ms);
        if (rc) return rc;
        total += l;
        if (ms == -2.0f) unknown = true;
        else if (ms >= 0.0f) ms_sum += ms;
        else host = true;
    }
    *launches = total;
    *kernel_ms = unknown ? -2.0f : (host ? -1.0f : ms_sum);
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

template <typename T, int S>
static void shape_info(int *core, int *halo, int *threads, int *hit_bytes, int *lds_bytes)
{
    *core = TileShape<T, S>::C;
    *halo = TileShape<T, S>::H;
    *threads = TileShape<T, S>::NTH;
    *hit_bytes = TileShape<T, S>::HB;
    *lds_bytes = (int)tile_lds_bytes<T, S>();
}

extern "C" int bmx_tile_shape_info(int img_dtype, int shape, int *core, int *halo, int *threads, int *hit_bytes,
                                   int *lds_bytes)
{
    BMX_CHECK(core && halo && threads && hit_bytes && lds_bytes, "bmx_tile_shape_info: null pointer");
    BMX_CHECK(img_dtype == BMX_IMG_U8 || img_dtype == BMX_IMG_F32, "bmx_tile_shape_info: bad image dtype %d",
              img_dtype);
    BMX_CHECK(shape >= 0 && shape < BMX_TILE_SHAPES, "bmx_tile_shape_info: tile_shape %d not in [0,%d)", shape,
              BMX_TILE_SHAPES);
#define BMX_INFO_CASE(TT, SS) \
    case SS: shape_info<TT, SS>(core, halo, threads, hit_bytes, lds_bytes); return BMX_OK;
    if (img_dtype == BMX_IMG_U8) {
        switch (shape) {
            BMX_INFO_CASE(uint8_t, 0) BMX_INFO_CASE(uint8_t, 1) BMX_INFO_CASE(uint8_t, 2)
            BMX_INFO_CASE(uint8_t, 3) BMX_INFO_CASE(uint8_t, 4) BMX_INFO_CASE(uint8_t, 5)
        }
    } else {
        switch (shape) {
            BMX_INFO_CASE(float, 0) BMX_INFO_CASE(float, 1) BMX_INFO_CASE(float, 2)
            BMX_INFO_CASE(float, 3) BMX_INFO_CASE(float, 4) BMX_INFO_CASE(float, 5)
        }
    }
#undef BMX_INFO_CASE
    return bmx_set_error(BMX_E_INVALID, "bmx_tile_shape_info: bad tile shape");
}

// Entry from bmx_walk_ex / bmx_walk_batch (bmx_walk.hip).  A single volume is a
// one-entry table.
int bmx_walk_tiled(bmx_ctx *ctx, const void *img, int img_dtype, const bmx_volume *vols, int nvol,
                   const bmx_seed *seeds, int64_t nseeds, int nlabels, int nbrw, int sorw, int cap,
                   uint32_t weight, int overrides, uint32_t k0, uint32_t k1, uint32_t *hits,
                   unsigned long long *acc, int shape, int graph_pairs, hipEvent_t *ev, int max_timed,
                   int *launches, float *kernel_ms)
{
    BMX_CHECK(shape >= 0 && shape < BMX_TILE_SHAPES, "bmx_walk: tile_shape %d not in [0,%d)", shape,
              BMX_TILE_SHAPES);
    // per-hit weight: hit_weight, times a seed's own weight (< 2^16) with overrides
    const bool hit16 = overrides ? weight <= 1u : weight <= kHit16MaxWeight;
#define BMX_TILED_CASE(TT, SS)                                                                                \
    case SS: {                                                                                                \
        constexpr int S2 = TileShape<TT, SS>::HB == 2 ? Hit32Sibling<TT, SS>::value : SS;                     \
        if (!hit16 && S2 != SS)                                                                               \
            return walk_tiled_chunked<TT, S2>(ctx, (const TT *)img, vols, nvol, seeds, nseeds, nlabels, nbrw, \
                                              sorw, cap, weight, overrides, k0, k1, hits, acc, graph_pairs, ev, \
                                              max_timed, launches, kernel_ms);                                \
        return walk_tiled_chunked<TT, SS>(ctx, (const TT *)img, vols, nvol, seeds, nseeds, nlabels, nbrw, sorw, \
                                          cap, weight, overrides, k0, k1, hits, acc, graph_pairs, ev, max_timed, \
                                          launches, kernel_ms);                                               \
    }
    if (img_dtype == BMX_IMG_U8) {
        switch (shape) {
            BMX_TILED_CASE(uint8_t, 0)
            BMX_TILED_CASE(uint8_t, 1)
            BMX_TILED_CASE(uint8_t, 2)
            BMX_TILED_CASE(uint8_t, 3)
            BMX_TILED_CASE(uint8_t, 4)
            BMX_TILED_CASE(uint8_t, 5)
        }
    } else {
        switch (shape) {
            BMX_TILED_CASE(float, 0)
            BMX_TILED_CASE(float, 1)
            BMX_TILED_CASE(float, 2)
            BMX_TILED_CASE(float, 3)
            BMX_TILED_CASE(float, 4)
            BMX_TILED_CASE(float, 5)
        }
    }
#undef BMX_TILED_CASE
    return bmx_set_error(BMX_E_INVALID, "bmx_walk: bad tile shape");
}
