/*
 * bmx_oracle.c -- CPU restatement of biomedisa's smart-interpolation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or, for ref_walk, as the timed CPU baseline).  The product path
 * (biomedisa_amd) never calls it.
 *
 * Two walkers live here:
 *   orc_ref_walk    -- literal restatement of the reference OpenCL kernel
 *                      `randomWalk` (biomedisa/features/random_walk/
 *                      pyopencl_small.py:192-314), including its float-valued
 *                      MRG32k3a-like recurrence seeded with the voxel index
 *                      (:213-223, :256-282) and its direction chain (:285-290).
 *                      Hits per label as in the per-label loop (:128-134).
 *   orc_philox_walk -- the same transition law driven by a counter-based
 *                      Philox4x32-10 stream per (seed voxel, walk).  This is
 *                      the bit-exact target of the HIP walker.
 * Both use orc_expf (Cody-Waite reduction + degree-6 near-minimax polynomial
 * with fused multiply-adds, ldexp scaling) in place of the vendor exp the reference JIT links
 * (pyopencl_small.py:187-190), so that CPU and GPU agree bit for bit.
 *
 * Arithmetic contract (build with -ffp-contract=off): every float operation
 * below is a single IEEE binary32 operation in the order written; fmaf is
 * used only where written explicitly.
 *
 * Parity pins (tests/test_oracle.py): Philox known-answer vectors (Random123),
 * orc_expf vs float64 exp, seed extraction vs brute force, transition law vs
 * its closed form (chi^2), hit conservation, argmax ties, curvature closed
 * forms, uncertainty closed forms.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_IGN 0xFFFFu   /* label index of ignored voxels (uint16 label volumes) */

/* ------------------------------------------------------------------------ */
/* Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as   */
/* easy as 1, 2, 3", SC'11.  Multipliers and Weyl constants from the paper. */
/* ------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                       uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------------ */
/* exp for binary32.  Cody & Waite argument reduction x = n*ln2 + r with a   */
/* two-part ln2 (hi part exact in 16 bits), fused with fmaf; e^r by a degree-5 */
/* polynomial 1 + r + c2 r^2 + ... + c5 r^5 whose coefficients are a near-     */
/* minimax (relative error) fit on |r| <= ln2/2 (fit error 1.06e-7 with the    */
/* float32 coefficients); 2^n applied with ldexpf, so results below FLT_MIN    */
/* underflow gradually as IEEE scaling does.  n = round(x log2 e) is read off  */
/* one fmaf onto 1.5 * 2^23 (its low mantissa bits hold n).  Max error 2.2 ulp */
/* on [-104, 0] against a long double exp.                                     */
/* ------------------------------------------------------------------------ */
float orc_expf(float x)
{
    const float LOG2E = 1.44269502925872802734375f;     /* 0x3fb8aa3b */
    const float LN2_HI = 0.693145751953125f;            /* 0x3f317200 */
    const float LN2_LO = 1.428606765330187045e-06f;     /* 0x35bfbe8e */
    const float SHIFT = 12582912.0f;                    /* 0x4b400000 */
    if (x > 88.0f) return INFINITY;
    if (x < -1000.0f) return 0.0f;
    const float s = fmaf(x, LOG2E, SHIFT);
    const float n = s - SHIFT;
    uint32_t su, shu;
    const float sh = SHIFT;
    memcpy(&su, &s, 4);
    memcpy(&shu, &sh, 4);
    const int ni = (int)(su - shu);
    float r = fmaf(-n, LN2_HI, x);
    r = fmaf(-n, LN2_LO, r);
    float p = 0.00831252523f;
    p = fmaf(p, r, 0.0418901145f);
    p = fmaf(p, r, 0.166671142f);
    p = fmaf(p, r, 0.499992311f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, ni);
}

/* weight(B, A, div1) of pyopencl_small.py:187-190: exp(-tmp*tmp*div1). */
static inline float orc_weight(float B, float A, float div1)
{
    float tmp = B - A;
    float e = ((-tmp) * tmp) * div1;
    return orc_expf(e);
}

/* ------------------------------------------------------------------------ */
/* Image access: dtype 0 = uint8, 1 = float32 (reference normalises every    */
/* other dtype to float32 in [0,255], interpolation.py:135-144).             */
/* ------------------------------------------------------------------------ */
static inline float orc_img(const void *img, int dtype, int64_t i)
{
    if (dtype == 0) return (float)((const uint8_t *)img)[i];
    return ((const float *)img)[i];
}

/* ------------------------------------------------------------------------ */
/* Labelled-plane detection.                                                  */
/*   XY mode (read_labeled_slices, biomedisa_helper.py:1015-1022): plane z   */
/*   is labelled iff it holds any non-zero label value.                      */
/*   allaxis (read_labeled_slices_allx, :1024-1045): per axis, voxels of     */
/*   1-voxel-thin lines are removed first along the first in-plane axis, then */
/*   along the second; a plane is labelled iff anything survives.  In-plane  */
/*   axis order: plane z -> (y, x); plane y -> (z, x); plane x -> (z, y).    */
/* labidx: uint16 label index per voxel (up to 256 labels, the reference's  */
/* short labels of pyopencl_small.py:90-95), 0 = background, ORC_IGN =      */
/* negative label; "non-zero label value" <=> labidx != 0.                   */
/* ------------------------------------------------------------------------ */
static inline int orc_nz(const uint16_t *lab, int64_t i) { return lab[i] != 0; }

/* ones after the first pass, at voxel (a,b) of a plane whose first in-plane  */
/* axis has extent na and stride sa. */
static inline int orc_pass1(const uint16_t *lab, int64_t base, int a, int na, int64_t sa)
{
    if (!orc_nz(lab, base)) return 0;
    if (a > 0 && a < na - 1 && !orc_nz(lab, base - sa) && !orc_nz(lab, base + sa)) return 0;
    return 1;
}

static inline int orc_pass2(const uint16_t *lab, int64_t base, int a, int na, int64_t sa,
                            int b, int nb, int64_t sb)
{
    if (!orc_pass1(lab, base, a, na, sa)) return 0;
    if (b > 0 && b < nb - 1 && !orc_pass1(lab, base - sb, a, na, sa)
        && !orc_pass1(lab, base + sb, a, na, sa)) return 0;
    return 1;
}

void orc_plane_flags(const uint16_t *lab, int Z, int Y, int X, int allaxis,
                     uint8_t *fz, uint8_t *fy, uint8_t *fx)
{
    const int64_t sy = X, sz = (int64_t)X * Y;
    memset(fz, 0, Z); memset(fy, 0, Y); memset(fx, 0, X);
    for (int z = 0; z < Z; ++z)
        for (int y = 0; y < Y; ++y)
            for (int x = 0; x < X; ++x) {
                int64_t i = z * sz + y * sy + x;
                if (!orc_nz(lab, i)) continue;
                if (!allaxis) { fz[z] = 1; continue; }
                if (!fz[z] && orc_pass2(lab, i, y, Y, sy, x, X, 1)) fz[z] = 1;
                if (!fy[y] && orc_pass2(lab, i, z, Z, sz, x, X, 1)) fy[y] = 1;
                if (!fx[x] && orc_pass2(lab, i, z, Z, sz, y, Y, sy)) fx[x] = 1;
            }
}

/* Mask value of the volume the reference rebuilds from the labelled slices  */
/* (pyopencl_small.py:97-107): label index inside a labelled plane, ORC_IGN  */
/* elsewhere (negative labels are ORC_IGN already). */
static inline uint16_t orc_mask(const uint16_t *lab, const uint8_t *fz, const uint8_t *fy,
                                const uint8_t *fx, int z, int y, int x, int64_t i)
{
    return (fz[z] || fy[y] || fx[x]) ? lab[i] : (uint16_t)ORC_IGN;
}

/* _calc_var of pyopencl_small.py:140-185 for the seed at (z,y,x). */
static float orc_calc_var(const uint16_t *lab, const uint8_t *fz, const uint8_t *fy,
                          const uint8_t *fx, const void *img, int dtype,
                          int Y, int X, int z, int y, int x)
{
    const int64_t sy = X, sz = (int64_t)X * Y;
    const int64_t i0 = z * sz + y * sy + x;
    const uint16_t seg = orc_mask(lab, fz, fy, fx, z, y, x, i0);
    const float B = orc_img(img, dtype, i0);
    float dev = 0.0f, summe = 0.0f;
    for (int n = -1; n < 2; ++n)            /* XY plane */
        for (int o = -1; o < 2; ++o) {
            int64_t i = i0 + n * sy + o;
            if (orc_mask(lab, fz, fy, fx, z, y + n, x + o, i) == seg) {
                float t = B - orc_img(img, dtype, i);
                dev = fmaf(t, t, dev); summe += 1.0f;
            }
        }
    for (int n = -1; n <= 1; n += 2)        /* "XZ": (z+n, y, x+o) */
        for (int o = -1; o < 2; ++o) {
            int64_t i = i0 + n * sz + o;
            if (summe < 9.0f && orc_mask(lab, fz, fy, fx, z + n, y, x + o, i) == seg) {
                float t = B - orc_img(img, dtype, i);
                dev = fmaf(t, t, dev); summe += 1.0f;
            }
        }
    for (int n = -1; n <= 1; n += 2)        /* "YZ": (z+n, y+o, x) */
        for (int o = -1; o <= 1; o += 2) {
            int64_t i = i0 + n * sz + o * sy;
            if (summe < 9.0f && orc_mask(lab, fz, fy, fx, z + n, y + o, x, i) == seg) {
                float t = B - orc_img(img, dtype, i);
                dev = fmaf(t, t, dev); summe += 1.0f;
            }
        }
    float var = dev / summe;
    if (var < 1.0f) var = 1.0f;
    return var;
}

/* Seeds: every interior voxel of a labelled plane whose label is not        */
/* negative (pyopencl_small.py:205-207), in ascending voxel order.           */
/* Returns the seed count; with pos == NULL only counts.                     */
int64_t orc_seeds(const uint16_t *lab, const void *img, int dtype, int Z, int Y, int X,
                  const uint8_t *fz, const uint8_t *fy, const uint8_t *fx,
                  int64_t *pos, uint8_t *label, float *var)
{
    const int64_t sy = X, sz = (int64_t)X * Y;
    int64_t n = 0;
    for (int z = 1; z < Z - 1; ++z)
        for (int y = 1; y < Y - 1; ++y)
            for (int x = 1; x < X - 1; ++x) {
                int64_t i = z * sz + y * sy + x;
                uint16_t m = orc_mask(lab, fz, fy, fx, z, y, x, i);
                if (m == ORC_IGN) continue;
                if (pos) {
                    pos[n] = i;
                    label[n] = (uint8_t)m;
                    var[n] = orc_calc_var(lab, fz, fy, fx, img, dtype, Y, X, z, y, x);
                }
                ++n;
            }
    return n;
}

/* ------------------------------------------------------------------------ */
/* Direction chain of pyopencl_small.py:285-290.  Returns 0..5 for           */
/* +z,-z,+y,-y,+x,-x, or -1 when no branch matches (rand > W5).              */
/* ------------------------------------------------------------------------ */
static inline int orc_direction(float rand, const float W[6])
{
    if (rand < W[0] || rand == 0.0f) return 0;
    if (rand >= W[0] && rand < W[1]) return 1;
    if (rand >= W[1] && rand < W[2]) return 2;
    if (rand >= W[2] && rand < W[3]) return 3;
    if (rand >= W[3] && rand < W[4]) return 4;
    if (rand >= W[4] && rand <= W[5]) return 5;
    return -1;
}

/* Six weights at voxel i, prefix-summed in the order W0..W5 (:241-252). */
static inline void orc_weights(const void *img, int dtype, int64_t i, int64_t sy, int64_t sz,
                               float B, float div1, float W[6])
{
    W[0] = orc_weight(B, orc_img(img, dtype, i + sz), div1);
    W[1] = orc_weight(B, orc_img(img, dtype, i - sz), div1);
    W[2] = orc_weight(B, orc_img(img, dtype, i + sy), div1);
    W[3] = orc_weight(B, orc_img(img, dtype, i - sy), div1);
    W[4] = orc_weight(B, orc_img(img, dtype, i + 1), div1);
    W[5] = orc_weight(B, orc_img(img, dtype, i - 1), div1);
    W[1] += W[0];
    W[2] += W[1];
    W[3] += W[2];
    W[4] += W[3];
    W[5] += W[4];
}

static const int ORC_DZ[6] = {1, -1, 0, 0, 0, 0};
static const int ORC_DY[6] = {0, 0, 1, -1, 0, 0};
static const int ORC_DX[6] = {0, 0, 0, 0, 1, -1};

/* ------------------------------------------------------------------------ */
/* Walker, Philox stream.  One independent walk per (seed, w): counter       */
/* (pos_lo, pos_hi, w, step/4), key (key_lo, key_hi); step s uses word s%4.  */
/* u = (word >> 8) * 2^-24 in [0,1); rand = W5 * u.                          */
/* hits: uint32 [nlabels][V].  Returns the number of accepted moves.          */
/* ------------------------------------------------------------------------ */
int64_t orc_philox_walk(const void *img, int dtype, int Z, int Y, int X,
                        const int64_t *pos, const uint8_t *label, const float *var,
                        int64_t nseeds, int nbrw, int sorw, uint64_t key,
                        uint32_t *hits)
{
    const int64_t sy = X, sz = (int64_t)X * Y, V = sz * Z;
    const uint32_t kk[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    int64_t accepted = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+:accepted)
    for (int64_t s = 0; s < nseeds; ++s) {
        const int64_t p0 = pos[s];
        const int z0 = (int)(p0 / sz), y0 = (int)((p0 / sy) % Y), x0 = (int)(p0 % X);
        const float B = orc_img(img, dtype, p0);
        const float div1 = 1.0f / (2.0f * var[s]);
        uint32_t *h = hits + (int64_t)label[s] * V;
        for (int w = 0; w < nbrw; ++w) {
            int64_t i = p0;
            int k = z0, l = y0, m = x0;
            uint32_t r[4] = {0, 0, 0, 0};
            for (int step = 0; step < sorw; ++step) {
                if ((step & 3) == 0) {
                    const uint32_t c[4] = {(uint32_t)p0, (uint32_t)((uint64_t)p0 >> 32),
                                           (uint32_t)w, (uint32_t)(step >> 2)};
                    orc_philox4x32_10(c, kk, r);
                }
                float W[6];
                orc_weights(img, dtype, i, sy, sz, B, div1, W);
                float u = (float)(r[step & 3] >> 8) * 5.9604644775390625e-08f;
                float rand = W[5] * u;
                int d = orc_direction(rand, W);
                if (d < 0) d = 5;   /* unreachable: u < 1 implies rand <= W5 */
                int kn = k + ORC_DZ[d], ln = l + ORC_DY[d], mn = m + ORC_DX[d];
                if (kn > 0 && kn < Z - 1 && ln > 0 && ln < Y - 1 && mn > 0 && mn < X - 1) {
                    k = kn; l = ln; m = mn;
                    i = k * sz + l * sy + m;
#pragma omp atomic
                    h[i] += 1u;
                    ++accepted;
                }
            }
        }
    }
    return accepted;
}

/* ------------------------------------------------------------------------ */
/* Walker, reference stream (pyopencl_small.py:209-311), all labels in one    */
/* pass over the seeds: each seed only ever writes its own label's hits, so  */
/* this equals the per-label launches of :128-134.  The recurrence state is  */
/* seeded with the voxel index (as float) and runs across all nbrw walks.    */
/* The direction registers start at (0,0,0) (uninitialised in the kernel)    */
/* and keep their value when no branch of the chain matches.                 */
/* ------------------------------------------------------------------------ */
/* One step of the reference's float MRG32k3a-like recurrence
 * (pyopencl_small.py:213-223, :256-282) as its GPU compilers build it: both
 * backends compile with floating-point contraction on (OpenCL C's FP_CONTRACT
 * defaults to ON, nvcc's --fmad to true), so a*b - c*d becomes
 * fma(a, b, -(c*d)) and p -= k*m becomes fma(-k, m, p), the forms clang's
 * contraction gives (the left product fused).  Written without contraction
 * the recurrence loses most of its state bits: on a uniform image its walks
 * drift by tens of voxels per 4000 steps.  s = {s10, s11, s12, s20, s21, s22};
 * returns the uniform (p1 - p2 [+ m1]) * norm, rand = W5 * that. */
static float orc_mrg_next(float s[6])
{
    const float norm = 2.328306549295728e-10f;
    const float m1 = 4294967087.0f, m2 = 4294944443.0f;
    const float a12 = 1403580.0f, a13n = 810728.0f, a21 = 527612.0f, a23n = 1370589.0f;
    float p1 = fmaf(a12, s[1], -(a13n * s[0]));
    int64_t k1 = (int64_t)(p1 / m1);
    p1 = fmaf(-(float)k1, m1, p1);
    if (p1 < 0.0f) p1 += m1;
    s[0] = s[1]; s[1] = s[2]; s[2] = p1;
    float p2 = fmaf(a21, s[5], -(a23n * s[3]));
    k1 = (int64_t)(p2 / m2);
    p2 = fmaf(-(float)k1, m2, p2);
    if (p2 < 0.0f) p2 += m2;
    s[3] = s[4]; s[4] = s[5]; s[5] = p2;
    return (p1 <= p2) ? ((p1 - p2 + m1) * norm) : ((p1 - p2) * norm);
}

/* The first n uniforms of the stream seeded with voxel index `index`. */
void orc_mrg_stream(int64_t index, int n, float *u)
{
    const float fi = (float)(int32_t)index;
    float s[6] = {fi, fi, fi, fi, fi, fi};
    for (int t = 0; t < n; ++t) u[t] = orc_mrg_next(s);
}

int64_t orc_ref_walk(const void *img, int dtype, int Z, int Y, int X,
                     const int64_t *pos, const uint8_t *label, const float *var,
                     int64_t nseeds, int nbrw, int sorw, uint32_t *hits)
{
    const int64_t sy = X, sz = (int64_t)X * Y, V = sz * Z;
    int64_t accepted = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+:accepted)
    for (int64_t s = 0; s < nseeds; ++s) {
        const int64_t p0 = pos[s];
        const int z0 = (int)(p0 / sz), y0 = (int)((p0 / sy) % Y), x0 = (int)(p0 % X);
        uint32_t *h = hits + (int64_t)label[s] * V;
        const float fi = (float)(int32_t)p0;
        float st[6] = {fi, fi, fi, fi, fi, fi};
        const float B = orc_img(img, dtype, p0);
        const float div1 = 1.0f / (2.0f * var[s]);
        int k = z0, l = y0, m = x0;
        int64_t i = p0;
        int n = 0, o = 0, p = 0;
        int step = 0, n_rw = 0;
        while (n_rw < nbrw) {
            float W[6];
            orc_weights(img, dtype, i, sy, sz, B, div1, W);
            const float rand = W[5] * orc_mrg_next(st);
            int d = orc_direction(rand, W);
            if (d >= 0) { n = ORC_DZ[d]; o = ORC_DY[d]; p = ORC_DX[d]; }
            if (k + n > 0 && k + n < Z - 1 && l + o > 0 && l + o < Y - 1 && m + p > 0 && m + p < X - 1) {
                k += n; l += o; m += p;
                i = k * sz + l * sy + m;
#pragma omp atomic
                h[i] += 1u;
                ++accepted;
            }
            step += 1;
            if (step == sorw) {
                k = z0; l = y0; m = x0;
                i = p0;
                n_rw += 1;
                step = 0;
            }
        }
    }
    return accepted;
}

/* ------------------------------------------------------------------------ */
/* Argmax merge (rw_small.py:200-201 + get_labels, biomedisa_helper.py:      */
/* 1209-1213): first maximum wins; index -> label value.                     */
/* ------------------------------------------------------------------------ */
void orc_argmax_u32(const uint32_t *hits, int nlabels, int64_t V,
                    const uint8_t *values, uint8_t *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < V; ++v) {
        int best = 0;
        uint32_t bv = hits[v];
        for (int l = 1; l < nlabels; ++l) {
            uint32_t h = hits[(int64_t)l * V + v];
            if (h > bv) { bv = h; best = l; }
        }
        out[v] = values[best];
    }
}

void orc_argmax_f32(const float *phi, int nlabels, int64_t V,
                    const uint8_t *values, uint8_t *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < V; ++v) {
        int best = 0;
        float bv = phi[v];
        for (int l = 1; l < nlabels; ++l) {
            float h = phi[(int64_t)l * V + v];
            if (h > bv) { bv = h; best = l; }
        }
        out[v] = values[best];
    }
}

/* ------------------------------------------------------------------------ */
/* Curvature smoothing: CURVATURE_SRC then UPDATE_SRC                        */
/* (opencl_kernels.py:136-224), `iters` times, interior voxels only.         */
/* tmp: scratch of V floats.                                                 */
/* ------------------------------------------------------------------------ */
void orc_smooth(float *phi, float *curv, int Z, int Y, int X, int iters)
{
    const int64_t xsh = X, flat = (int64_t)X * Y;
    for (int it = 0; it < iters; ++it) {
#pragma omp parallel for schedule(static)
        for (int z = 1; z < Z - 1; ++z)
            for (int y = 1; y < Y - 1; ++y)
                for (int x = 1; x < X - 1; ++x) {
                    const int64_t index = z * flat + y * xsh + x;
                    const float eps = 1e-10f;
                    float dx = (phi[index - 1] - phi[index + 1]) * 0.5f;
                    float dxx = (phi[index - 1] - 2.0f * phi[index] + phi[index + 1]);
                    float dx2 = dx * dx;
                    float dy = (phi[index - xsh] - phi[index + xsh]) * 0.5f;
                    float dyy = (phi[index - xsh] - 2.0f * phi[index] + phi[index + xsh]);
                    float dy2 = dy * dy;
                    float dz = (phi[index - flat] - phi[index + flat]) * 0.5f;
                    float dzz = (phi[index - flat] - 2.0f * phi[index] + phi[index + flat]);
                    float dz2 = dz * dz;
                    float dxy = (phi[index - xsh - 1] + phi[index + xsh + 1]
                                 - phi[index - xsh + 1] - phi[index + xsh - 1]) * 0.25f;
                    float dxz = (phi[index + flat - 1] + phi[index - flat + 1]
                                 - phi[index + flat + 1] - phi[index - flat - 1]) * 0.25f;
                    float dyz = (phi[index + flat - xsh] + phi[index - flat + xsh]
                                 - phi[index + flat + xsh] - phi[index - flat - xsh]) * 0.25f;
                    curv[index] = (dxx * (dy2 + dz2)
                                   + dyy * (dx2 + dz2)
                                   + dzz * (dx2 + dy2)
                                   - 2.0f * dx * dy * dxy
                                   - 2.0f * dx * dz * dxz
                                   - 2.0f * dy * dz * dyz)
                                  / (dx2 + dy2 + dz2 + eps);
                }
#pragma omp parallel for schedule(static)
        for (int z = 1; z < Z - 1; ++z)
            for (int y = 1; y < Y - 1; ++y)
                for (int x = 1; x < X - 1; ++x) {
                    const int64_t index = z * flat + y * xsh + x;
                    phi[index] += 0.01f * curv[index];
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Uncertainty: MAX_SRC per label in label order, then UNCERTAINTY_SRC       */
/* (opencl_kernels.py:54-133), interior voxels; out = u (float).             */
/* Voxels outside the interior keep the last label's hits                    */
/* (the kernel writes into that label's buffer, rw_small.py:170-175).        */
/* ------------------------------------------------------------------------ */
void orc_uncertainty(const float *hits, int nlabels, int Z, int Y, int X, float *out)
{
    const int64_t flat = (int64_t)X * Y, V = flat * Z;
    memcpy(out, hits + (int64_t)(nlabels - 1) * V, (size_t)V * sizeof(float));
#pragma omp parallel for schedule(static)
    for (int z = 1; z < Z - 1; ++z)
        for (int y = 1; y < Y - 1; ++y)
            for (int x = 1; x < X - 1; ++x) {
                const int64_t index = z * flat + (int64_t)y * X + x;
                float m0 = 0.0f, mid = 0.0f, mx = 0.0f;
                for (int l = 0; l < nlabels; ++l) {
                    float tmp = hits[(int64_t)l * V + index];
                    if (tmp > mx) { m0 = mid; mid = mx; mx = tmp; }
                    else if (tmp > mid) { m0 = mid; mid = tmp; }
                    else if (tmp > m0) { m0 = tmp; }
                }
                float max2 = mx;
                if (max2 == 0.0f) max2 = 1.0f;
                float t = m0 / max2;
                float t1 = mid / max2;
                out[index] = 1.0f - (1.0f - t) * (1.0f - t1);
            }
}

/* The `found` test of the large-volume kernel, as written in
 * pyopencl_large.py:570-588: a seed at (plane, row, column) walks if
 * (column + plane) and (row + plane) are even; otherwise it scans the offsets
 * -100..100 in steps of 5 along each axis and walks if an interior voxel there
 * holds a label of the seeds volume (labelled planes; ignore elsewhere) that is
 * neither its own nor the ignore value.  out[k] = 1 if seed k walks. */
void orc_adaptive(const uint16_t *lab, int Z, int Y, int X, const uint8_t *fz, const uint8_t *fy,
                  const uint8_t *fx, const int64_t *pos, int64_t n, uint8_t *out)
{
    const int64_t sy = X, sz = (int64_t)X * Y;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < n; ++k) {
        const int64_t p = pos[k];
        const int plane = (int)(p / sz), row = (int)((p % sz) / sy), column = (int)(p % sy);
        const uint16_t segment = orc_mask(lab, fz, fy, fx, plane, row, column, p);
        int found = 0;
        if ((column + plane) % 2 == 0 && (row + plane) % 2 == 0) {
            found = 1;
        } else {
            for (int z = -100; z < 101; z = z + 5)
                for (int y = -100; y < 101; y = y + 5)
                    for (int x = -100; x < 101; x = x + 5)
                        if (plane + z > 0 && row + y > 0 && column + x > 0 && plane + z < Z - 1 &&
                            row + y < Y - 1 && column + x < X - 1) {
                            const int64_t q = (int64_t)(plane + z) * sz + (int64_t)(row + y) * sy + column + x;
                            const uint16_t v = orc_mask(lab, fz, fy, fx, plane + z, row + y, column + x, q);
                            if (v != segment && v != ORC_IGN) found = 1;
                        }
        }
        out[k] = (uint8_t)found;
    }
}

/* 6-connected components of mask (non-zero = foreground), numbered 1, 2, ...
 * in the raster order of their first voxel (breadth-first from each unvisited
 * foreground voxel in raster order); labels[i] = 0 for background.  Returns the
 * component count.  queue: scratch of Z*Y*X int64. */
int64_t orc_label6(const uint8_t *mask, int Z, int Y, int X, int32_t *labels, int64_t *queue)
{
    const int64_t sy = X, sz = (int64_t)X * Y, V = sz * Z;
    for (int64_t i = 0; i < V; ++i) labels[i] = 0;
    int32_t n = 0;
    for (int64_t s = 0; s < V; ++s) {
        if (!mask[s] || labels[s]) continue;
        ++n;
        int64_t head = 0, tail = 0;
        queue[tail++] = s;
        labels[s] = n;
        while (head < tail) {
            const int64_t i = queue[head++];
            const int z = (int)(i / sz), y = (int)((i % sz) / sy), x = (int)(i % sy);
            const int64_t nb[6] = {x > 0 ? i - 1 : -1, x < X - 1 ? i + 1 : -1, y > 0 ? i - sy : -1,
                                   y < Y - 1 ? i + sy : -1, z > 0 ? i - sz : -1, z < Z - 1 ? i + sz : -1};
            for (int k = 0; k < 6; ++k) {
                const int64_t j = nb[k];
                if (j >= 0 && mask[j] && !labels[j]) {
                    labels[j] = n;
                    queue[tail++] = j;
                }
            }
        }
    }
    return n;
}

/* evolution of curvop_numba.py:32-57, as written: B = A, then every interior
 * voxel takes the first neighbour label (27 positions, dz/dy/dx -1..1) with a
 * strictly smaller cost |mean[l] - img| (alpha x that for label 0) than the
 * best so far, starting from its own label.  Costs in double. */
void orc_evolution(const float *mean, const uint8_t *A, const uint8_t *data, int Z, int Y, int X, double alpha,
                   uint8_t *B)
{
    const int64_t sy = X, sz = (int64_t)X * Y, V = sz * Z;
    for (int64_t i = 0; i < V; ++i) B[i] = A[i];
#pragma omp parallel for schedule(static)
    for (int k = 1; k < Z - 1; ++k)
        for (int l = 1; l < Y - 1; ++l)
            for (int m = 1; m < X - 1; ++m) {
                const int64_t i = k * sz + l * sy + m;
                const uint8_t refLabel = A[i];
                uint8_t finLabel = refLabel;
                const double img = data[i];
                double ref = refLabel == 0 ? alpha * fabs((double)mean[refLabel] - img)
                                           : fabs((double)mean[refLabel] - img);
                for (int n = -1; n < 2; ++n)
                    for (int o = -1; o < 2; ++o)
                        for (int p = -1; p < 2; ++p) {
                            const uint8_t tmpLabel = A[i + n * sz + o * sy + p];
                            if (tmpLabel != refLabel) {
                                const double val = tmpLabel == 0 ? alpha * fabs((double)mean[tmpLabel] - img)
                                                                 : fabs((double)mean[tmpLabel] - img);
                                if (val < ref) {
                                    ref = val;
                                    finLabel = tmpLabel;
                                }
                            }
                        }
                B[i] = finLabel;
            }
}

/* The planar patterns _P3 of curvop (curvop_numba.py:130-138), P3[t][k][l][m]. */
static void orc_patterns(int P3[9][3][3][3])
{
    for (int t = 0; t < 9; ++t)
        for (int k = 0; k < 3; ++k)
            for (int l = 0; l < 3; ++l)
                for (int m = 0; m < 3; ++m) P3[t][k][l][m] = 0;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            P3[0][a][b][1] = 1;
            P3[1][a][1][b] = 1;
            P3[2][1][a][b] = 1;
            P3[3][a][b][b] = 1;
            P3[4][a][b][2 - b] = 1;
            P3[5][b][a][b] = 1;
            P3[6][b][a][2 - b] = 1;
            P3[7][b][b][a] = 1;
            P3[8][b][2 - b][a] = 1;
        }
}

/* erosion and dilation of curvop_numba.py:59-121, as written (reads start,
 * writes final). */
void orc_erosion(const uint8_t *start, uint8_t *final_, int zsh, int ysh, int xsh, int label)
{
    int P3[9][3][3][3];
    orc_patterns(P3);
    const int64_t sy = xsh, sz = (int64_t)xsh * ysh;
#pragma omp parallel for schedule(static)
    for (int plane = 1; plane < zsh - 1; ++plane)
        for (int row = 1; row < ysh - 1; ++row)
            for (int column = 1; column < xsh - 1; ++column) {
                const int64_t c = plane * sz + row * sy + column;
                int found = 0;
                for (int n = -1; n < 2; ++n)
                    for (int o = -1; o < 2; ++o)
                        for (int p = -1; p < 2; ++p)
                            if (start[c + n * sz + o * sy + p] != label) found = 1;
                if (start[c] == label && found == 1) {
                    int t = 0;
                    found = 0;
                    while (t < 9 && found == 0) {
                        long value = 0;
                        for (int k = 0; k < 3; ++k)
                            for (int l = 0; l < 3; ++l)
                                for (int m = 0; m < 3; ++m)
                                    if (P3[t][k][l][m] == 1)
                                        value += labs((long)start[c + (k - 1) * sz + (l - 1) * sy + (m - 1)] - label);
                        if (value == 0) found = 1;
                        t += 1;
                    }
                    if (found == 0) {
                        int subLabel = label;
                        for (int n = -1; n < 2; ++n)
                            for (int o = -1; o < 2; ++o)
                                for (int p = -1; p < 2; ++p) {
                                    const int tmpLabel = start[c + n * sz + o * sy + p];
                                    if (tmpLabel != label) subLabel = tmpLabel;
                                }
                        final_[c] = (uint8_t)subLabel;
                    }
                }
            }
}

void orc_dilation(const uint8_t *start, uint8_t *final_, int zsh, int ysh, int xsh, int label)
{
    int P3[9][3][3][3];
    orc_patterns(P3);
    const int64_t sy = xsh, sz = (int64_t)xsh * ysh;
#pragma omp parallel for schedule(static)
    for (int plane = 1; plane < zsh - 1; ++plane)
        for (int row = 1; row < ysh - 1; ++row)
            for (int column = 1; column < xsh - 1; ++column) {
                const int64_t c = plane * sz + row * sy + column;
                int found = 0;
                for (int n = -1; n < 2; ++n)
                    for (int o = -1; o < 2; ++o)
                        for (int p = -1; p < 2; ++p)
                            if (start[c + n * sz + o * sy + p] == label) found = 1;
                if (start[c] != label && found == 1) {
                    int t = 0;
                    found = 0;
                    while (t < 9 && found == 0) {
                        long value = 0;
                        for (int k = 0; k < 3; ++k)
                            for (int l = 0; l < 3; ++l)
                                for (int m = 0; m < 3; ++m) {
                                    if (P3[t][k][l][m] == 1 &&
                                        start[c + (k - 1) * sz + (l - 1) * sy + (m - 1)] != label)
                                        value += 0;
                                    else
                                        value += 1;
                                }
                        if (value == 0) found = 1;
                        t += 1;
                    }
                    if (found == 0) final_[c] = (uint8_t)label;
                }
            }
}

/* geodis of active_contour.py:42-89, as written: `iterations` forward and
 * backward raster scans in place; every candidate in double, the minimum stored
 * as float. */
void orc_geodis(float *c, int zsh, int ysh, int xsh, double sqrt2, double sqrt3, int iterations)
{
    const int64_t sy = xsh, sz = (int64_t)xsh * ysh;
#define C3(z, y, x) ((double)c[(int64_t)(z) * sz + (int64_t)(y) * sy + (x)])
    for (int it = 0; it < iterations; ++it) {
        for (int z = 1; z < zsh; ++z)
            for (int y = 1; y < ysh - 1; ++y)
                for (int x = 1; x < xsh - 1; ++x) {
                    double m = C3(z, y, x);
                    const double a[13] = {C3(z - 1, y - 1, x - 1) + sqrt3, C3(z - 1, y - 1, x) + sqrt2,
                                          C3(z - 1, y - 1, x + 1) + sqrt3, C3(z - 1, y, x - 1) + sqrt2,
                                          C3(z - 1, y, x) + 1, C3(z - 1, y, x + 1) + sqrt2,
                                          C3(z - 1, y + 1, x - 1) + sqrt3, C3(z - 1, y + 1, x) + sqrt2,
                                          C3(z - 1, y + 1, x + 1) + sqrt3, C3(z, y - 1, x - 1) + sqrt2,
                                          C3(z, y - 1, x) + 1, C3(z, y - 1, x + 1) + sqrt2, C3(z, y, x - 1) + 1};
                    for (int k = 0; k < 13; ++k)
                        if (a[k] < m) m = a[k];
                    c[(int64_t)z * sz + (int64_t)y * sy + x] = (float)m;
                }
        for (int z = zsh - 2; z > -1; --z)
            for (int y = ysh - 2; y > 0; --y)
                for (int x = xsh - 2; x > 0; --x) {
                    double m = C3(z, y, x);
                    const double a[13] = {C3(z + 1, y - 1, x - 1) + sqrt3, C3(z + 1, y - 1, x) + sqrt2,
                                          C3(z + 1, y - 1, x + 1) + sqrt3, C3(z + 1, y, x - 1) + sqrt2,
                                          C3(z + 1, y, x) + 1, C3(z + 1, y, x + 1) + sqrt2,
                                          C3(z + 1, y + 1, x - 1) + sqrt3, C3(z + 1, y + 1, x) + sqrt2,
                                          C3(z + 1, y + 1, x + 1) + sqrt3, C3(z, y + 1, x + 1) + sqrt2,
                                          C3(z, y + 1, x) + 1, C3(z, y + 1, x - 1) + sqrt2, C3(z, y, x + 1) + 1};
                    for (int k = 0; k < 13; ++k)
                        if (a[k] < m) m = a[k];
                    c[(int64_t)z * sz + (int64_t)y * sy + x] = (float)m;
                }
    }
#undef C3
}

/* smooth_img_3x3 of features/biomedisa_helper.py:309-323, as written; dtype 0 = uint8
 * (the mean truncated into uint8), 1 = float32. */
void orc_smooth_img(const void *img, int dtype, int zsh, int ysh, int xsh, void *out)
{
    const int64_t sy = xsh, sz = (int64_t)xsh * ysh;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < zsh; ++z)
        for (int y = 0; y < ysh; ++y)
            for (int x = 0; x < xsh; ++x) {
                double tmp = 0;
                int i = 0;
                for (int k = -1; k < 2; ++k)
                    for (int l = -1; l < 2; ++l)
                        for (int m = -1; m < 2; ++m)
                            if (0 <= z + k && z + k < zsh && 0 <= y + l && y + l < ysh && 0 <= x + m && x + m < xsh) {
                                const int64_t q = (z + k) * sz + (y + l) * sy + x + m;
                                tmp += dtype == 0 ? (double)((const uint8_t *)img)[q] : (double)((const float *)img)[q];
                                i += 1;
                            }
                const int64_t c = z * sz + y * sy + x;
                if (dtype == 0)
                    ((uint8_t *)out)[c] = (uint8_t)(tmp / i);
                else
                    ((float *)out)[c] = (float)(tmp / i);
            }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_expf_array(const float *x, float *y, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) y[i] = orc_expf(x[i]);
}

void orc_philox_array(const uint32_t *ctr, const uint32_t *key, uint32_t *out, int64_t n)
{
    for (int64_t i = 0; i < n; ++i) orc_philox4x32_10(ctr + 4 * i, key + 2 * i, out + 4 * i);
}
