/*
 * bmx.h -- C ABI of libbmx_hip.so, the MI355X smart-interpolation engine.
 *
 * This library replaces the platform backends and device kernels of
 * biomedisa's smart interpolation (biomedisa/features/random_walk/
 * pyopencl_small.py, pycuda_small*.py, opencl_kernels.py, gpu_kernels.py)
 * and the host merge of rw_small.py.  Each entry point names the reference
 * interface it stands in for.
 *
 * Conventions
 *   - Volumes are C-contiguous (Z, Y, X) arrays in device memory (allocate
 *     with bmx_malloc, fill with bmx_upload).  V = Z*Y*X must be < 2^32.
 *   - Label volumes hold a label INDEX per voxel (uint8): 0 = background,
 *     1..nlabels-1 = the sorted non-zero label values, BMX_LABEL_NONE (0xFF)
 *     = negative (ignored) label: at most 255 label indices (0..254).  A
 *     volume with all 256 label values 0..255 (and no negative label) uses
 *     every index as a label, with the *_lab entry points' label_none = 256;
 *     the walk, merge, smoothing and uncertainty take nlabels up to 256.
 *   - Images are uint8 (BMX_IMG_U8) or float32 in [0,255] (BMX_IMG_F32);
 *     the reference normalises every other dtype to the latter
 *     (biomedisa/interpolation.py:135-144).
 *   - Return value 0 = success, negative = error; bmx_last_error() gives the
 *     message of the last error on the calling thread.
 *   - A context owns one HIP stream; calls on one context are ordered on it.
 *     Contexts are not shared between threads.
 */
#ifndef BMX_H
#define BMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BMX_OK 0
#define BMX_E_INVALID (-1)
#define BMX_E_OOM (-2)
#define BMX_E_DEVICE (-3)

#define BMX_IMG_U8 0
#define BMX_IMG_F32 1

/* Random streams of the walker. */
#define BMX_RNG_PHILOX 0 /* counter-based Philox4x32-10 per (seed voxel, walk) */
#define BMX_RNG_REF 1    /* the reference's float MRG recurrence per seed     */

#define BMX_LABEL_NONE 0xFF

typedef struct bmx_ctx bmx_ctx;

/* One walker seed: voxel index, label index, intensity variance of the
 * seed's neighbourhood (_calc_var, pyopencl_small.py:140-185).  `reserved` is 0
 * as written by bmx_seeds; with bmx_walk_opts.seed_overrides the tiled walker
 * reads it as (walks | hit_weight << 16), see bmx_seed_walks. */
typedef struct bmx_seed {
    uint32_t pos;
    uint32_t label;
    float var;
    uint32_t reserved;
} bmx_seed;

/* Message of the last failed call on this thread ("" if none). */
const char *bmx_last_error(void);
int bmx_version(void);

/* Context: replaces the per-rank OpenCL context/queue selection of
 * _get_device (biomedisa/features/biomedisa_helper.py:1003-1013) and the
 * pycuda context of rw_small.py:63-72. */
int bmx_open(int device, bmx_ctx **out);
int bmx_close(bmx_ctx *ctx);
int bmx_sync(bmx_ctx *ctx);

/* Device memory: replaces cl.Buffer / gpuarray allocation and the blocking
 * enqueue_copy transfers of pyopencl_small.py:118-121,133. */
int bmx_malloc(bmx_ctx *ctx, size_t bytes, void **dptr);
int bmx_free(bmx_ctx *ctx, void *dptr);
int bmx_upload(bmx_ctx *ctx, void *dst_dev, const void *src_host, size_t bytes);
int bmx_download(bmx_ctx *ctx, void *dst_host, const void *src_dev, size_t bytes);
int bmx_zero(bmx_ctx *ctx, void *dst_dev, size_t bytes);
/* Device-to-device copy on the context stream. */
int bmx_copy(bmx_ctx *ctx, void *dst_dev, const void *src_dev, size_t bytes);

/* ------------------------------------------------------------------------ */
/* Pre- and post-processing of one call on the device.  The caller's arrays  */
/* are uploaded as they are (raw dtypes below); the cropped, padded image    */
/* and label-index volume of the walk are written from them.                 */
/* ------------------------------------------------------------------------ */
#define BMX_RAW_U8 0
#define BMX_RAW_I8 1
#define BMX_RAW_U16 2
#define BMX_RAW_I16 3
#define BMX_RAW_U32 4
#define BMX_RAW_I32 5
#define BMX_RAW_F32 6

/* Label classes: the values 0..255, then "negative", then "above 255". */
#define BMX_LABEL_CLASSES 258

/* Label set and label boxes of a raw label volume (integer dtype, device,
 * Z*Y*X < 2^32): counts[c] = voxels of class c (counts[0] is left 0: zero
 * voxels are not counted) and boxes[6c..6c+5] = zmin, zmax, ymin, ymax, xmin,
 * xmax of its voxels (zmin > zmax when absent), host arrays of
 * BMX_LABEL_CLASSES and 6*BMX_LABEL_CLASSES entries.  Replaces the unique()
 * of pre_processing (biomedisa/features/biomedisa_helper.py:82-103, 668-717)
 * and the non-zero scan of predict_blocksize (:1105-1124), whose box is the
 * union of the boxes of the classes that stay non-zero. */
int bmx_label_stats(bmx_ctx *ctx, const void *lab, int lab_dtype, int64_t Z, int64_t Y, int64_t X,
                    uint64_t *counts, int32_t *boxes);

/* Smallest and largest value of a raw image (device, n < 2^32 elements), for
 * the float32 normalisation of interpolation.py:140-144. */
int bmx_image_range(bmx_ctx *ctx, const void *img, int img_raw_dtype, int64_t n, double *vmin, double *vmax);

/* The walk's image: the 1-voxel zero pad of interpolation.py:157-158 and the
 * crop to the label box of :203, in one pass from the raw (Z,Y,X) image.
 * box = [z0, z1, y0, y1, x0, x1) in padded coordinates; out has
 * (z1-z0)(y1-y0)(x1-x0) voxels.  normalize = 0: uint8 copied, int8 + 128 as
 * uint8 (:136-139), float32 copied; normalize = 1: float32
 * ((float)v - vmin) / vrange * 255 in that order of IEEE operations, with
 * vmin = amin and vrange = amax of (image - vmin) as float32 (:140-144); pad
 * voxels are 0 either way. */
int bmx_crop_image(bmx_ctx *ctx, const void *img, int img_raw_dtype, int normalize, int64_t Z, int64_t Y,
                   int64_t X, const int64_t *box, float vmin, float vrange, void *out);

/* The walk's label-index volume: pad and crop as bmx_crop_image, every voxel
 * mapped through lut[BMX_LABEL_CLASSES] (host; class as for bmx_label_stats,
 * pad voxels as value 0).  Replaces the label edits of pre_processing
 * (values > 255, `only` and `ignore` set to 0: biomedisa_helper.py:682-717)
 * and the label volume rebuild of pyopencl_small.py:89-107 (label index per
 * voxel, negative = ignored = BMX_LABEL_NONE). */
int bmx_crop_labels(bmx_ctx *ctx, const void *lab, int lab_dtype, int64_t Z, int64_t Y, int64_t X,
                    const int64_t *box, const uint8_t *lut, uint8_t *out);

/* A result of the cropped walk volume placed into the caller's (Z,Y,X)
 * volume (device): out = crop where the box covers the padded voxel, else 0.
 * Replaces the result placement of rw_small.py:212-216 (zeros, box slice,
 * [1:-1] on every axis). */
int bmx_place(bmx_ctx *ctx, const uint8_t *crop, int64_t Z, int64_t Y, int64_t X, const int64_t *box,
              uint8_t *out);

/* Counts of the sanity check of rw_small.py:203-211, Dice_score(labels,
 * result * (labels > 0)) (biomedisa_helper.py:155-165), over n voxels of the
 * label-index volume and the regular result (label values): counts[0] =
 * labelled voxels (index 1..nlabels-1) whose result equals their value v and
 * counted[v] != 0 (the caller clears counted[v] where (labels + result) wraps
 * to 0 in the label dtype), counts[1] = labelled voxels, counts[2] = labelled
 * voxels with a non-zero result; Dice = 2 c0 / (c1 + c2). */
int bmx_dice_counts(bmx_ctx *ctx, const uint8_t *labidx, const uint8_t *result, int64_t n, int nlabels,
                    const uint8_t *values, const uint8_t *counted, uint64_t *counts);

/* Labelled planes: replaces read_labeled_slices / read_labeled_slices_allx /
 * read_indices_allx (biomedisa_helper.py:1015-1066).  flags (device, Z+Y+X
 * bytes) receives 1 for every labelled z-plane, then y-plane, then x-plane. */
int bmx_plane_flags(bmx_ctx *ctx, const uint8_t *lab, int64_t Z, int64_t Y, int64_t X,
                    int allaxis, uint8_t *flags);

/* bmx_plane_flags with large = 1: the large-volume path's planes, replacing
 * read_labeled_slices_large (biomedisa_helper.py:1068-1080: scanning z, a
 * labelled plane hides the 4 planes after it) and read_labeled_slices_allx_large
 * (:1082-1103: as allaxis, but only positive labels count). */
int bmx_plane_flags_ex(bmx_ctx *ctx, const uint8_t *lab, int64_t Z, int64_t Y, int64_t X,
                       int allaxis, int large, uint8_t *flags);

/* Adaptive walks of the large-volume path: replaces the `found` test of the
 * randomWalk kernel in pyopencl_large.py:570-590.  mask[p] (uint8 over the
 * volume) = 1 where a seed at p walks: x+z and y+z even, or an interior voxel
 * on the stride-5 grid within +-100 voxels of p (per axis) lies on a labelled
 * plane with a label other than p's.  tmp: device scratch of 4*Z*Y*X bytes.
 * Select the walking seeds with bmx_select_seeds. */
int bmx_adaptive_mask(bmx_ctx *ctx, const uint8_t *lab, int64_t Z, int64_t Y, int64_t X,
                      const uint8_t *flags, void *tmp, uint8_t *mask);

/* The three entry points above and bmx_seeds with the label volume's index
 * for ignored voxels given: label_none = BMX_LABEL_NONE (what the others
 * assume) or 256 when all 256 label indices are labels (a volume using every
 * label value 0..255 and no negative label).  Replace the same reference code;
 * the 256-label case is the reference's switch to short labels with -1 as
 * ignore value (pyopencl_small.py:90-95). */
int bmx_plane_flags_lab(bmx_ctx *ctx, const uint8_t *lab, int label_none, int64_t Z, int64_t Y, int64_t X,
                        int allaxis, int large, uint8_t *flags);
int bmx_adaptive_mask_lab(bmx_ctx *ctx, const uint8_t *lab, int label_none, int64_t Z, int64_t Y, int64_t X,
                          const uint8_t *flags, void *tmp, uint8_t *mask);
int bmx_seeds_lab(bmx_ctx *ctx, const uint8_t *lab, int label_none, const void *img, int img_dtype, int64_t Z,
                  int64_t Y, int64_t X, const uint8_t *flags, bmx_seed *seeds, int64_t capacity,
                  int64_t *nseeds);

/* Seed extraction: replaces the label-volume rebuild of pyopencl_small.py:
 * 89-107 and the per-work-item seed test and _calc_var of :140-185, :205-228.
 * Writes the seeds in ascending voxel order.  seeds == NULL counts only.
 * *nseeds receives the count; capacity bounds the writes. */
int bmx_seeds(bmx_ctx *ctx, const uint8_t *lab, const void *img, int img_dtype,
              int64_t Z, int64_t Y, int64_t X, const uint8_t *flags,
              bmx_seed *seeds, int64_t capacity, int64_t *nseeds);

/* Keep the seeds whose voxel has mask[pos] != 0 (uint8 mask over the
 * volume), in unspecified order; *nout receives the count.  out must hold
 * nseeds entries.  No counterpart in the reference: used by the
 * surrogate-guided mode to walk only seeds that can reach uncertain voxels. */
int bmx_select_seeds(bmx_ctx *ctx, const bmx_seed *seeds, int64_t nseeds, const uint8_t *mask,
                     bmx_seed *out, int64_t *nout);

/* Random walks of every seed, all labels in one launch: replaces the
 * per-label loop fill -> randomWalk -> D2H of pyopencl_small.py:128-134 and
 * the randomWalk kernel :192-314.  hits (device, uint32 [nlabels][V]) is
 * ACCUMULATED into (zero it first with bmx_zero).  rng selects the stream;
 * key is the Philox key (ignored for BMX_RNG_REF).  *accepted (may be NULL)
 * receives the number of accepted moves = the hits added.  The tiled walker
 * (chosen automatically from 2^16 walkers up) walks 2^32 walkers (nseeds *
 * nbrw) or more in chunks of whole seeds, with the same hits; the sorted
 * walker refuses them. */
int bmx_walk(bmx_ctx *ctx, const void *img, int img_dtype, int64_t Z, int64_t Y, int64_t X,
             const bmx_seed *seeds, int64_t nseeds, int nlabels, int nbrw, int sorw,
             int rng, uint64_t key, uint32_t *hits, uint64_t *accepted);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the Philox walk of bmx_walk_ex with algorithm BMX_WALK_DIRECT
 * (identical hits and accepted moves), whose packed neighbour records and
 * move counter live in a static device buffer of BMX_NOALLOC_SCRATCH_BYTES per
 * GPU instead of context memory: 256 bytes plus V * 8 bytes for uint8 images
 * or V * 32 bytes for float32, so volumes up to 8 Mi voxels (uint8) or 2 Mi
 * voxels (float32); larger ones are rejected with BMX_E_INVALID.  Calls no
 * hipMalloc, hipFree or allocating routine.  Every call on one GPU uses that
 * GPU's buffer: calls that overlap on one GPU (from several contexts or
 * threads) race.  hits is accumulated into, as for bmx_walk. */
#define BMX_NOALLOC_SCRATCH_BYTES (64u << 20)
int bmx_walk_noalloc(bmx_ctx *ctx, const void *img, int img_dtype, int64_t Z, int64_t Y, int64_t X,
                     const bmx_seed *seeds, int64_t nseeds, int nlabels, int nbrw, int sorw, uint64_t key,
                     uint32_t *hits, uint64_t *accepted);

/* Walker algorithms (identical hit counts; they differ only in speed). */
#define BMX_WALK_AUTO 0   /* tiled for >= 2^16 walkers, hit weights or per-seed overrides
                             (Philox), direct otherwise */
#define BMX_WALK_DIRECT 1 /* one launch, each thread walks its whole walk */
#define BMX_WALK_SORTED 2 /* lockstep segments, walkers re-sorted by brick between them */
#define BMX_WALK_TILED 3  /* tile-owner rounds, hit counters privatised in LDS */

typedef struct bmx_walk_opts {
    int algorithm;     /* BMX_WALK_* */
    int segment_steps; /* sorted: steps per launch; tiled: max steps per walker per round;
                          multiple of 4 (0 = 128 sorted, 256 tiled) */
    int brick_log2;    /* sorted: log2 of the sort-brick side in voxels (0 = 3) */
    int hit_weight;    /* tiled: count added per accepted move (0 = 1); with nbrw = n/k walks
                          and weight k, a seed subset's hits estimate its n-walk hits */
    int seed_overrides; /* tiled: 1 = each seed's `reserved` gives its walk count (low 16
                           bits, 0 = nbrw, at most nbrw) and a hit weight (high 16 bits,
                           0 = 1) multiplying hit_weight */
    int tile_shape;     /* tiled: tile shape index in [0, BMX_TILE_SHAPES) (core side, halo,
                           workgroup size; 0 = the default shape) */
    int graph_rounds;   /* tiled: rounds per replay of one captured HIP graph whose walk
                           kernel is persistent (resident workgroups take buckets from a
                           device counter), replayed until every walk has ended; k > 0 =
                           k rounds rounded up to even, 0 = 8; -1 = rounds driven by the
                           host, one walk workgroup launched per bucket */
} bmx_walk_opts;

/* Tile shapes compiled into the tiled walker. */
#define BMX_TILE_SHAPES 6

/* Geometry of tile shape `shape` for images of type img_dtype (BMX_IMG_*):
 * core side and halo in voxels, threads per workgroup, bytes per LDS hit
 * counter (2: 16-bit counters, two workgroups per CU for u8; 4: 32-bit) and
 * LDS bytes per workgroup.  No counterpart in the reference, whose kernel has
 * no tiles; for the autotuner and for callers choosing tile_shape. */
int bmx_tile_shape_info(int img_dtype, int shape, int *core, int *halo, int *threads, int *hit_bytes,
                        int *lds_bytes);

/* bmx_walk with a choice of algorithm (opts may be NULL = automatic).  No
 * counterpart in the reference, which has one kernel per backend. */
int bmx_walk_ex(bmx_ctx *ctx, const void *img, int img_dtype, int64_t Z, int64_t Y, int64_t X,
                const bmx_seed *seeds, int64_t nseeds, int nlabels, int nbrw, int sorw, int rng,
                uint64_t key, uint32_t *hits, uint64_t *accepted, const bmx_walk_opts *opts);

/* One volume of a batched walk.  Its image occupies elements [offset, offset +
 * Z*Y*X) of the batch image (z-major, x fastest); its seeds are
 * seeds[seed_offset, seed_offset + nseeds) with volume-local positions and
 * labels below nlabels, as bmx_seeds writes them; its hits accumulate into the
 * uint32 block [nlabels][Z*Y*X] at element hits_offset of the batch hits. */
typedef struct bmx_volume {
    uint64_t offset;
    int32_t Z, Y, X;
    int32_t nlabels;
    uint64_t hits_offset;
    uint64_t seed_offset;
    uint64_t nseeds;
} bmx_volume;

/* Random walks of many volumes of uneven size and label count in one tiled
 * walk: per volume the hits equal those of bmx_walk_ex with the same key and
 * the Philox stream (bit-exact).  Volumes are given in ascending, disjoint
 * offset and seed order, the batch image below 2^32 elements.  opts as for
 * bmx_walk_ex (algorithm must be AUTO or TILED).  Replaces the per-volume
 * calls a caller of smart_interpolation (biomedisa/interpolation.py:77-203)
 * would make; no counterpart in the reference. */
int bmx_walk_batch(bmx_ctx *ctx, const void *img, int img_dtype, const bmx_volume *vols, int nvol,
                   const bmx_seed *seeds, int nbrw, int sorw, uint64_t key, uint32_t *hits,
                   uint64_t *accepted, const bmx_walk_opts *opts);

/* Milliseconds of the last bmx_walk (whole walk phase: record packing
 * excluded, sorting included), measured with HIP events on the context
 * stream. */
float bmx_last_walk_ms(bmx_ctx *ctx);

/* Timing of the last bmx_walk: whole phase, sum of the walker-kernel
 * launches alone, number of walker-kernel launches. */
int bmx_last_walk_stats(bmx_ctx *ctx, float *phase_ms, float *kernel_ms, int *launches);

/* Per-voxel argmax over labels, first maximum wins, mapped through
 * values[nlabels] (label index -> label value): replaces np.argmax +
 * get_labels (rw_small.py:200-201, biomedisa_helper.py:1209-1213). */
int bmx_argmax(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int64_t V,
               const uint8_t *values, uint8_t *out);

/* Curvature smoothing of one float32 volume in place: `iters` rounds of
 * CURVATURE_SRC then UPDATE_SRC (opencl_kernels.py:136-224), fused into one
 * stencil pass per round.  tmp: device scratch of V floats. */
int bmx_smooth(bmx_ctx *ctx, float *phi, float *tmp, int64_t Z, int64_t Y, int64_t X, int iters);

/* The 'smooth' result of rw_small.py:114-140: every label's hits smoothed
 * `iters` times, then first-maximum argmax mapped through values. */
int bmx_smooth_argmax(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int64_t Z, int64_t Y,
                      int64_t X, int iters, const uint8_t *values, uint8_t *out);

/* bmx_smooth_argmax with clamp_first = 1: the large-volume path's 'smooth'
 * (pyopencl_large.py:455-469), where the first label's smoothed map has its
 * negative values set to 0 before the running maximum (max_to_label, :233-241:
 * strictly greater wins, as the first-maximum argmax). */
int bmx_smooth_argmax_ex(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int64_t Z, int64_t Y,
                         int64_t X, int iters, const uint8_t *values, uint8_t *out, int clamp_first);

/* Outlier removal: replaces clean (biomedisa/features/remove_outlier.py:60-95)
 * on a label VALUE volume (device uint8, Z*Y*X).  For every present value but
 * the smallest, 6-connected components of that value smaller than or equal to
 * threshold x its largest component become 0.  out may equal in. */
int bmx_clean(bmx_ctx *ctx, const uint8_t *in, int64_t Z, int64_t Y, int64_t X, double threshold,
              uint8_t *out);

/* Hole filling: replaces fill (remove_outlier.py:97-139).  For every present
 * value k but the smallest, in ascending order and on the volume as changed by
 * the values before it: in the box of k grown by one voxel, 6-connected
 * components of "not k" other than the largest entry of [voxels of k,
 * component sizes] become k unless larger than threshold x (voxels of k);
 * when the voxels of k are that largest entry they become 0.  out may equal in. */
int bmx_fill(bmx_ctx *ctx, const uint8_t *in, int64_t Z, int64_t Y, int64_t X, double threshold,
             uint8_t *out);

/* HxByteRLE codec of AmiraMesh byte lattices (host memory; no device work),
 * replacing byterle_decoder / byterle_encoder
 * (features/amira_to_np/amira_data_stream.py:21-141).  decode: n_out bytes
 * from n_in encoded bytes, returns n_out or a negative BMX_E_*; encode: needs
 * cap >= n + ceil(n / 127), returns the encoded length. */
int64_t bmx_hxbyterle_decode(const uint8_t *in, int64_t n_in, uint8_t *out, int64_t n_out);
int64_t bmx_hxbyterle_encode(const uint8_t *in, int64_t n, uint8_t *out, int64_t cap);
/* PNG scanline unfiltering (host memory), for slice stacks read by load_data
 * (biomedisa_helper.py:505-591): rows x (1 + rowbytes) filtered bytes -> rows x
 * rowbytes; bpp = bytes per complete pixel (1..8). */
int bmx_png_unfilter(const uint8_t *in, int64_t rows, int64_t rowbytes, int bpp, uint8_t *out);

/* Image denoising: replaces smooth_img_3x3 (features/biomedisa_helper.py:309-323), the
 * mean of each voxel's 3x3x3 neighbourhood clipped to the volume (summed in
 * double, stored in the image's type; uint8 truncates).  out != img. */
int bmx_denoise(bmx_ctx *ctx, const void *img, int img_dtype, int64_t Z, int64_t Y, int64_t X, void *out);

/* Measurement aid, no counterpart in the reference: `threads` threads (a
 * multiple of 256) run iters x 128 independent float32 fused multiply-adds
 * each into out[threads]; *ms receives the kernel time.  Its known count of
 * full-rate VALU instructions (threads * iters * 2 wave-level instructions)
 * scales VALU counters read in the same profiling run. */
int bmx_valu_calibrate(bmx_ctx *ctx, float *out, int64_t threads, int iters, float *ms);

/* Counts of the 256 values of a uint8 device array (host uint64[256] out);
 * below 2^32 per value.  Used for the label sets of the post-processing. */
int bmx_histogram(bmx_ctx *ctx, const uint8_t *vol, int64_t n, uint64_t *counts);

/* Active contour: replaces the acwe loop of activeContour
 * (biomedisa/features/active_contour.py:158-169) with evolution and curvop
 * (curvop_numba.py:32-148).  img (uint8) and lab (label values) are the
 * padded, cropped volumes the reference works on; labels[0..nlabels) its
 * allLabels (ascending).  steps times: label means, evolution (alpha weights
 * label 0), then `smooth` erosion/dilation rounds per label.  out receives the
 * labels; tmp is device scratch of Z*Y*X bytes. */
int bmx_acwe(bmx_ctx *ctx, const uint8_t *img, const uint8_t *lab, int64_t Z, int64_t Y, int64_t X,
             const uint8_t *labels, int nlabels, double alpha, int smooth, int steps, uint8_t *out,
             uint8_t *tmp);

/* Refinement (active_contour.py:193-225, activeContour with simple=True), in
 * three steps around a host-side decision per plane:
 *   bmx_refine_distance: dist (float32, Z*Y*X) = geodis of (lab == 0 ? 100 : 0),
 *     the chamfer scans of active_contour.py:42-89 reproduced bit for bit;
 *     d (uint8) = dist <= 10 ? 0 : 1.
 *   bmx_refine_stats: per z-plane and label value, over the voxels with d == 0,
 *     count, sum and sum of squares of img: stats[z][256][3] (uint64).
 *   bmx_refine_apply: out = (d == 0 and lut[z][img]) ? label[z] : 0, with the
 *     host's per-plane label (0 = none) and uint8 lut[z][256]. */
int bmx_refine_distance(bmx_ctx *ctx, const uint8_t *lab, int64_t Z, int64_t Y, int64_t X, float *dist,
                        uint8_t *d);
int bmx_refine_stats(bmx_ctx *ctx, const uint8_t *img, const uint8_t *lab, const uint8_t *d, int64_t Z,
                     int64_t Y, int64_t X, uint64_t *stats);
int bmx_refine_apply(bmx_ctx *ctx, const uint8_t *img, const uint8_t *d, int64_t Z, int64_t Y, int64_t X,
                     const uint8_t *lut, const uint8_t *label, uint8_t *out);

/* The 'uncertainty' result of rw_small.py:152-176 (MAX_SRC + UNCERTAINTY_SRC,
 * opencl_kernels.py:54-133): uint8(255 * (1 - (1-m0/m2)(1-m1/m2))) from the
 * three largest hit counts, interior voxels; 0 elsewhere. */
int bmx_uncertainty(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int64_t Z, int64_t Y,
                    int64_t X, uint8_t *out);

/* ------------------------------------------------------------------------ */
/* Learned proposal (surrogate) kernels.  No counterpart in the reference:   */
/* a 3-D conv encoder-decoder that labels confident voxels so the walker can */
/* skip seeds far from uncertain ones.  Activations: bf16 (uint16 bits),     */
/* channels-last [N][D][H][W][C]; conv weights bf16 [Co][27][Ci]; FP32       */
/* accumulation on MFMA; channel counts multiples of 16.                     */
/* ------------------------------------------------------------------------ */
int bmx_conv3d_fwd(bmx_ctx *ctx, const uint16_t *x, int N, int D, int H, int W, int Ci,
                   const uint16_t *w, const float *b, uint16_t *y, int Co, int relu);
/* dx = conv^T(dy, w) [* (xact > 0) if xact]; wt_scratch: Co*27*Ci bf16. */
/* Forward convolution of concat(a, up2(b)) (b: Cb channels on the half-size grid,
 * nearest 2x upsampling), without materialising either: the decoder input of the
 * encoder-decoder at inference time.  Same output as bmx_up2 + bmx_concat +
 * bmx_conv3d_fwd.  No counterpart in the reference. */
int bmx_conv3d_fwd_cat(bmx_ctx *ctx, const uint16_t *a, int N, int D, int H, int W, int Ca, const uint16_t *b,
                       int Cb, const uint16_t *w, const float *bias, uint16_t *y, int Co, int relu);
int bmx_conv3d_dgrad(bmx_ctx *ctx, const uint16_t *dy, int N, int D, int H, int W, int Co,
                     const uint16_t *w, const uint16_t *xact, uint16_t *dx, int Ci,
                     uint16_t *wt_scratch);
/* dw += sum_v dy x(shifted), db += sum_v dy (FP32, zero them first). */
int bmx_conv3d_wgrad(bmx_ctx *ctx, const uint16_t *x, int N, int D, int H, int W, int Ci,
                     const uint16_t *dy, int Co, float *dw, float *db);
int bmx_pool2(bmx_ctx *ctx, const uint16_t *x, int N, int D, int H, int W, int C, uint16_t *y);
int bmx_pool2_bwd(bmx_ctx *ctx, const uint16_t *dy, int N, int D, int H, int W, int C, uint16_t *dx);
int bmx_up2(bmx_ctx *ctx, const uint16_t *x, int N, int D, int H, int W, int C, uint16_t *y);
int bmx_up2_bwd(bmx_ctx *ctx, const uint16_t *dy, int N, int D, int H, int W, int C, uint16_t *dx);
int bmx_concat(bmx_ctx *ctx, const uint16_t *a, int Ca, const uint16_t *b, int Cb, int64_t nvox,
               uint16_t *y);
int bmx_split(bmx_ctx *ctx, const uint16_t *y, int64_t nvox, uint16_t *a, int Ca, uint16_t *b,
              int Cb, int accumulate_a);
int bmx_relu_mask(bmx_ctx *ctx, uint16_t *g, const uint16_t *a, int64_t n);
/* 1x1x1 softmax head over 16 channels, L <= 32 outputs.  With target (soft
 * labels [v][L]) or target_idx (one label index per voxel) also dx, dW, db
 * (accumulated) and the summed cross-entropy (accumulated into *loss); vmask
 * (may be NULL) excludes voxels. */
int bmx_head(bmx_ctx *ctx, const uint16_t *x, int64_t nvox, int L, const float *W, const float *b,
             float *prob, const float *target, const uint8_t *target_idx, const uint8_t *vmask, float scale,
             uint16_t *dx, float *dW, float *db, float *loss);
int bmx_adam(bmx_ctx *ctx, float *p, const float *g, float *m, float *v, uint16_t *pbf, int64_t n,
             float lr, float beta1, float beta2, float eps, int step, float weight_decay);
int bmx_f32_to_bf16(bmx_ctx *ctx, const float *a, uint16_t *b, int64_t n);
/* y += x elementwise (bf16). */
int bmx_add_bf16(bmx_ctx *ctx, uint16_t *y, const uint16_t *x, int64_t n);
/* Training targets on a padded grid from walker hit counts.
 * hard = 0: t[v][l] = hits[l][v] / sum_l hits[.][v]; vmask = 1 where the voxel is
 *           inside (Z,Y,X) and some walker arrived.
 * hard = 1: t[v] = one-hot of the walker's argmax (first maximum; index 0 when all
 *           counts are 0, as bmx_argmax); vmask = 1 inside (Z,Y,X). */
/* Training labels on a padded grid: idx = first argmax index of the hits (0 when
 * all counts are 0), vmask = 1 inside (Z,Y,X).  No counterpart in the reference. */
int bmx_hits_index(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int Z, int Y, int X, int Zp, int Yp, int Xp,
                   uint8_t *idx, uint8_t *vmask);
int bmx_hits_target(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int Z, int Y, int X, int Zp,
                    int Yp, int Xp, int L, int hard, float *target, uint8_t *vmask);

/* Surrogate-guided walking (no counterpart in the reference). */
/* Network input on a padded grid (Zp,Yp,Xp) >= (Z,Y,X), 32 bf16 channels:
 * image/127.5-1, seed indicator, one-hot label index (< 30) on seed voxels. */
int bmx_net_input(bmx_ctx *ctx, const void *img, int img_dtype, const uint8_t *lab,
                  const uint8_t *flags, int Z, int Y, int X, int Zp, int Yp, int Xp, uint16_t *out);
/* Head on the padded grid -> per voxel of (Z,Y,X): argmax label index and
 * uncertainty flag (top softmax probability < thr). */
int bmx_head_predict(bmx_ctx *ctx, const uint16_t *x, int Z, int Y, int X, int Zp, int Yp, int Xp,
                     int L, const float *W, const float *b, float thr, uint8_t *pred, uint8_t *unc);
/* U = unc | dilate(label boundaries of pred, brad); need = dilate(U, srad)
 * (Chebyshev dilations); tmp: V bytes of scratch. */
int bmx_guided_masks(bmx_ctx *ctx, const uint8_t *pred, const uint8_t *unc, int Z, int Y, int X,
                     int brad, int srad, uint8_t *U, uint8_t *need, uint8_t *tmp);
/* out = U ? values[argmax hits] : values[pred]. */
/* seeds[i].reserved = walks | weight << 16 for i < n (per-seed walk count and hit
 * weight read by bmx_walk_ex with seed_overrides).  No counterpart in the reference. */
int bmx_seed_walks(bmx_ctx *ctx, bmx_seed *seeds, int64_t n, int walks, int weight);
/* ring = dilate(need, r) & ~need (Chebyshev r): the seeds of the far-field ring. */
int bmx_ring_mask(bmx_ctx *ctx, const uint8_t *need, int Z, int Y, int X, int r, uint8_t *ring, uint8_t *tmp);
int bmx_merge_guided(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int64_t V,
                     const uint8_t *values, const uint8_t *pred, const uint8_t *U, uint8_t *out);
/* a = a & ~b over n mask bytes (0/1), e.g. to split the far-field ring
 * (no counterpart in the reference). */
int bmx_mask_andnot(bmx_ctx *ctx, uint8_t *a, const uint8_t *b, int64_t n);
/* Repair of the guided result (no counterpart in the reference).
 * D = !U & (sum of hits >= minhits, and > 0) & (first argmax of hits != pred). */
int bmx_guided_disagree(bmx_ctx *ctx, const uint32_t *hits, int nlabels, int64_t V, const uint8_t *pred,
                        const uint8_t *U, uint32_t minhits, uint8_t *D);
/* D <- dilate(D, r) & ~U (the repaired voxels U2); need = dilate(U2, srad). */
int bmx_repair_masks(bmx_ctx *ctx, uint8_t *D, const uint8_t *U, int Z, int Y, int X, int r, int srad,
                     uint8_t *need, uint8_t *tmp);
/* On U2: hits[:, v] = hits2[:, v] and U[v] = 1. */
int bmx_repair_hits(bmx_ctx *ctx, uint32_t *hits, const uint32_t *hits2, int nlabels, int64_t V,
                    const uint8_t *U2, uint8_t *U);
/* out = U ? a : values[pred]  (a holds label values, e.g. a smoothed result). */
int bmx_select_labels(bmx_ctx *ctx, const uint8_t *a, const uint8_t *pred, const uint8_t *U,
                      int nlabels, const uint8_t *values, int64_t V, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif /* BMX_H */
