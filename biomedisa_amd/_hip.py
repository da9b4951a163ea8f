"""ctypes binding of libbmx_hip.so (C ABI declared in include/bmx.h).

The library is built in place by ``biomedisa_amd.build`` (``make -C
biomedisa_amd/csrc``).  Loading fails loudly when it is missing: there is no
CPU fallback in the product path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbmx_hip.so")

BMX_OK = 0
NOALLOC_SCRATCH_BYTES = 64 << 20   # BMX_NOALLOC_SCRATCH_BYTES
BMX_E_INVALID, BMX_E_OOM, BMX_E_DEVICE = -1, -2, -3
IMG_U8, IMG_F32 = 0, 1
RNG_PHILOX, RNG_REF = 0, 1
WALK_AUTO, WALK_DIRECT, WALK_SORTED, WALK_TILED = 0, 1, 2, 3
LABEL_NONE = 0xFF
LABEL_ALL = 256     # label_none of a volume whose 256 indices are all labels
LABEL_CLASSES = 258          # BMX_LABEL_CLASSES: values 0..255, negative, above 255
CLASS_NEG, CLASS_BIG = 256, 257
# raw dtypes of the device pre-processing (BMX_RAW_*)
RAW_DTYPES = {np.dtype(np.uint8): 0, np.dtype(np.int8): 1, np.dtype(np.uint16): 2, np.dtype(np.int16): 3,
              np.dtype(np.uint32): 4, np.dtype(np.int32): 5, np.dtype(np.float32): 6}
RAW_U8, RAW_I8, RAW_F32 = 0, 1, 6


class WalkOpts(C.Structure):
    _fields_ = [("algorithm", C.c_int), ("segment_steps", C.c_int), ("brick_log2", C.c_int),
                ("hit_weight", C.c_int), ("seed_overrides", C.c_int), ("tile_shape", C.c_int),
                ("graph_rounds", C.c_int)]

TILE_SHAPES = 6   # BMX_TILE_SHAPES

# bmx_volume: one volume of a batched walk.
VOLUME_DTYPE = np.dtype([("offset", "<u8"), ("Z", "<i4"), ("Y", "<i4"), ("X", "<i4"), ("nlabels", "<i4"),
                         ("hits_offset", "<u8"), ("seed_offset", "<u8"), ("nseeds", "<u8")])

SEED_DTYPE = np.dtype([("pos", "<u4"), ("label", "<u4"), ("var", "<f4"), ("reserved", "<u4")])

# Every symbol include/bmx.h declares, with (restype, argtypes).
_p, _i64, _i = C.c_void_p, C.c_int64, C.c_int
SIGNATURES = {
    "bmx_last_error": (C.c_char_p, []),
    "bmx_tile_shape_info": (_i, [_i, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)]),
    "bmx_version": (_i, []),
    "bmx_open": (_i, [_i, C.POINTER(_p)]),
    "bmx_close": (_i, [_p]),
    "bmx_sync": (_i, [_p]),
    "bmx_malloc": (_i, [_p, C.c_size_t, C.POINTER(_p)]),
    "bmx_free": (_i, [_p, _p]),
    "bmx_upload": (_i, [_p, _p, _p, C.c_size_t]),
    "bmx_download": (_i, [_p, _p, _p, C.c_size_t]),
    "bmx_zero": (_i, [_p, _p, C.c_size_t]),
    "bmx_copy": (_i, [_p, _p, _p, C.c_size_t]),
    "bmx_label_stats": (_i, [_p, _p, _i, _i64, _i64, _i64, _p, _p]),
    "bmx_image_range": (_i, [_p, _p, _i, _i64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "bmx_crop_image": (_i, [_p, _p, _i, _i, _i64, _i64, _i64, _p, C.c_float, C.c_float, _p]),
    "bmx_crop_labels": (_i, [_p, _p, _i, _i64, _i64, _i64, _p, _p, _p]),
    "bmx_place": (_i, [_p, _p, _i64, _i64, _i64, _p, _p]),
    "bmx_dice_counts": (_i, [_p, _p, _p, _i64, _i, _p, _p, _p]),
    "bmx_plane_flags": (_i, [_p, _p, _i64, _i64, _i64, _i, _p]),
    "bmx_seeds": (_i, [_p, _p, _p, _i, _i64, _i64, _i64, _p, _p, _i64, C.POINTER(_i64)]),
    "bmx_plane_flags_ex": (_i, [_p, _p, _i64, _i64, _i64, _i, _i, _p]),
    "bmx_adaptive_mask": (_i, [_p, _p, _i64, _i64, _i64, _p, _p, _p]),
    "bmx_plane_flags_lab": (_i, [_p, _p, _i, _i64, _i64, _i64, _i, _i, _p]),
    "bmx_adaptive_mask_lab": (_i, [_p, _p, _i, _i64, _i64, _i64, _p, _p, _p]),
    "bmx_seeds_lab": (_i, [_p, _p, _i, _p, _i, _i64, _i64, _i64, _p, _p, _i64, C.POINTER(_i64)]),
    "bmx_smooth_argmax_ex": (_i, [_p, _p, _i, _i64, _i64, _i64, _i, _p, _p, _i]),
    "bmx_clean": (_i, [_p, _p, _i64, _i64, _i64, C.c_double, _p]),
    "bmx_fill": (_i, [_p, _p, _i64, _i64, _i64, C.c_double, _p]),
    "bmx_acwe": (_i, [_p, _p, _p, _i64, _i64, _i64, _p, _i, C.c_double, _i, _i, _p, _p]),
    "bmx_refine_distance": (_i, [_p, _p, _i64, _i64, _i64, _p, _p]),
    "bmx_refine_stats": (_i, [_p, _p, _p, _p, _i64, _i64, _i64, _p]),
    "bmx_refine_apply": (_i, [_p, _p, _p, _i64, _i64, _i64, _p, _p, _p]),
    "bmx_histogram": (_i, [_p, _p, _i64, _p]),
    "bmx_valu_calibrate": (_i, [_p, _p, _i64, _i, C.POINTER(C.c_float)]),
    "bmx_denoise": (_i, [_p, _p, _i, _i64, _i64, _i64, _p]),
    "bmx_hxbyterle_decode": (_i64, [_p, _i64, _p, _i64]),
    "bmx_hxbyterle_encode": (_i64, [_p, _i64, _p, _i64]),
    "bmx_png_unfilter": (_i, [_p, _i64, _i64, _i, _p]),
    "bmx_guided_disagree": (_i, [_p, _p, _i, _i64, _p, _p, C.c_uint32, _p]),
    "bmx_mask_andnot": (_i, [_p, _p, _p, _i64]),
    "bmx_repair_masks": (_i, [_p, _p, _p, _i, _i, _i, _i, _i, _p, _p]),
    "bmx_repair_hits": (_i, [_p, _p, _p, _i, _i64, _p, _p]),
    "bmx_select_seeds": (_i, [_p, _p, _i64, _p, _p, C.POINTER(_i64)]),
    "bmx_walk": (_i, [_p, _p, _i, _i64, _i64, _i64, _p, _i64, _i, _i, _i, _i, C.c_uint64, _p,
                      C.POINTER(C.c_uint64)]),
    "bmx_walk_ex": (_i, [_p, _p, _i, _i64, _i64, _i64, _p, _i64, _i, _i, _i, _i, C.c_uint64, _p,
                         C.POINTER(C.c_uint64), C.POINTER(WalkOpts)]),
    "bmx_walk_noalloc": (_i, [_p, _p, _i, _i64, _i64, _i64, _p, _i64, _i, _i, _i, C.c_uint64, _p,
                              C.POINTER(C.c_uint64)]),
    "bmx_walk_batch": (_i, [_p, _p, _i, _p, _i, _p, _i, _i, C.c_uint64, _p, C.POINTER(C.c_uint64),
                            C.POINTER(WalkOpts)]),
    "bmx_last_walk_ms": (C.c_float, [_p]),
    "bmx_last_walk_stats": (_i, [_p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(_i)]),
    "bmx_argmax": (_i, [_p, _p, _i, _i64, _p, _p]),
    "bmx_smooth": (_i, [_p, _p, _p, _i64, _i64, _i64, _i]),
    "bmx_smooth_argmax": (_i, [_p, _p, _i, _i64, _i64, _i64, _i, _p, _p]),
    "bmx_uncertainty": (_i, [_p, _p, _i, _i64, _i64, _i64, _p]),
    "bmx_conv3d_fwd": (_i, [_p, _p, _i, _i, _i, _i, _i, _p, _p, _p, _i, _i]),
    "bmx_conv3d_fwd_cat": (_i, [_p, _p, _i, _i, _i, _i, _i, _p, _i, _p, _p, _p, _i, _i]),
    "bmx_conv3d_dgrad": (_i, [_p, _p, _i, _i, _i, _i, _i, _p, _p, _p, _i, _p]),
    "bmx_conv3d_wgrad": (_i, [_p, _p, _i, _i, _i, _i, _i, _p, _i, _p, _p]),
    "bmx_pool2": (_i, [_p, _p, _i, _i, _i, _i, _i, _p]),
    "bmx_pool2_bwd": (_i, [_p, _p, _i, _i, _i, _i, _i, _p]),
    "bmx_up2": (_i, [_p, _p, _i, _i, _i, _i, _i, _p]),
    "bmx_up2_bwd": (_i, [_p, _p, _i, _i, _i, _i, _i, _p]),
    "bmx_concat": (_i, [_p, _p, _i, _p, _i, _i64, _p]),
    "bmx_split": (_i, [_p, _p, _i64, _p, _i, _p, _i, _i]),
    "bmx_relu_mask": (_i, [_p, _p, _p, _i64]),
    "bmx_head": (_i, [_p, _p, _i64, _i, _p, _p, _p, _p, _p, _p, C.c_float, _p, _p, _p, _p]),
    "bmx_adam": (_i, [_p, _p, _p, _p, _p, _p, _i64, C.c_float, C.c_float, C.c_float, C.c_float, _i,
                      C.c_float]),
    "bmx_f32_to_bf16": (_i, [_p, _p, _p, _i64]),
    "bmx_add_bf16": (_i, [_p, _p, _p, _i64]),
    "bmx_hits_target": (_i, [_p, _p, _i, _i, _i, _i, _i, _i, _i, _i, _i, _p, _p]),
    "bmx_hits_index": (_i, [_p, _p, _i, _i, _i, _i, _i, _i, _i, _p, _p]),
    "bmx_ring_mask": (_i, [_p, _p, _i, _i, _i, _i, _p, _p]),
    "bmx_seed_walks": (_i, [_p, _p, _i64, _i, _i]),
    "bmx_net_input": (_i, [_p, _p, _i, _p, _p, _i, _i, _i, _i, _i, _i, _p]),
    "bmx_head_predict": (_i, [_p, _p, _i, _i, _i, _i, _i, _i, _i, _p, _p, C.c_float, _p, _p]),
    "bmx_guided_masks": (_i, [_p, _p, _p, _i, _i, _i, _i, _i, _p, _p, _p]),
    "bmx_merge_guided": (_i, [_p, _p, _i, _i64, _p, _p, _p, _p]),
    "bmx_select_labels": (_i, [_p, _p, _p, _p, _i, _p, _i64, _p]),
}

_lib = None


class BmxError(RuntimeError):
    """A libbmx_hip call returned an error code."""

    def __init__(self, code, msg):
        super().__init__("libbmx_hip error %d: %s" % (code, msg))
        self.code = code


def load():
    """Load libbmx_hip.so and bind every declared entry point."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libbmx_hip.so not found at %s: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C %s/csrc`"
                % (LIB_PATH, _HERE))
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _check(rc):
    if rc != BMX_OK:
        msg = load().bmx_last_error().decode(errors="replace")
        raise BmxError(rc, msg)


# Device allocations of at least POOL_MIN_BYTES go back to their Engine's pool
# when freed and are handed out again for a request of the same size: the
# pipelines allocate the same hit volumes and masks call after call, and a
# hipMalloc of tens of GB that follows frees has taken seconds on this GPU
# (a 73 GB hit volume: 2.2 s on its third allocation, against 13 ms to clear
# it).  The pool holds at most POOL_MAX_BYTES (BMX_POOL_MAX_GB in the
# environment; 0 disables the pool), room for the two hit volumes of a guided
# 1024^3 call; an allocation that fails for lack of memory empties the pool and
# is tried once more, and Engine.release_pool() hands the memory back to the
# device for other users.  Every kernel of an Engine runs on its one stream, so
# a buffer handed out again is only written after the work queued on its
# previous owner.
POOL_MIN_BYTES = 32 << 20
POOL_MAX_BYTES = int(float(os.environ.get("BMX_POOL_MAX_GB", "160")) * (1 << 30))


class DeviceBuffer:
    """Device allocation owned by an Engine; freed (or pooled by the Engine)
    on free()/garbage collection."""

    def __init__(self, engine, nbytes):
        self.engine = engine
        self.nbytes = int(nbytes)
        self.ptr = None
        self.ptr = engine._alloc_raw(self.nbytes)

    def free(self):
        if self.ptr is not None and self.engine.ctx is not None:
            self.engine._free_raw(self.ptr, self.nbytes)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def zero(self):
        _check(load().bmx_zero(self.engine.ctx, self.ptr, self.nbytes))

    def download(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _check(load().bmx_download(self.engine.ctx, out.ctypes.data, self.ptr, out.nbytes))
        return out


class Engine:
    """One HIP device context (bmx_ctx) and its stream."""

    def __init__(self, device=0):
        lib = load()
        ctx = C.c_void_p()
        _check(lib.bmx_open(int(device), C.byref(ctx)))
        self.ctx = ctx.value
        self.device = device
        self._pool = {}      # nbytes -> device pointers free for reuse
        self._pooled = 0     # bytes held by the pool

    def close(self):
        if self.ctx is not None:
            self.release_pool()
            _check(load().bmx_close(self.ctx))
            self.ctx = None

    def _alloc_raw(self, nbytes):
        if nbytes >= POOL_MIN_BYTES:
            free = self._pool.get(nbytes)
            if free:
                self._pooled -= nbytes
                return free.pop()
        ptr = C.c_void_p()
        rc = load().bmx_malloc(self.ctx, nbytes, C.byref(ptr))
        if rc == BMX_E_OOM and self._pooled:
            self.release_pool()
            rc = load().bmx_malloc(self.ctx, nbytes, C.byref(ptr))
        _check(rc)
        return ptr.value

    def _free_raw(self, ptr, nbytes):
        if nbytes >= POOL_MIN_BYTES and self._pooled + nbytes <= POOL_MAX_BYTES:
            self._pool.setdefault(nbytes, []).append(ptr)
            self._pooled += nbytes
            return
        _check(load().bmx_free(self.ctx, ptr))

    def release_pool(self):
        """Return every pooled allocation to the device."""
        pool, self._pool, self._pooled = self._pool, {}, 0
        for ptrs in pool.values():
            for ptr in ptrs:
                _check(load().bmx_free(self.ctx, ptr))

    def pooled_bytes(self):
        return self._pooled

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync(self):
        _check(load().bmx_sync(self.ctx))

    # memory ---------------------------------------------------------------
    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        buf = DeviceBuffer(self, arr.nbytes)
        _check(load().bmx_upload(self.ctx, buf.ptr, arr.ctypes.data, arr.nbytes))
        return buf

    def upload_into(self, buf, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= buf.nbytes
        _check(load().bmx_upload(self.ctx, buf.ptr, arr.ctypes.data, arr.nbytes))

    # pre- and post-processing ---------------------------------------------
    def label_stats(self, lab, raw_dtype, dims):
        """(counts[258], boxes[258, 6]) of the label classes of a raw device label volume."""
        Z, Y, X = dims
        counts = np.zeros(LABEL_CLASSES, np.uint64)
        boxes = np.zeros((LABEL_CLASSES, 6), np.int32)
        _check(load().bmx_label_stats(self.ctx, lab.ptr, int(raw_dtype), Z, Y, X, counts.ctypes.data,
                                      boxes.ctypes.data))
        return counts, boxes

    def image_range(self, img, raw_dtype, n):
        lo, hi = C.c_double(), C.c_double()
        _check(load().bmx_image_range(self.ctx, img.ptr, int(raw_dtype), int(n), C.byref(lo), C.byref(hi)))
        return lo.value, hi.value

    def crop_image(self, img, raw_dtype, normalize, dims, box, vmin, vrange, out):
        Z, Y, X = dims
        b = np.ascontiguousarray(box, dtype=np.int64)
        _check(load().bmx_crop_image(self.ctx, img.ptr, int(raw_dtype), int(bool(normalize)), Z, Y, X,
                                     b.ctypes.data, float(vmin), float(vrange), out.ptr))

    def crop_labels(self, lab, raw_dtype, dims, box, lut, out):
        Z, Y, X = dims
        b = np.ascontiguousarray(box, dtype=np.int64)
        lut = np.ascontiguousarray(lut, dtype=np.uint8)
        assert lut.size == LABEL_CLASSES
        _check(load().bmx_crop_labels(self.ctx, lab.ptr, int(raw_dtype), Z, Y, X, b.ctypes.data,
                                      lut.ctypes.data, out.ptr))

    def place(self, crop, dims, box, out):
        Z, Y, X = dims
        b = np.ascontiguousarray(box, dtype=np.int64)
        _check(load().bmx_place(self.ctx, crop.ptr, Z, Y, X, b.ctypes.data, out.ptr))

    def dice_counts(self, labidx, result, n, nlabels, values, counted):
        values = np.ascontiguousarray(values, dtype=np.uint8)
        counted = np.ascontiguousarray(counted, dtype=np.uint8)
        assert counted.size == 256
        out = np.zeros(3, np.uint64)
        _check(load().bmx_dice_counts(self.ctx, labidx.ptr, result.ptr, int(n), int(nlabels), values.ctypes.data,
                                      counted.ctypes.data, out.ctypes.data))
        return [int(v) for v in out]

    # path -----------------------------------------------------------------
    # label_none: the label volume's index for ignored voxels (LABEL_NONE, or
    # LABEL_ALL when all 256 indices are labels)
    def plane_flags(self, lab, dims, allaxis, out=None, large=False, label_none=LABEL_NONE):
        Z, Y, X = dims
        flags = out if out is not None else DeviceBuffer(self, Z + Y + X)
        _check(load().bmx_plane_flags_lab(self.ctx, lab.ptr, int(label_none), Z, Y, X, int(bool(allaxis)),
                                          int(bool(large)), flags.ptr))
        return flags

    def adaptive_mask(self, lab, dims, flags, out=None, label_none=LABEL_NONE):
        """uint8 mask of the voxels whose seeds walk in the large-volume path."""
        Z, Y, X = dims
        V = Z * Y * X
        mask = out if out is not None else DeviceBuffer(self, V)
        tmp = DeviceBuffer(self, 4 * V)
        try:
            _check(load().bmx_adaptive_mask_lab(self.ctx, lab.ptr, int(label_none), Z, Y, X, flags.ptr, tmp.ptr,
                                                mask.ptr))
        finally:
            self.sync()
            tmp.free()
        return mask

    def count_seeds(self, lab, dims, flags, label_none=LABEL_NONE):
        Z, Y, X = dims
        n = C.c_int64()
        _check(load().bmx_seeds_lab(self.ctx, lab.ptr, int(label_none), None, IMG_U8, Z, Y, X, flags.ptr, None, 0,
                                    C.byref(n)))
        return n.value

    def seeds(self, lab, img, img_dtype, dims, flags, capacity=None, out=None, label_none=LABEL_NONE):
        """Seed list on the device -> (buffer, count)."""
        Z, Y, X = dims
        if capacity is None:
            capacity = self.count_seeds(lab, dims, flags, label_none)
        buf = out if out is not None else DeviceBuffer(self, max(1, capacity) * SEED_DTYPE.itemsize)
        n = C.c_int64()
        _check(load().bmx_seeds_lab(self.ctx, lab.ptr, int(label_none), img.ptr, img_dtype, Z, Y, X, flags.ptr,
                                    buf.ptr, capacity, C.byref(n)))
        return buf, n.value

    def select_seeds(self, seeds, nseeds, mask, out=None):
        """Seeds whose voxel has mask != 0 -> (buffer, count)."""
        buf = out if out is not None else DeviceBuffer(self, max(1, nseeds) * SEED_DTYPE.itemsize)
        n = C.c_int64()
        _check(load().bmx_select_seeds(self.ctx, seeds.ptr, int(nseeds), mask.ptr, buf.ptr,
                                       C.byref(n)))
        return buf, n.value

    def walk(self, img, img_dtype, dims, seeds, nseeds, nlabels, nbrw, sorw, rng, key, hits,
             algorithm=WALK_AUTO, segment_steps=0, brick_log2=0, hit_weight=1, seed_overrides=False,
             tile_shape=0, graph_rounds=0):
        """Accumulate the walks of `nseeds` seeds into `hits` (each accepted move adds
        `hit_weight`, times the seed's own weight with seed_overrides); returns the
        accepted moves."""
        Z, Y, X = dims
        acc = C.c_uint64()
        opts = WalkOpts(int(algorithm), int(segment_steps), int(brick_log2), int(hit_weight),
                        int(bool(seed_overrides)), int(tile_shape), int(graph_rounds))
        _check(load().bmx_walk_ex(self.ctx, img.ptr, img_dtype, Z, Y, X,
                                  seeds.ptr if seeds is not None else None, int(nseeds),
                                  int(nlabels), int(nbrw), int(sorw), int(rng),
                                  C.c_uint64(int(key) & (2**64 - 1)), hits.ptr, C.byref(acc),
                                  C.byref(opts)))
        return acc.value

    def walk_noalloc(self, img, img_dtype, dims, seeds, nseeds, nlabels, nbrw, sorw, key, hits):
        """bmx_walk_noalloc: the direct Philox walk with its scratch in the library's
        static device buffer (no allocation); returns the accepted moves."""
        Z, Y, X = dims
        acc = C.c_uint64()
        _check(load().bmx_walk_noalloc(self.ctx, img.ptr, img_dtype, Z, Y, X,
                                       seeds.ptr if seeds is not None else None, int(nseeds), int(nlabels),
                                       int(nbrw), int(sorw), C.c_uint64(int(key) & (2**64 - 1)), hits.ptr,
                                       C.byref(acc)))
        return acc.value

    def walk_batch(self, img, img_dtype, volumes, seeds, nbrw, sorw, key, hits, segment_steps=0,
                   hit_weight=1, seed_overrides=False, tile_shape=0, graph_rounds=0):
        """Tiled walk of every volume of a batch (`volumes`: VOLUME_DTYPE array, image,
        seeds and hit blocks at the offsets it gives); returns the accepted moves."""
        vols = np.ascontiguousarray(volumes, dtype=VOLUME_DTYPE)
        acc = C.c_uint64()
        opts = WalkOpts(WALK_TILED, int(segment_steps), 0, int(hit_weight), int(bool(seed_overrides)),
                        int(tile_shape), int(graph_rounds))
        _check(load().bmx_walk_batch(self.ctx, img.ptr, img_dtype, vols.ctypes.data, len(vols),
                                     seeds.ptr if seeds is not None else None, int(nbrw), int(sorw),
                                     C.c_uint64(int(key) & (2**64 - 1)), hits.ptr, C.byref(acc),
                                     C.byref(opts)))
        return acc.value

    def last_walk_ms(self):
        return float(load().bmx_last_walk_ms(self.ctx))

    def last_walk_stats(self):
        """(whole walk phase ms, walker-kernel ms, walker-kernel launches) of the last walk."""
        a, b, n = C.c_float(), C.c_float(), C.c_int()
        _check(load().bmx_last_walk_stats(self.ctx, C.byref(a), C.byref(b), C.byref(n)))
        return a.value, b.value, n.value

    def argmax(self, hits, nlabels, V, values, out):
        values = np.ascontiguousarray(values, dtype=np.uint8)
        _check(load().bmx_argmax(self.ctx, hits.ptr, int(nlabels), int(V), values.ctypes.data, out.ptr))

    def smooth(self, phi, tmp, dims, iters):
        Z, Y, X = dims
        _check(load().bmx_smooth(self.ctx, phi.ptr, tmp.ptr, Z, Y, X, int(iters)))

    def smooth_argmax(self, hits, nlabels, dims, iters, values, out, clamp_first=False):
        Z, Y, X = dims
        values = np.ascontiguousarray(values, dtype=np.uint8)
        _check(load().bmx_smooth_argmax_ex(self.ctx, hits.ptr, int(nlabels), Z, Y, X, int(iters),
                                           values.ctypes.data, out.ptr, int(bool(clamp_first))))

    def clean(self, vol, dims, threshold, out):
        """Outlier removal of a uint8 label-value volume (device) into out."""
        Z, Y, X = dims
        _check(load().bmx_clean(self.ctx, vol.ptr, Z, Y, X, float(threshold), out.ptr))

    def fill(self, vol, dims, threshold, out):
        """Hole filling of a uint8 label-value volume (device) into out."""
        Z, Y, X = dims
        _check(load().bmx_fill(self.ctx, vol.ptr, Z, Y, X, float(threshold), out.ptr))

    def denoise(self, img, dtype, dims, out):
        """3x3x3 mean filter of a uint8/float32 device image into out."""
        Z, Y, X = dims
        _check(load().bmx_denoise(self.ctx, img.ptr, int(dtype), Z, Y, X, out.ptr))

    def histogram(self, vol, n):
        """Counts of the 256 values of a uint8 device buffer."""
        counts = np.zeros(256, np.uint64)
        _check(load().bmx_histogram(self.ctx, vol.ptr, int(n), counts.ctypes.data))
        return counts

    def acwe(self, img, lab, dims, labels, alpha, smooth, steps, out, tmp):
        """Active-contour refinement of the label volume lab (device) into out."""
        Z, Y, X = dims
        labels = np.ascontiguousarray(labels, dtype=np.uint8)
        _check(load().bmx_acwe(self.ctx, img.ptr, lab.ptr, Z, Y, X, labels.ctypes.data, len(labels),
                               float(alpha), int(smooth), int(steps), out.ptr, tmp.ptr))

    def uncertainty(self, hits, nlabels, dims, out):
        Z, Y, X = dims
        _check(load().bmx_uncertainty(self.ctx, hits.ptr, int(nlabels), Z, Y, X, out.ptr))


class BufferView:
    """A sub-range of a DeviceBuffer (no ownership): .ptr / .nbytes like a buffer."""

    def __init__(self, buf, offset, nbytes=None):
        assert 0 <= offset <= buf.nbytes
        self.base = buf
        self.ptr = buf.ptr + int(offset)
        self.nbytes = buf.nbytes - int(offset) if nbytes is None else int(nbytes)
        assert offset + self.nbytes <= buf.nbytes

    def download(self, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        _check(load().bmx_download(self.base.engine.ctx, out.ctypes.data, self.ptr, out.nbytes))
        return out


def ptr(buf):
    """Device pointer of a DeviceBuffer, or None."""
    return None if buf is None else buf.ptr


def call(name, *args):
    """Call a bmx_* entry point with a context as first argument; raise on error."""
    _check(getattr(load(), name)(*args))


def tile_shape_info(img_dtype, shape):
    """Geometry of a tiled-walker shape (bmx_tile_shape_info) as a dict: core,
    halo, threads, hit_bytes (2 = 16-bit LDS counters), lds_bytes."""
    vals = [C.c_int() for _ in range(5)]
    _check(load().bmx_tile_shape_info(int(img_dtype), int(shape), *[C.byref(v) for v in vals]))
    return dict(zip(("core", "halo", "threads", "hit_bytes", "lds_bytes"), (v.value for v in vals)))


_engines = {}


def engine(device=0):
    """Process-wide Engine per device (created on first use)."""
    eng = _engines.get(device)
    if eng is None or eng.ctx is None:
        eng = Engine(device)
        _engines[device] = eng
    return eng
