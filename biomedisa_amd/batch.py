"""Batched smart interpolation: many volumes of uneven size and label count in
one tiled walk.

The reference interpolates one volume per call (biomedisa/interpolation.py:43,
with one OpenCL launch per label, pyopencl_small.py:128-134); a caller with many
volumes loops over it.  Here the volumes share one device image, one seed list
and one hit buffer, and bmx_walk_batch walks every seed of every volume in the
same rounds, so small volumes fill the GPU together.  A seed's Philox stream
depends only on its volume-local position, the walk key and the walk index, so
each volume's hits - and labels - equal those of smart_interpolation on that
volume alone with the same key.

    smart_interpolation_batch(datas, labelDatas, ...) -> [result dict, ...]

Per volume the checks, pre-processing and result dictionary are those of
smart_interpolation; the walk parameters (nbrw, sorw, allaxis, smooth, key)
are the batch's.
"""
import time

import numpy as np

from . import _hip
from .helpers import adjacent_planes
from .interpolation import DEFAULT_KEY, PipelineError, _prepare, _results

_ALIGN = 256   # element alignment of each volume in the batch image


class DeviceBatch:
    """Images, label-index volumes, seeds and the volume table of a batch on the device.

    items: [(img, labidx, nlabels, allaxis)] with img uint8 or float32 (Z,Y,X) and
    labidx the uint8 label-index volume of the same shape.  When the images mix
    uint8 and float32 the uint8 ones are widened to float32 (same values, same
    walk).
    """

    def __init__(self, eng, items):
        if not items:
            raise ValueError("empty batch")
        f32 = any(img.dtype != np.uint8 for img, _, _, _ in items)
        self.eng = eng
        self.img_dtype = _hip.IMG_F32 if f32 else _hip.IMG_U8
        esize = 4 if f32 else 1
        self.dims, self.nlabels, self.offsets, self.hits_offsets = [], [], [], []
        off = hoff = 0
        for img, lab, nl, _ in items:
            if img.shape != lab.shape or img.ndim != 3:
                raise ValueError("image and label volume must have the same 3-D shape")
            if img.dtype not in (np.uint8, np.float32):
                raise TypeError("images must be uint8 or float32 after normalisation")
            Z, Y, X = (int(s) for s in img.shape)
            self.dims.append((Z, Y, X))
            self.nlabels.append(int(nl))
            self.offsets.append(off)
            self.hits_offsets.append(hoff)
            off += -(-(Z * Y * X) // _ALIGN) * _ALIGN
            hoff += int(nl) * Z * Y * X
        if off >= 2 ** 32:
            raise PipelineError("batch image exceeds 2^32 voxels")
        self.total_voxels = sum(Z * Y * X for Z, Y, X in self.dims)
        self.img = eng.alloc(max(1, off * esize))
        self.lab = eng.alloc(max(1, off))
        for (img, lab, _, _), o in zip(items, self.offsets):
            a = np.ascontiguousarray(img, dtype=np.float32 if f32 else np.uint8)
            eng.upload_into(_hip.BufferView(self.img, o * esize, a.nbytes), a)
            eng.upload_into(_hip.BufferView(self.lab, o, lab.size), np.ascontiguousarray(lab, dtype=np.uint8))
        self.hits_elems = hoff
        # plane flags and seeds per volume, into one seed list in volume order
        nflags = sum(Z + Y + X for Z, Y, X in self.dims)
        self.flags = eng.alloc(max(1, nflags))
        counts, fo = [], 0
        self._flag_views = []
        # a volume of all 256 label values uses every index as a label (no ignored index)
        self.label_none = [_hip.LABEL_ALL if nl == 256 else _hip.LABEL_NONE for nl in self.nlabels]
        for (Z, Y, X), (_, _, _, allaxis), o, ln in zip(self.dims, items, self.offsets, self.label_none):
            fv = _hip.BufferView(self.flags, fo, Z + Y + X)
            eng.plane_flags(self.view_lab(o, Z * Y * X), (Z, Y, X), allaxis, out=fv, label_none=ln)
            flags = fv.download((Z + Y + X,), np.uint8)
            if allaxis and adjacent_planes((flags[:Z], flags[Z:Z + Y], flags[Z + Y:])):
                raise PipelineError("At least one empty slice between labels required.")
            counts.append(eng.count_seeds(self.view_lab(o, Z * Y * X), (Z, Y, X), fv, label_none=ln))
            self._flag_views.append(fv)
            fo += Z + Y + X
        self.nseeds = counts
        self.seed_offsets = list(np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64))
        total = int(sum(counts))
        isz = _hip.SEED_DTYPE.itemsize
        self.seeds = eng.alloc(max(1, total) * isz)
        for k, ((Z, Y, X), o) in enumerate(zip(self.dims, self.offsets)):
            if counts[k] == 0:
                continue
            out = _hip.BufferView(self.seeds, int(self.seed_offsets[k]) * isz, counts[k] * isz)
            _, n = eng.seeds(self.view_lab(o, Z * Y * X), self.view_img(o, Z * Y * X), self.img_dtype, (Z, Y, X),
                             self._flag_views[k], capacity=counts[k], out=out, label_none=self.label_none[k])
            assert n == counts[k]
        self.table = np.zeros(len(items), dtype=_hip.VOLUME_DTYPE)
        for k, (Z, Y, X) in enumerate(self.dims):
            self.table[k] = (self.offsets[k], Z, Y, X, self.nlabels[k], self.hits_offsets[k],
                             self.seed_offsets[k], counts[k])

    def view_img(self, off, n):
        esize = 4 if self.img_dtype == _hip.IMG_F32 else 1
        return _hip.BufferView(self.img, off * esize, n * esize)

    def view_lab(self, off, n):
        return _hip.BufferView(self.lab, off, n)

    def free(self):
        for b in (self.img, self.lab, self.flags, self.seeds):
            b.free()


def run_batch(batch, values, nbrw=10, sorw=4000, smooth=0, key=DEFAULT_KEY, walk_opts=None, timings=None,
              hits=None):
    """Walk every volume of `batch` in one bmx_walk_batch call, then the argmax
    (and smoothing) per volume.  values: per volume the uint8 label values.
    Returns [{'regular': view, 'smooth': view}] over one output buffer (kept
    alive as out[k]['_buffer'])."""
    eng = batch.eng
    t = timings if timings is not None else {}
    opts = dict(walk_opts or {})
    t0 = time.perf_counter()
    own_hits = hits is None
    if own_hits:
        hits = eng.alloc(max(1, batch.hits_elems) * 4)
    hits.zero()
    accepted = eng.walk_batch(batch.img, batch.img_dtype, batch.table, batch.seeds, nbrw, sorw, key, hits,
                              **opts)
    eng.sync()
    t1 = time.perf_counter()
    phase_ms, kernel_ms, launches = eng.last_walk_stats()
    outs = []
    regular = eng.alloc(max(1, batch.total_voxels))
    sm = eng.alloc(max(1, batch.total_voxels)) if smooth else None
    vo = 0
    for k, (Z, Y, X) in enumerate(batch.dims):
        V = Z * Y * X
        h = _hip.BufferView(hits, batch.hits_offsets[k] * 4, batch.nlabels[k] * V * 4)
        r = _hip.BufferView(regular, vo, V)
        eng.argmax(h, batch.nlabels[k], V, values[k], r)
        o = {"regular": r, "_buffer": regular}
        if smooth:
            s = _hip.BufferView(sm, vo, V)
            eng.smooth_argmax(h, batch.nlabels[k], (Z, Y, X), smooth, values[k], s)
            o["smooth"] = s
            o["_sbuffer"] = sm
        outs.append(o)
        vo += V
    eng.sync()
    t2 = time.perf_counter()
    if own_hits:
        hits.free()
    t.update(walk_s=t1 - t0, merge_s=t2 - t1, walk_phase_ms=phase_ms, walk_kernel_ms=kernel_ms,
             walk_launches=launches, accepted=accepted, nseeds=int(sum(batch.nseeds)))
    return outs


def smart_interpolation_batch(datas, labelDatas, nbrw=10, sorw=4000, allaxis=False, smooth=0,
                              key=DEFAULT_KEY, device=0, tune=True, walk_opts=None, timings=None):
    """smart_interpolation over a list of volumes in one batched walk.

    Returns one entry per volume: its smart_interpolation result dict, or None
    where that volume's input is rejected (after printing the error, as
    smart_interpolation does).  tune=True picks the walk's tile shape, round
    cap and round driver with the walk autotuner (cached per workload class);
    walk_opts gives them explicitly instead.
    """
    if len(datas) != len(labelDatas):
        raise ValueError("datas and labelDatas differ in length")
    preps = [_prepare(d, l, nbrw, sorw, False, None, allaxis, "none", "all", smooth, False, False)
             for d, l in zip(datas, labelDatas)]
    live = [k for k, p in enumerate(preps) if p is not None]
    results = [None] * len(preps)
    if not live:
        return results
    eng = _hip.engine(device)
    items = [(preps[k].img, preps[k].labidx, preps[k].bm.nol, bool(allaxis)) for k in live]
    batch = DeviceBatch(eng, items)
    try:
        opts = walk_opts
        if opts is None and tune:
            from .autotune import tune_walk
            opts = tune_walk(batch, nbrw, sorw, key)
        outs = run_batch(batch, [preps[k].values for k in live], nbrw, sorw, smooth, key, opts, timings)
        for k, o in zip(live, outs):
            dev = {"regular": o["regular"]}
            if "smooth" in o:
                dev["smooth"] = o["smooth"]
            results[k] = _results(preps[k], dev)
        for o in outs[:1]:
            o["_buffer"].free()
            if "_sbuffer" in o:
                o["_sbuffer"].free()
    finally:
        batch.free()
    return results
