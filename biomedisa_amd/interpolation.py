"""Drop-in ``smart_interpolation`` running on one MI355X through libbmx_hip.

Mirrors ``biomedisa.interpolation.smart_interpolation``
(biomedisa/interpolation.py:43-284) and the small-volume driver
``rw_small._diffusion_child`` (biomedisa/features/random_walk/rw_small.py:
39-281): same arguments, same pre-processing, same result dictionary.  The
MPI split, the per-label OpenCL launches and the host argmax are replaced by
one device pipeline per call:

    plane flags -> seeds -> walk (all labels, one launch) -> argmax
                 [-> smoothing + argmax] [-> uncertainty]

Extra keyword arguments (after the reference's):
    rng     'philox' (default; counter-based stream, walks of a seed run in
            parallel) or 'reference' (the reference's float MRG recurrence,
            one thread per seed, all walks in sequence).
    key     64-bit Philox key.
    device  HIP device ordinal.
"""
import time

import numpy as np

from . import _hip
from .helpers import (Biomedisa, Dice_score, _error_, adjacent_planes, label_index_volume,
                      normalize_image, pre_processing, predict_blocksize)

DEFAULT_KEY = 0x5EED0B10D1CE0001
# The guided walk's sparser, ring-weighted seed set runs best with 128 steps
# per walker per round (the plain walk with the library's 256).
GUIDED_WALK_OPTS = {"segment_steps": 128}
# Guided parameters that trade speed for agreement with the plain walker:
# wider near set, two ring walks, a wider repair ring (pass as `guide=`).
GUIDED_ROBUST = {"srad": 8, "far_walks": 2, "repair": 2, "repair_far_walks": 1, "repair_far": 32, "repair_srad": 8}


def guided_defaults(nbrw, allaxis):
    """Guided parameters the drop-in call uses unless `guide=` sets them.
    Labels on XY planes only are sparser than the all-axis labels the network
    was mostly trained on, and its confident errors are more frequent there:
    such volumes use a stricter confidence threshold (0.95), a wider boundary
    band (6) and less noisy ring estimates (the ring walks about a fifth of
    the walks, its inner part half of them, each count dividing nbrw).  On 8
    XY-labelled volumes (256^3, 5-12 structures, and three of the batch
    workload's) this kept the minimum per-label Dice against the plain walker
    at >= 0.995, where the all-axis defaults fell to 0.958."""
    if allaxis:
        return {}

    def divisor_at_most(q):
        return max([d for d in range(1, nbrw + 1) if nbrw % d == 0 and d <= max(q, 1)], default=1)

    return {"thr": 0.95, "brad": 6, "far_walks": divisor_at_most(nbrw // 5), "mid_walks": divisor_at_most(nbrw // 2)}

_RNG = {"philox": _hip.RNG_PHILOX, "reference": _hip.RNG_REF}


class DeviceVolume:
    """Image + label-index volume resident on the device, ready for the pipeline."""

    def __init__(self, eng, img, labidx):
        assert img.shape == labidx.shape and img.ndim == 3
        if img.dtype == np.uint8:
            self.img_dtype = _hip.IMG_U8
        elif img.dtype == np.float32:
            self.img_dtype = _hip.IMG_F32
        else:
            raise TypeError("image must be uint8 or float32 after normalisation, got %s" % img.dtype)
        self.eng = eng
        self.dims = tuple(int(s) for s in img.shape)
        self.img = eng.upload(img)
        self.lab = eng.upload(np.ascontiguousarray(labidx, dtype=np.uint8))

    @classmethod
    def from_buffers(cls, eng, img, lab, dims, img_dtype):
        """A volume whose image (uint8 or float32) and label-index buffers are
        already on the device (the volume takes them over)."""
        self = cls.__new__(cls)
        self.eng, self.img, self.lab = eng, img, lab
        self.dims = tuple(int(s) for s in dims)
        self.img_dtype = img_dtype
        return self

    @property
    def V(self):
        Z, Y, X = self.dims
        return Z * Y * X


class PipelineError(ValueError):
    """Input rejected by the pipeline (the reference's _error_ cases)."""


# rw_small.py:147-149 and :181-183: an allocation failure while smoothing or
# computing the uncertainty drops that result with a warning.
# A label volume with all 256 values 0..255 holds no index for ignored
# (negative) labels: the reference's short-label path (pyopencl_small.py:90-95)
# needs one more value than a uint8 label index has.
NEG_WITH_256_LABELS = "Negative (ignored) labels together with all 256 label values are not supported."
SMOOTH_OOM_WARNING = "Warning: GPU out of memory to allocate smooth array. Process continues without smoothing."
UNCERTAINTY_OOM_WARNING = ("Warning: GPU out of memory to allocate uncertainty array. Process continues without "
                           "uncertainty.")


def optional_stage(stage, warning):
    """Run an optional result stage; on a device out-of-memory error print the
    reference's warning and return None (every other error propagates)."""
    try:
        return stage()
    except _hip.BmxError as e:
        if e.code != _hip.BMX_E_OOM:
            raise
        print(warning)
        return None


def select_path(dims, nlabels):
    """The reference's choice between its small- and large-volume drivers
    (interpolation.py:192-196): 'small' when the cropped box as float32 times the
    label count stays below 1e10 bytes and the box below 2e9 bytes."""
    nbytes = int(np.prod(dims)) * 4
    return "small" if nbytes * nlabels < 1e10 and nbytes < 2e9 else "large"


def run_pipeline(vol, nlabels, values, nbrw=10, sorw=4000, allaxis=False, smooth=0,
                 uncertainty=False, rng="philox", key=DEFAULT_KEY, keep_hits=False, timings=None,
                 walk_opts=None, path="small"):
    """Seeds, walk, merge (and optional smoothing / uncertainty) on device-resident inputs.

    walk_opts: keyword options of Engine.walk (algorithm, segment_steps,
    tile_shape, graph_rounds); they change the speed of the walk, not its hits.

    path: 'small' follows rw_small / pyopencl_small (every seed walks); 'large'
    the large-volume driver on one GPU (rw_large.py:40-123, pyopencl_large.py:
    243-496): its plane readers (a labelled XY plane hides the next 4; allaxis
    planes count positive labels only), adaptive walks (a seed walks if it lies
    on the sub-lattice x+z, y+z even, or sees another label on the stride-5 grid
    within +-100 voxels) and the clamp of the first label's smoothed map at 0.
    Volumes beyond the reference's OpenCL range (1.8e9 voxels), which it splits
    into z-subdomains that confine the walks, are walked whole here.

    Returns a dict of device buffers: 'regular' (uint8 V), optional 'smooth',
    'uncertainty', 'hits' (uint32 [nlabels][V]), plus host ints 'nseeds',
    'accepted', 'flags' (uint8 Z+Y+X).
    """
    eng = vol.eng
    Z, Y, X = vol.dims
    V = vol.V
    if nlabels > 256:
        raise PipelineError("at most 256 labels (including background) are supported")
    # a volume of all 256 label values uses every index as a label (no ignored index)
    lnone = _hip.LABEL_ALL if nlabels == 256 else _hip.LABEL_NONE
    t = timings if timings is not None else {}
    t0 = time.perf_counter()
    if path not in ("small", "large"):
        raise ValueError("path must be 'small' or 'large'")
    large = path == "large"
    flags_buf = eng.plane_flags(vol.lab, vol.dims, allaxis, label_none=lnone)
    flags = flags_buf.download((Z + Y + X,), np.uint8)
    if allaxis and adjacent_planes((flags[:Z], flags[Z:Z + Y], flags[Z + Y:])):
        raise PipelineError("At least one empty slice between labels required.")
    if large:
        eng.plane_flags(vol.lab, vol.dims, allaxis, out=flags_buf, large=True, label_none=lnone)
        flags = flags_buf.download((Z + Y + X,), np.uint8)
    seeds, nseeds = eng.seeds(vol.lab, vol.img, vol.img_dtype, vol.dims, flags_buf, label_none=lnone)
    nwalk = nseeds
    if large and nseeds:
        mask = eng.adaptive_mask(vol.lab, vol.dims, flags_buf, label_none=lnone)
        walkers = eng.alloc(max(1, nseeds) * _hip.SEED_DTYPE.itemsize)
        _, nwalk = eng.select_seeds(seeds, nseeds, mask, out=walkers)
        mask.free()
        seeds.free()
        seeds = walkers
    t1 = time.perf_counter()
    hits = eng.alloc(nlabels * V * 4)
    hits.zero()
    accepted = eng.walk(vol.img, vol.img_dtype, vol.dims, seeds, nwalk, nlabels, nbrw, sorw,
                        _RNG[rng], key, hits, **(walk_opts or {}))
    t2 = time.perf_counter()
    out = {"nseeds": nseeds, "walked_seeds": nwalk, "accepted": accepted, "flags": flags, "path": path}
    regular = eng.alloc(V)
    eng.argmax(hits, nlabels, V, values, regular)
    out["regular"] = regular
    if smooth:
        def smooth_stage():
            sm = eng.alloc(V)
            eng.smooth_argmax(hits, nlabels, vol.dims, smooth, values, sm, clamp_first=large)
            return sm
        sm = optional_stage(smooth_stage, SMOOTH_OOM_WARNING)
        if sm is not None:
            out["smooth"] = sm
    if uncertainty:
        def uncertainty_stage():
            uq = eng.alloc(V)
            eng.uncertainty(hits, nlabels, vol.dims, uq)
            return uq
        uq = optional_stage(uncertainty_stage, UNCERTAINTY_OOM_WARNING)
        if uq is not None:
            out["uncertainty"] = uq
    eng.sync()
    t3 = time.perf_counter()
    if keep_hits:
        out["hits"] = hits
    else:
        hits.free()
    seeds.free()
    phase_ms, kernel_ms, launches = eng.last_walk_stats()
    t.update(seeds_s=t1 - t0, walk_s=t2 - t1, merge_s=t3 - t2, walk_phase_ms=phase_ms,
             walk_kernel_ms=kernel_ms, walk_launches=launches)
    return out


def run_pipeline_guided(vol, nlabels, values, net, nbrw=10, sorw=4000, allaxis=False, smooth=0,
                        key=DEFAULT_KEY, thr=0.9, brad=4, srad=6, far=48, far_walks=1, repair=2, repair_min=10,
                        repair_far_walks=None, repair_far=24, repair_srad=3, mid=16, mid_walks=None,
                        plain_above=0.5, walk_opts=None,
                        repair_walk_opts=None, timings=None, keep=False):
    """Surrogate-guided pipeline on device-resident inputs.

    The network proposes a label and a confidence for every voxel.  U = voxels
    whose top (temperature-scaled) probability is below `thr` plus the proposed
    label boundaries dilated by `brad`.  Seeds within `srad` (Chebyshev) of U
    walk all `nbrw` walks, with the same Philox streams as the plain walker, so
    those walks are exactly the plain walker's.  Seeds in the ring from `srad`
    to `far` walk their first `far_walks` walks with every hit counted
    nbrw/far_walks times: an unbiased estimate of what they add to U
    (far=0 disables the ring); ring seeds within `mid` walk `mid_walks` walks
    instead (mid=None: none; mid_walks=None: 2 when nbrw is even, else
    far_walks), which lowers the noise of the estimate where it weighs most,
    next to U.  Near and ring seeds walk in one call (per-seed
    walk counts and weights), so ring walkers fill the same tile buckets.
    Repair (repair = a radius; None disables it): outside U, a voxel whose hits
    from this walk number at least `repair_min` and whose walker argmax is not the
    proposed label marks a confident network error.  Those voxels dilated by
    `repair` (minus U) form U2; every seed within `repair_srad` (None: srad)
    of U2 walks all its walks again into a second hit volume, whose counts replace the first walk's
    on U2, and U2 joins U.  The repair walk has its own ring out to
    `repair_far` (None: far), whose seeds walk `repair_far_walks` walks (None:
    far_walks); repair_walk_opts overrides walk_opts for it.  GUIDED_ROBUST
    holds slower, more conservative settings.
    plain_above: when more than this fraction of the seeds falls in the near
    set (the proposal is uncertain almost everywhere, as on volumes unlike its
    training data), the plain pipeline runs instead: its result is exact and
    costs little more (None: never).  So does a volume of more labels than the
    network predicts (surrogate.L_MAX).
    walk_opts: as for run_pipeline.  The result takes the walker's argmax on U and
    the network's label elsewhere ('smooth': the smoothed walker argmax on U).
    Returns the dict of run_pipeline plus 'walked_seeds', 'ring_seeds' (of
    which 'mid_ring_seeds' walk mid_walks walks), 'repair_seeds',
    'repair_ring_seeds', 'repair_voxels' and 'fallback'.
    """
    from . import surrogate as S
    eng = vol.eng
    Z, Y, X = vol.dims
    V = vol.V
    if far > srad and (far_walks < 1 or nbrw % far_walks):
        raise PipelineError("far_walks must divide nbrw")
    if mid_walks is None:
        mid_walks = 2 if nbrw % 2 == 0 else far_walks
    # the inner ring reaches min(mid, far); it is empty when mid <= srad
    mid = None if far <= srad or mid is None or mid <= srad else min(mid, far)
    if mid is not None and (mid_walks < 1 or nbrw % mid_walks):
        raise PipelineError("mid_walks must divide nbrw")
    t = timings if timings is not None else {}
    if nlabels > S.L_MAX:
        # the network's head predicts at most L_MAX labels: the plain walk
        # labels the volume instead (its result is exact)
        out = run_pipeline(vol, nlabels, values, nbrw, sorw, allaxis, smooth, key=key, keep_hits=keep,
                           walk_opts=walk_opts, timings=t)
        out.update(ring_seeds=0, mid_ring_seeds=0, repair_seeds=0, repair_ring_seeds=0, repair_voxels=0,
                   fallback=True)
        return out
    t0 = time.perf_counter()
    flags_buf = eng.plane_flags(vol.lab, vol.dims, allaxis)
    flags = flags_buf.download((Z + Y + X,), np.uint8)
    if allaxis and adjacent_planes((flags[:Z], flags[Z:Z + Y], flags[Z + Y:])):
        raise PipelineError("At least one empty slice between labels required.")
    seeds, nseeds = eng.seeds(vol.lab, vol.img, vol.img_dtype, vol.dims, flags_buf)
    t1 = time.perf_counter()
    pred, unc = eng.alloc(V), eng.alloc(V)
    net.predict_volume(vol, flags, nlabels, thr, pred, unc)
    U, need, tmp = eng.alloc(V), eng.alloc(V), eng.alloc(V)
    _hip.call("bmx_guided_masks", eng.ctx, pred.ptr, unc.ptr, Z, Y, X, int(brad), int(srad), U.ptr, need.ptr,
              tmp.ptr)
    ring = eng.alloc(V) if far > srad else None
    inner_ring = eng.alloc(V) if mid is not None else None

    def walkers(near, far_walks, far=far, inner=srad, mid=None):
        """Seeds in `near` with all walks, then the ring seeds out to `far` with
        far_walks weighted walks; with `mid`, the ring seeds out to mid walk
        mid_walks weighted walks instead -> (list, near count, ring count, of
        which out to mid)."""
        sub = eng.alloc(max(1, nseeds) * _hip.SEED_DTYPE.itemsize)
        _, nsub = eng.select_seeds(seeds, nseeds, near, out=sub)
        nring = nmid = 0
        if ring is not None and far > inner:
            # ring seeds follow the near seeds in the same list (the sets are disjoint)
            _hip.call("bmx_ring_mask", eng.ctx, near.ptr, Z, Y, X, int(far - inner), ring.ptr, tmp.ptr)
            if mid is not None:
                _hip.call("bmx_ring_mask", eng.ctx, near.ptr, Z, Y, X, int(mid - inner), inner_ring.ptr, tmp.ptr)
                _hip.call("bmx_mask_andnot", eng.ctx, ring.ptr, inner_ring.ptr, V)
                tail = _hip.BufferView(sub, nsub * _hip.SEED_DTYPE.itemsize)
                _, nmid = eng.select_seeds(seeds, nseeds, inner_ring, out=tail)
                _hip.call("bmx_seed_walks", eng.ctx, tail.ptr, nmid, int(mid_walks), nbrw // mid_walks)
            tail = _hip.BufferView(sub, (nsub + nmid) * _hip.SEED_DTYPE.itemsize)
            _, nout = eng.select_seeds(seeds, nseeds, ring, out=tail)
            _hip.call("bmx_seed_walks", eng.ctx, tail.ptr, nout, int(far_walks), nbrw // far_walks)
            nring = nmid + nout
        return sub, nsub, nring, nmid

    sub, nsub, nring, nmid = walkers(need, far_walks, mid=mid)
    eng.sync()
    if plain_above is not None and nseeds and nsub > plain_above * nseeds:
        out = run_pipeline(vol, nlabels, values, nbrw, sorw, allaxis, smooth, key=key, keep_hits=keep,
                           walk_opts=walk_opts, timings=t)
        out.update(walked_seeds=nseeds, ring_seeds=0, mid_ring_seeds=0, repair_seeds=0, repair_ring_seeds=0, repair_voxels=0,
                   fallback=True)
        if keep:
            every = eng.upload(np.ones(V, np.uint8))     # every voxel carries the walker's label
            out.update(pred=pred, U=every, need=every)
        t.update(proposal_s=time.perf_counter() - t1)
        return out
    t2 = time.perf_counter()
    hits = eng.alloc(nlabels * V * 4)
    hits.zero()
    wopts = dict(GUIDED_WALK_OPTS, **(walk_opts or {}))
    accepted = eng.walk(vol.img, vol.img_dtype, vol.dims, sub, nsub + nring, nlabels, nbrw, sorw,
                        _hip.RNG_PHILOX, key, hits, seed_overrides=nring > 0, **wopts)
    phase_ms, kernel_ms, launches = eng.last_walk_stats()
    t3 = time.perf_counter()
    nrep, nrep_ring, rep_vox = 0, 0, 0
    if repair is not None:
        U2 = eng.alloc(V)
        _hip.call("bmx_guided_disagree", eng.ctx, hits.ptr, int(nlabels), V, pred.ptr, U.ptr, int(repair_min),
                  U2.ptr)
        need2 = eng.alloc(V)
        rsrad = srad if repair_srad is None else repair_srad
        _hip.call("bmx_repair_masks", eng.ctx, U2.ptr, U.ptr, Z, Y, X, int(repair), int(rsrad), need2.ptr, tmp.ptr)
        rfw = far_walks if repair_far_walks is None else repair_far_walks
        if far > srad and (rfw < 1 or nbrw % rfw):
            raise PipelineError("repair_far_walks must divide nbrw")
        sub2, nrep, nrep_ring, _ = walkers(need2, rfw, far if repair_far is None else repair_far, rsrad)
        need2.free()
        t["repair_masks_s"] = time.perf_counter() - t3
        if nrep:
            rep_vox = int(eng.histogram(U2, V)[1])
            hits2 = eng.alloc(nlabels * V * 4)
            hits2.zero()
            ropts = wopts if repair_walk_opts is None else dict(GUIDED_WALK_OPTS, **repair_walk_opts)
            accepted += eng.walk(vol.img, vol.img_dtype, vol.dims, sub2, nrep + nrep_ring, nlabels, nbrw, sorw,
                                 _hip.RNG_PHILOX, key, hits2, seed_overrides=nrep_ring > 0, **ropts)
            p2, k2, l2 = eng.last_walk_stats()
            t["repair_walk_ms"] = p2
            phase_ms += p2
            kernel_ms = kernel_ms + k2 if kernel_ms >= 0 and k2 >= 0 else -2.0
            launches += l2
            _hip.call("bmx_repair_hits", eng.ctx, hits.ptr, hits2.ptr, int(nlabels), V, U2.ptr, U.ptr)
            eng.sync()
            hits2.free()
        sub2.free()
        U2.free()
    t3r = time.perf_counter()
    out = {"nseeds": nseeds, "walked_seeds": nsub, "ring_seeds": nring, "mid_ring_seeds": nmid, "repair_seeds": nrep,
           "repair_ring_seeds": nrep_ring, "repair_voxels": rep_vox, "accepted": accepted, "flags": flags,
           "fallback": False}
    values = np.ascontiguousarray(values, dtype=np.uint8)
    regular = eng.alloc(V)
    _hip.call("bmx_merge_guided", eng.ctx, hits.ptr, int(nlabels), V, values.ctypes.data, pred.ptr, U.ptr,
              regular.ptr)
    out["regular"] = regular
    if smooth:
        def smooth_stage():
            sm_all, sm = eng.alloc(V), eng.alloc(V)
            eng.smooth_argmax(hits, nlabels, vol.dims, smooth, values, sm_all)
            _hip.call("bmx_select_labels", eng.ctx, sm_all.ptr, pred.ptr, U.ptr, int(nlabels), values.ctypes.data,
                      V, sm.ptr)
            return sm
        sm = optional_stage(smooth_stage, SMOOTH_OOM_WARNING)
        if sm is not None:
            out["smooth"] = sm
    eng.sync()
    t4 = time.perf_counter()
    if keep:
        out.update(pred=pred, U=U, need=need, hits=hits)
    t.update(seeds_s=t1 - t0, proposal_s=t2 - t1, walk_s=t3 - t2, repair_s=t3r - t3, merge_s=t4 - t3r,
             walk_phase_ms=phase_ms, walk_kernel_ms=kernel_ms, walk_launches=launches)
    return out


def smart_interpolation(data, labelData, nbrw=10, sorw=4000, acwe=False, acwe_alpha=1.0,
                        acwe_smooth=1, acwe_steps=3, path_to_data=None, path_to_labels=None,
                        denoise=False, uncertainty=False, platform=None, allaxis=False,
                        ignore='none', only='all', smooth=0, compression=True, return_hits=False,
                        img_id=None, label_id=None, remote=False, queue=0, clean=None, fill=None,
                        slicer=False, rng="philox", key=DEFAULT_KEY, device=0, method="plain",
                        guide=None, path="auto", timings=None):
    """Smart interpolation of sparsely labelled slices (biomedisa/interpolation.py:43).

    Returns the reference's result dict ({'regular': uint8 (Z,Y,X)} plus
    'smooth', 'uncertainty', 'hits' when requested), or None when the input is
    rejected (after printing "Error: ..." as the reference does).

    method='guided' runs the surrogate-guided pipeline (run_pipeline_guided;
    Philox stream only, no 'uncertainty'/'hits'); `guide` is a dict of its
    keyword parameters (thr, brad, srad, far, far_walks, mid, repair, ...),
    which override guided_defaults(nbrw, allaxis) (stricter settings for
    labels on XY planes only).

    path: 'auto' (default) takes the reference's small- or large-volume driver
    by its size rule (interpolation.py:192-196), 'small' or 'large' forces one;
    see run_pipeline.  The guided method always walks as the small driver.

    timings: a dict that receives the phase times (prepare_s, pipeline_s,
    results_s), the pipeline's own timings and counts (seeds, accepted moves).
    """
    files = None
    if data is None or labelData is None:
        # pre_processing (features/biomedisa_helper.py:639-655): read what was not
        # passed; result files are written only when the image came from a file
        # (interpolation.py:76-79)
        from . import io as bio
        final_image_type, label_header = ".tif", None
        image_from_file = data is None
        if data is None:
            data = bio.load_data(path_to_data)[0] if path_to_data is not None else None
            if data is None:
                print("Error: Invalid image data.")
                return None
        if labelData is None:
            if path_to_labels is not None:
                labelData, final_image_type = bio.load_data(path_to_labels)
                if labelData is not None:
                    label_header = bio.read_header(path_to_labels)
            if labelData is None:
                print("Error: Invalid label data.")
                return None
        if image_from_file:
            files = bio.result_paths(path_to_data, final_image_type)
    if rng not in _RNG:
        raise ValueError("rng must be 'philox' or 'reference'")
    if method not in ("plain", "guided"):
        raise ValueError("method must be 'plain' or 'guided'")
    if path not in ("auto", "small", "large"):
        raise ValueError("path must be 'auto', 'small' or 'large'")
    if method == "guided" and (rng != "philox" or uncertainty or return_hits):
        raise ValueError("guided mode runs the Philox walker and returns 'regular'/'smooth' only")

    eng = _hip.engine(device)
    got = _interpolate(eng, data, labelData, nbrw, sorw, uncertainty, platform, allaxis, ignore, only, smooth,
                       return_hits, slicer, denoise, rng, key, method, guide, path, device, timings)
    if got is None:
        return None
    prep, results = got
    bm = prep.bm
    post_process(results, clean, fill, device)
    if acwe:
        # rw_small.py:244-253: the (normalised) image of the walk's box, zeros
        # elsewhere, and the regular result
        from .post import active_contour, refinement
        z0, z1, y0, y1, x0, x1 = prep.box
        img_c = prep.host_image()
        data = np.zeros((bm.zsh, bm.ysh, bm.xsh), dtype=img_c.dtype)
        data[z0:z1, y0:y1, x0:x1] = img_c
        data = data[1:-1, 1:-1, 1:-1]
        results["acwe"] = active_contour(data, results["regular"], acwe_alpha, acwe_smooth, acwe_steps, device)
        results["refined"] = refinement(data, results["regular"], device)
    if files is not None:
        # the result files of rw_small.py:212-253, named as interpolation.py:115-133
        from . import io as bio
        for key, path in files.items():
            if results.get(key) is not None:
                if key == "uncertainty":
                    bio.save_data(path, results[key], compression)
                else:
                    bio.save_data(path, results[key], compression, header=label_header,
                                  file_type=final_image_type)
    return results


class DeviceArray:
    """A caller's array resident in device memory (shape and dtype as the
    NumPy array it stands for, C order), as smart_interpolation_device takes
    its inputs and returns its results."""

    def __init__(self, buf, shape, dtype):
        self.buf, self.shape, self.dtype = buf, tuple(int(s) for s in shape), np.dtype(dtype)

    @classmethod
    def upload(cls, eng, arr):
        arr = np.ascontiguousarray(arr)
        return cls(eng.upload(arr), arr.shape, arr.dtype)

    def download(self):
        return self.buf.download(self.shape, self.dtype)

    def free(self):
        self.buf.free()


def smart_interpolation_device(data, labelData, nbrw=10, sorw=4000, allaxis=False, ignore='none', only='all',
                               smooth=0, uncertainty=False, denoise=False, rng="philox", key=DEFAULT_KEY, device=0,
                               method="plain", guide=None, path="auto", timings=None):
    """smart_interpolation on arrays already in device memory (DeviceArray image
    and labels in the dtypes the drop-in takes: 8/16/32-bit integers or
    float32), with the results left there: {'regular': DeviceArray (Z,Y,X)
    uint8, optional 'smooth', 'uncertainty'}, or None when the input is
    rejected.  Every step after the inputs are resident runs on the device:
    label set and box, normalisation, pad and crop, the walk and its merges,
    result placement and the sanity check."""
    if method == "guided" and (rng != "philox" or uncertainty):
        raise ValueError("guided mode runs the Philox walker and returns 'regular'/'smooth' only")
    eng = _hip.engine(device)
    got = _interpolate(eng, data, labelData, nbrw, sorw, uncertainty, None, allaxis, ignore, only, smooth, False,
                       False, denoise, rng, key, method, guide, path, device, timings, to_host=False)
    if got is NotImplemented:
        raise TypeError("smart_interpolation_device takes 8/16/32-bit integer or float32 arrays below 2^32 voxels")
    return None if got is None else got[1]


def _interpolate(eng, data, labelData, nbrw, sorw, uncertainty, platform, allaxis, ignore, only, smooth,
                 return_hits, slicer, denoise, rng, key, method, guide, path, device, timings, to_host=True):
    """Pre-processing, pipeline and results of one call -> (prep, results), or
    None when the input is rejected (after printing the error)."""
    t = timings if timings is not None else {}
    t0 = time.perf_counter()
    prep = _prepare_device(eng, data, labelData, nbrw, sorw, uncertainty, platform, allaxis, ignore, only,
                           smooth, return_hits, slicer, denoise)
    if prep is NotImplemented:
        if not to_host:
            return NotImplemented
        # label or image types the device pre-processing does not take: on the host
        prep = _prepare(data, labelData, nbrw, sorw, uncertainty, platform, allaxis, ignore, only, smooth,
                        return_hits, slicer, denoise, device)
        if prep is not None:
            prep.vol = DeviceVolume(eng, prep.img, prep.labidx)
    if prep is None:
        return None
    bm = prep.bm
    vol = prep.vol
    t1 = time.perf_counter()
    try:
        if method == "guided":
            from . import surrogate
            dev = run_pipeline_guided(vol, bm.nol, prep.values, surrogate.net(device), bm.nbrw, bm.sorw,
                                      bm.allaxis, bm.smooth, key, timings=t,
                                      **dict(guided_defaults(bm.nbrw, bm.allaxis), **(guide or {})))
        else:
            if path == "auto":
                path = select_path(vol.dims, bm.nol)
            dev = run_pipeline(vol, bm.nol, prep.values, bm.nbrw, bm.sorw, bm.allaxis, bm.smooth,
                               bm.uncertainty, rng, key, keep_hits=bm.return_hits, path=path, timings=t)
    except PipelineError as e:
        _error_(bm, str(e))
        return None
    t2 = time.perf_counter()
    results = _results(prep, dev) if to_host else _results_device(prep, dev, to_host=False)
    t3 = time.perf_counter()
    t.update(prepare_s=t1 - t0, pipeline_s=t2 - t1, results_s=t3 - t2,
             stats={k: v for k, v in dev.items() if isinstance(v, (int, bool, str))})
    return prep, results


def post_process(results, clean=None, fill=None, device=0):
    """The reference's label post-processing (rw_small.py:222-241) on the device:
    'cleaned' (and 'smooth_cleaned') by bmx_clean, 'filled' by bmx_fill, and
    'cleaned_filled' = cleaned + (filled - regular) in uint8 arithmetic, as the
    reference computes it."""
    if not (clean or fill):
        return results
    eng = _hip.engine(device)

    def run(name, vol, threshold):
        buf = eng.upload(np.ascontiguousarray(vol, dtype=np.uint8))
        out = eng.alloc(vol.size)
        getattr(eng, name)(buf, vol.shape, threshold, out)
        res = out.download(vol.shape, np.uint8)
        buf.free()
        out.free()
        return res

    final = results["regular"]
    if clean:
        results["cleaned"] = run("clean", final, clean)
        if "smooth" in results:
            results["smooth_cleaned"] = run("clean", results["smooth"], clean)
    if fill:
        results["filled"] = run("fill", final, fill)
    if clean and fill:
        results["cleaned_filled"] = results["cleaned"] + (results["filled"] - final)
    return results


class _Prepared:
    """A validated input on the host: the padded, cropped image and label-index
    volume the pipeline runs on, and what is needed to place its results."""

    def __init__(self, bm, img, lab, labidx, values, box):
        self.bm, self.img, self.lab_c, self.labidx, self.values, self.box = bm, img, lab, labidx, values, box
        self.vol = None

    def host_image(self):
        return self.img


class _DevicePrepared:
    """A validated input prepared on the device: the walk's volume (cropped,
    padded image and label-index volume), the label values and the box."""

    def __init__(self, bm, vol, values, box, shape, label_dtype):
        self.bm, self.vol, self.values, self.box = bm, vol, values, box
        self.shape, self.label_dtype = shape, label_dtype

    def host_image(self):
        return self.vol.img.download(self.vol.dims, np.uint8 if self.vol.img_dtype == _hip.IMG_U8 else np.float32)


def _prepare_device(eng, data, labelData, nbrw, sorw, uncertainty, platform, allaxis, ignore, only, smooth,
                    return_hits, slicer, denoise=False):
    """Checks and pre-processing of smart_interpolation (interpolation.py:77-204)
    with the arrays uploaded once and every step after the checks on the device:
    label set and label box (bmx_label_stats), normalisation, denoise, pad and
    crop (bmx_crop_image, bmx_crop_labels).  Returns a _DevicePrepared, None
    when the input is rejected (after printing the error), or NotImplemented
    for inputs the device steps do not take (8-byte or boolean labels, images
    of types other than 8/16/32-bit integers and float32 are normalised on the
    host first; volumes of 2^32 voxels and more), which the caller prepares on
    the host instead."""
    from .helpers import box_from_class_boxes, class_lut, pre_processing_counts
    bm = Biomedisa()
    bm.TIC = time.time()
    bm.success = True
    bm.data, bm.labelData = data, labelData
    bm.nbrw, bm.sorw, bm.allaxis, bm.smooth = int(nbrw), int(sorw), bool(allaxis), int(smooth)
    bm.uncertainty, bm.return_hits = bool(uncertainty), bool(return_hits)
    bm.ignore, bm.only, bm.platform, bm.slicer = ignore, only, platform, slicer
    # the checks of pre_processing (biomedisa_helper.py:639-661) that need no pass over the data
    if data is None:
        _error_(bm, "Invalid image data.")
        return None
    if labelData is None:
        _error_(bm, "Invalid label data.")
        return None
    on_device = isinstance(data, DeviceArray), isinstance(labelData, DeviceArray)
    if not on_device[0]:
        data = np.asarray(data)
    if not on_device[1]:
        labelData = np.asarray(labelData)
    if len(labelData.shape) != 3:
        _error_(bm, "Label data must be three-dimensional.")
        return None
    if data.shape != labelData.shape:
        _error_(bm, "Image and label data must have the same x,y,z-dimensions.")
        return None
    Z, Y, X = (int(s) for s in labelData.shape)
    if Z * Y * X >= 2 ** 32 or min(Z, Y, X) < 1:
        return NotImplemented
    if labelData.dtype not in _hip.RAW_DTYPES or labelData.dtype == np.float32:
        if labelData.dtype.kind != "f" or on_device[1]:
            return NotImplemented
        print(f"Warning: Potential label loss during conversion from {labelData.dtype} to int32.")
        labelData = labelData.astype(np.int32)
    if on_device[0] and data.dtype not in _hip.RAW_DTYPES:
        return NotImplemented
    lab_raw = _hip.RAW_DTYPES[labelData.dtype]
    lab_dev = labelData.buf if on_device[1] else eng.upload(labelData)
    try:
        counts, boxes = eng.label_stats(lab_dev, lab_raw, (Z, Y, X))
        if pre_processing_counts(bm, counts) is None:
            return NotImplemented
        if not bm.success:
            return None
        if labelData.dtype == np.uint32:
            # interpolation.py:145-147 (the values above 255 are already 0)
            print(f"Warning: Potential label loss during conversion from {labelData.dtype} to int32.")
        if bm.nol > 256 or (bm.nol == 256 and bm.neg_kept):
            _error_(bm, NEG_WITH_256_LABELS)
            return None
        box = box_from_class_boxes(boxes, bm.kept, bm.neg_kept, (Z, Y, X))
        dims_c = (box[1] - box[0], box[3] - box[2], box[5] - box[4])
        Vc = dims_c[0] * dims_c[1] * dims_c[2]
        if Vc >= 2 ** 32:
            return NotImplemented
        bm.zsh, bm.ysh, bm.xsh = Z + 2, Y + 2, X + 2
        labidx = eng.alloc(Vc)
        eng.crop_labels(lab_dev, lab_raw, (Z, Y, X), box, class_lut(bm.allLabels, bm.neg_kept), labidx)
    finally:
        if not on_device[1]:
            lab_dev.free()

    # the image: data types (interpolation.py:135-144), denoise (:149-151), pad and crop
    owned = True
    if on_device[0]:
        raw, src, owned = _hip.RAW_DTYPES[data.dtype], data.buf, False
    elif data.dtype in _hip.RAW_DTYPES:
        raw = _hip.RAW_DTYPES[data.dtype]
        src = eng.upload(data)
    else:
        raw, src = _hip.RAW_F32, eng.upload(normalize_image(data))
    normalize, vmin, vrange = raw not in (_hip.RAW_U8, _hip.RAW_I8), 0.0, 1.0
    if normalize and data.dtype in _hip.RAW_DTYPES:
        lo, hi = eng.image_range(src, raw, Z * Y * X)
        vmin = np.float32(lo)
        vrange = np.float32(np.float32(hi) - vmin)
    elif raw == _hip.RAW_F32:
        normalize = False   # normalised on the host above
    out_u8 = raw in (_hip.RAW_U8, _hip.RAW_I8)
    img_dtype = _hip.IMG_U8 if out_u8 else _hip.IMG_F32
    esize = 1 if out_u8 else 4
    if denoise:
        # the whole normalised image first, then smooth_img_3x3 on it
        full = eng.alloc(Z * Y * X * esize)
        eng.crop_image(src, raw, normalize, (Z, Y, X), (1, Z + 1, 1, Y + 1, 1, X + 1), vmin, vrange, full)
        if owned:
            src.free()
        src, owned = eng.alloc(Z * Y * X * esize), True
        eng.denoise(full, img_dtype, (Z, Y, X), src)
        full.free()
        raw, normalize = (_hip.RAW_U8 if out_u8 else _hip.RAW_F32), False
    img_c = eng.alloc(Vc * esize)
    eng.crop_image(src, raw, normalize, (Z, Y, X), box, vmin, vrange, img_c)
    eng.sync()
    if owned:
        src.free()
    vol = DeviceVolume.from_buffers(eng, img_c, labidx, dims_c, img_dtype)
    values = np.asarray(bm.allLabels, dtype=np.uint8)
    return _DevicePrepared(bm, vol, values, box, (Z, Y, X), labelData.dtype)


def denoise_image(img, device=0):
    """smooth_img_3x3 (features/biomedisa_helper.py:309-323) on the device
    (bmx_denoise): the mean of each voxel's 3x3x3 neighbourhood clipped to the
    volume, in the image's type (uint8 or float32, as normalize_image leaves it)."""
    img = np.ascontiguousarray(img)
    if img.dtype not in (np.uint8, np.float32):
        raise TypeError("denoise_image takes uint8 or float32 images")
    eng = _hip.engine(device)
    src = eng.upload(img)
    out = eng.alloc(img.nbytes)
    eng.denoise(src, _hip.IMG_U8 if img.dtype == np.uint8 else _hip.IMG_F32, img.shape, out)
    res = out.download(img.shape, img.dtype)
    src.free()
    out.free()
    return res


def _prepare(data, labelData, nbrw, sorw, uncertainty, platform, allaxis, ignore, only, smooth,
             return_hits, slicer, denoise=False, device=0):
    """Checks and pre-processing of smart_interpolation (interpolation.py:77-204);
    None (after printing the error) when the input is rejected."""
    bm = Biomedisa()
    bm.TIC = time.time()
    bm.success = True
    bm.data, bm.labelData = data, labelData
    bm.nbrw, bm.sorw, bm.allaxis, bm.smooth = int(nbrw), int(sorw), bool(allaxis), int(smooth)
    bm.uncertainty, bm.return_hits = bool(uncertainty), bool(return_hits)
    bm.ignore, bm.only, bm.platform, bm.slicer = ignore, only, platform, slicer

    bm = pre_processing(bm)
    if not bm.success:
        return None

    # data types (interpolation.py:135-147)
    bm.data = normalize_image(bm.data)
    if bm.labelData.dtype in ["uint32", "int64", "uint64"]:
        print(f"Warning: Potential label loss during conversion from {bm.labelData.dtype} to int32.")
        bm.labelData = bm.labelData.astype(np.int32)
    if denoise:
        # interpolation.py:149-151
        bm.data = denoise_image(bm.data, device)

    # 1-voxel pad (:157-159) and label bounding box +-100 (:192, :203-204)
    img = np.pad(bm.data, 1, mode="constant", constant_values=0)
    lab = np.pad(bm.labelData, 1, mode="constant", constant_values=0)
    bm.zsh, bm.ysh, bm.xsh = img.shape
    box = predict_blocksize(lab)
    z0, z1, y0, y1, x0, x1 = box
    img_c = np.ascontiguousarray(img[z0:z1, y0:y1, x0:x1])
    lab_c = np.ascontiguousarray(lab[z0:z1, y0:y1, x0:x1])
    labidx = label_index_volume(lab_c, bm.allLabels)
    values = np.asarray(bm.allLabels, dtype=np.uint8)
    if bm.nol > 256 or (bm.nol == 256 and np.any(lab_c < 0)):
        _error_(bm, NEG_WITH_256_LABELS)
        return None
    return _Prepared(bm, img_c, lab_c, labidx, values, box)


def _results(prep, dev):
    """The result dictionary of smart_interpolation from the pipeline's device
    buffers (rw_small.py:195-281)."""
    if isinstance(prep, _DevicePrepared):
        return _results_device(prep, dev)
    bm = prep.bm
    z0, z1, y0, y1, x0, x1 = prep.box
    shape_c = prep.img.shape
    results = {}

    def place(buf):
        crop = buf.download(shape_c, np.uint8)
        full = np.zeros((bm.zsh, bm.ysh, bm.xsh), dtype=np.uint8)
        full[z0:z1, y0:y1, x0:x1] = crop
        return full[1:-1, 1:-1, 1:-1], crop

    final_result, final_crop = place(dev["regular"])
    if "smooth" in dev:
        results["smooth"] = place(dev["smooth"])[0]
    if "uncertainty" in dev:
        results["uncertainty"] = place(dev["uncertainty"])[0]
    if bm.return_hits:
        h = dev["hits"].download((bm.nol,) + shape_c, np.uint32)
        hits = np.zeros((bm.nol, bm.zsh, bm.ysh, bm.xsh), dtype=np.float32)
        hits[:, z0:z1, y0:y1, x0:x1] = h
        results["hits"] = hits[:, 1:-1, 1:-1, 1:-1]

    # validate result (rw_small.py:203-211)
    mask = prep.lab_c > 0
    bad_result_warning(Dice_score(prep.lab_c, final_crop * mask))

    results["regular"] = final_result
    return results


def bad_result_warning(dice):
    """The sanity check of rw_small.py:203-211: a global Dice below 0.3 between
    the labelled voxels and the result prints the reference's warning."""
    if dice < 0.3:
        print('Warning: Bad result! Use "--allaxis" if you labeled axes other than the xy-plane.')
        return True
    return False


def _results_device(prep, dev, to_host=True):
    """The result dictionary from the pipeline's device buffers, placed into
    whole (Z,Y,X) volumes on the device (bmx_place) and copied out once each
    (to_host=False: left on the device as DeviceArrays); the sanity check's
    Dice from device counts (bmx_dice_counts)."""
    from .helpers import dice_counted, dice_from_counts
    bm = prep.bm
    eng = prep.vol.eng
    z0, z1, y0, y1, x0, x1 = prep.box
    shape_c = prep.vol.dims
    shape = prep.shape
    results = {}
    full = eng.alloc(int(np.prod(shape))) if to_host else None

    def place(buf):
        if to_host:
            eng.place(buf, shape, prep.box, full)
            return full.download(shape, np.uint8)
        out = eng.alloc(int(np.prod(shape)))
        eng.place(buf, shape, prep.box, out)
        return DeviceArray(out, shape, np.uint8)

    for key in ("regular", "smooth", "uncertainty"):
        if key in dev:
            results[key] = place(dev[key])
    if full is not None:
        full.free()
    if bm.return_hits:
        h = dev["hits"].download((bm.nol,) + shape_c, np.uint32)
        hits = np.zeros((bm.nol, bm.zsh, bm.ysh, bm.xsh), dtype=np.float32)
        hits[:, z0:z1, y0:y1, x0:x1] = h
        results["hits"] = hits[:, 1:-1, 1:-1, 1:-1]
    a, b, c = eng.dice_counts(prep.vol.lab, dev["regular"], prep.vol.V, bm.nol, prep.values,
                              dice_counted(prep.label_dtype))
    bad_result_warning(dice_from_counts(a, b, c))
    if not to_host:
        eng.sync()
    for key in ("regular", "smooth", "uncertainty", "hits"):
        if key in dev:
            dev[key].free()
    return results
