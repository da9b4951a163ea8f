"""Host side of the label post-processing run after the walk.

Mirrors biomedisa/features/active_contour.py (activeContour, refinement) and
biomedisa/features/remove_outlier.py (clean, fill), whose device work runs in
libbmx_hip (bmx_acwe, bmx_clean, bmx_fill).
"""
import numpy as np

from . import _hip


def img_to_uint8(img):
    """biomedisa_helper.py:345-352."""
    if img.dtype != np.uint8:
        img = img.astype(np.float32)
        img -= np.amin(img)
        img /= np.amax(img)
        img *= 255.0
        img = img.astype(np.uint8)
    return img


def _box(mask):
    """reduce_blocksize of active_contour.py:91-110: bbox of the non-zero labels
    (mask) grown by 100 and clamped."""
    zsh, ysh, xsh = mask.shape
    nz = [np.nonzero(np.any(mask, axis=tuple(a for a in range(3) if a != d)))[0] for d in range(3)]
    lo = lambda v: v - 100 if v - 100 > 0 else 0
    hi = lambda v, n: v + 100 if v + 100 < n else n
    return tuple(slice(lo(int(n.min())), hi(int(n.max()), s)) for n, s in zip(nz, (zsh, ysh, xsh)))


class _Inputs:
    """activeContour's inputs on the device (active_contour.py:112-156): the
    padded, cropped uint8 image and labels, the box, the padded shape and the
    label set."""

    def __init__(self, eng, data, labels):
        labels = np.asarray(labels)
        if labels.dtype.kind == "f":
            labels = labels.astype(np.int32)
        img = np.pad(img_to_uint8(data), 1)
        lab = np.pad(labels, 1)
        self.box = _box(lab != 0)
        self.shape = img.shape
        img_b = np.ascontiguousarray(img[self.box])
        lab_b = np.ascontiguousarray(lab[self.box])
        self.dims = lab_b.shape
        self.img = eng.upload(img_b)
        if lab_b.dtype == np.uint8:
            self.lab = eng.upload(lab_b)
            counts = eng.histogram(self.lab, lab_b.size)
            vals = np.flatnonzero(counts)
        else:
            vals = np.unique(lab_b)
            vals = vals[(vals >= 0) & (vals <= 255)]
            self.lab = eng.upload(lab_b.astype(np.uint8))
        if not np.any(vals == 0):
            vals = np.concatenate([[0], vals])
        self.all_labels = vals.astype(np.uint8)

    def place(self, res):
        final = np.zeros(self.shape, np.uint8)
        final[self.box] = res
        return np.ascontiguousarray(final[1:-1, 1:-1, 1:-1])

    def free(self):
        self.img.free()
        self.lab.free()


def _inputs(eng, data, labels):
    """None where the reference's pre_processing rejects the labels: shapes
    differ, or no non-zero label (the label set then has fewer than two values)."""
    if data.shape != labels.shape or labels.ndim != 3 or not np.any(labels != 0):
        return None
    return _Inputs(eng, data, labels)


def active_contour(data, labels, alpha=1.0, smooth=1, steps=3, device=0):
    """activeContour(data, labels, alpha, smooth, steps) with simple=False: the
    acwe loop on the device (bmx_acwe).  Returns the uint8 labels in the
    input's shape, or None where the reference's pre_processing rejects the
    labels (fewer than two values)."""
    eng = _hip.engine(device)
    inp = _inputs(eng, data, labels)
    if inp is None:
        return None
    n = int(np.prod(inp.dims))
    out, tmp = eng.alloc(n), eng.alloc(n)
    try:
        eng.acwe(inp.img, inp.lab, inp.dims, inp.all_labels, alpha, smooth, steps, out, tmp)
        res = out.download(inp.dims, np.uint8)
    finally:
        for b in (out, tmp):
            b.free()
        inp.free()
    return inp.place(res)


def _std_from_moments(n, s, q):
    """Population standard deviation of n integers with sum s and sum of squares
    q: the exact variance (n q - s^2) / n^2, rounded once, then sqrt.  np.std
    sums rounded squared deviations instead, which can differ in the last bits."""
    from fractions import Fraction
    return float(np.sqrt(float(Fraction(n * q - s * s, n * n))))


def _plane_rule(stats, all_labels):
    """refinement's per-plane decision (active_contour.py:211-224), on the
    per-plane moments: for each label l (in order) with voxels of l and of other
    values near the labels (d == 0), plane k becomes
    (d == 0 and p1 > p2) * l, p1 / p2 the Gaussian densities of the two groups
    at the voxel's value; a later label overwrites the whole plane.  Returns
    (lut uint8 [Z][256], label uint8 [Z])."""
    Z = stats.shape[0]
    lut = np.zeros((Z, 256), np.uint8)
    plane_label = np.zeros(Z, np.uint8)
    img = np.arange(256, dtype=np.uint8)
    total = stats[:, :, :].sum(axis=1, dtype=np.uint64)
    for l in all_labels[1:]:
        for k in range(Z):
            n1, s1_, q1 = (int(v) for v in stats[k, l])
            n2, s2_, q2 = (int(t) - int(v) for t, v in zip(total[k], stats[k, l]))
            if n1 == 0 or n2 == 0:
                continue
            m1, m2 = np.float64(s1_ / n1), np.float64(s2_ / n2)
            s1 = max(np.float64(_std_from_moments(n1, s1_, q1)), 1)
            s2 = max(np.float64(_std_from_moments(n2, s2_, q2)), 1)
            p1 = np.exp(-(img - m1) ** 2 / (2 * s1 ** 2)) / np.sqrt(2 * np.pi * s1 ** 2)
            p2 = np.exp(-(img - m2) ** 2 / (2 * s2 ** 2)) / np.sqrt(2 * np.pi * s2 ** 2)
            lut[k] = p1 > p2
            plane_label[k] = l
    return lut, plane_label


def refinement(data, labels, device=0):
    """activeContour(data, labels, simple=True): the labels refined plane by
    plane near their boundaries (active_contour.py:193-225).  The chamfer
    distance, the plane moments and the final assignment run on the device; the
    per-plane Gaussian decision is taken on the host for the 256 image values.
    None where the reference's pre_processing rejects the labels."""
    eng = _hip.engine(device)
    inp = _inputs(eng, data, labels)
    if inp is None:
        return None
    Z, Y, X = inp.dims
    n = Z * Y * X
    dist, d = eng.alloc(4 * n), eng.alloc(n)
    st, out = eng.alloc(Z * 256 * 3 * 8), eng.alloc(n)
    bufs = [dist, d, st, out]
    try:
        _hip.call("bmx_refine_distance", eng.ctx, inp.lab.ptr, Z, Y, X, dist.ptr, d.ptr)
        _hip.call("bmx_refine_stats", eng.ctx, inp.img.ptr, inp.lab.ptr, d.ptr, Z, Y, X, st.ptr)
        stats = st.download((Z, 256, 3), np.uint64)
        lut, plane_label = _plane_rule(stats, inp.all_labels)
        lt, pl = eng.upload(lut), eng.upload(plane_label)
        bufs += [lt, pl]
        _hip.call("bmx_refine_apply", eng.ctx, inp.img.ptr, d.ptr, Z, Y, X, lt.ptr, pl.ptr, out.ptr)
        res = out.download(inp.dims, np.uint8)
    finally:
        for b in bufs:
            b.free()
        inp.free()
    return inp.place(res)
