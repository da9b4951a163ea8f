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


def _all_labels(labels):
    """pre_processing's label set (biomedisa_helper.py:671-694): unique values,
    0 first."""
    vals = np.unique(labels)
    vals = vals[(vals >= 0) & (vals <= 255)]
    if not np.any(vals == 0):
        vals = np.concatenate([[0], vals])
    return vals.astype(np.uint8)


def _box(labels):
    """reduce_blocksize of active_contour.py:91-110: bbox of the non-zero labels
    grown by 100 and clamped."""
    zsh, ysh, xsh = labels.shape
    nz = [np.nonzero(np.any(labels != 0, axis=tuple(a for a in range(3) if a != d)))[0] for d in range(3)]
    lo = lambda v: v - 100 if v - 100 > 0 else 0
    hi = lambda v, n: v + 100 if v + 100 < n else n
    return tuple(slice(lo(int(n.min())), hi(int(n.max()), s)) for n, s in zip(nz, (zsh, ysh, xsh)))


def _prepare(data, labels):
    """activeContour's checks and inputs (active_contour.py:112-156): None when
    pre_processing rejects them, else (uint8 image box, label box, box, padded
    shape, label set)."""
    if data.shape != labels.shape or labels.ndim != 3:
        return None
    labels = np.asarray(labels)
    if labels.dtype.kind == "f":
        labels = labels.astype(np.int32)
    all_labels = _all_labels(labels)
    if len(all_labels) < 2:
        return None
    img = np.pad(img_to_uint8(data), 1)
    lab = np.pad(labels, 1)
    box = _box(lab)
    return (np.ascontiguousarray(img[box]), np.ascontiguousarray(lab[box]).astype(np.uint8), box, img.shape,
            all_labels)


def active_contour(data, labels, alpha=1.0, smooth=1, steps=3, device=0):
    """activeContour(data, labels, alpha, smooth, steps) with simple=False: the
    acwe loop on the device (bmx_acwe).  Returns the uint8 labels in the
    input's shape, or None where the reference's pre_processing rejects the
    labels (fewer than two values)."""
    prep = _prepare(data, labels)
    if prep is None:
        return None
    img, lab, box, shape, all_labels = prep
    eng = _hip.engine(device)
    ib, lb = eng.upload(img), eng.upload(lab)
    out, tmp = eng.alloc(lab.size), eng.alloc(lab.size)
    try:
        eng.acwe(ib, lb, lab.shape, all_labels, alpha, smooth, steps, out, tmp)
        res = out.download(lab.shape, np.uint8)
    finally:
        for b in (ib, lb, out, tmp):
            b.free()
    final = np.zeros(shape, np.uint8)
    final[box] = res
    return np.ascontiguousarray(final[1:-1, 1:-1, 1:-1])
