"""Host-side helpers of the smart-interpolation path (NumPy).

Each function restates the behaviour of the reference helper it names
(biomedisa/features/biomedisa_helper.py, biomedisa/interpolation.py); the
device work lives in libbmx_hip.
"""
import numpy as np


class Biomedisa:
    """Attribute bag carried through a call, as the reference's ``bm``."""


def unique(arr, return_counts=False):
    """Sorted unique values (biomedisa_helper.py:82-103).

    Non-negative integer arrays use a bincount, as the reference's Numba path.
    """
    if np.issubdtype(arr.dtype, np.integer) and arr.size and arr.min() >= 0:
        counts = np.bincount(arr.ravel().astype(np.int64, copy=False) if arr.dtype.itemsize > 4
                             or arr.dtype == np.uint32 else arr.ravel())
        labels = np.flatnonzero(counts)
        if return_counts:
            return labels, counts[labels]
        return labels
    return np.unique(arr, return_counts=return_counts)


def _error_(bm, message, level=1):
    """Print and flag an error (biomedisa_helper.py:622-637, no Django/Slicer sinks)."""
    if level == 0:
        print("Warning:", message)
    elif level == 1:
        print("Error:", message)
        bm.success = False
    return bm


def pre_processing(bm):
    """Label-set preparation of biomedisa_helper.py:639-724 (arrays given, no file loading).

    Sets bm.allLabels (sorted label values, 0 first) and bm.nol; may rewrite
    bm.labelData (float labels -> int32, values > 255 and removed labels -> 0).
    """
    if bm.data is None:
        return _error_(bm, "Invalid image data.")
    if bm.labelData is None:
        return _error_(bm, "Invalid label data.")
    if len(bm.labelData.shape) != 3:
        return _error_(bm, "Label data must be three-dimensional.")
    if bm.data.shape != bm.labelData.shape:
        return _error_(bm, "Image and label data must have the same x,y,z-dimensions.")

    if bm.labelData.dtype in ["float16", "float32", "float64"]:
        print(f"Warning: Potential label loss during conversion from {bm.labelData.dtype} to int32.")
        bm.labelData = bm.labelData.astype(np.int32)

    bm.allLabels = unique(bm.labelData)

    if np.any(bm.allLabels < 0):
        bm.allLabels = bm.allLabels[bm.allLabels >= 0]
        print("Warning: Negative labels indicate ignored regions.")

    if np.any(bm.allLabels > 255):
        bm.labelData = np.array(bm.labelData, copy=True)
        bm.labelData[bm.labelData > 255] = 0
        bm.allLabels = bm.allLabels[bm.allLabels <= 255]
        print("Warning: Only label values <=255 are allowed. Label values greater than 255 will be removed.")

    if not np.any(bm.allLabels == 0):
        bm.allLabels = np.append(0, bm.allLabels)

    labels_to_compute = str(bm.only).split(",")
    if not any(x in ["all", "All", "ALL"] for x in labels_to_compute):
        labels_to_remove = [k for k in bm.allLabels if str(k) not in labels_to_compute and k > 0]
        if labels_to_remove:
            bm.labelData = np.array(bm.labelData, copy=True)
        for k in labels_to_remove:
            bm.labelData[bm.labelData == k] = 0
            bm.allLabels = bm.allLabels[bm.allLabels != k]

    labels_to_remove = str(bm.ignore).split(",")
    if not any(x in ["none", "None", "NONE"] for x in labels_to_remove):
        for k in labels_to_remove:
            try:
                k = int(k)
            except ValueError:
                continue
            bm.labelData = np.array(bm.labelData, copy=True)
            bm.labelData[bm.labelData == k] = 0
            bm.allLabels = bm.allLabels[bm.allLabels != k]

    bm.nol = len(bm.allLabels)
    if bm.nol < 2:
        return _error_(bm, "No labeled slices found.")
    bm.success = True
    return bm


def pre_processing_counts(bm, counts):
    """pre_processing (biomedisa_helper.py:639-724) from the label-class counts
    of bmx_label_stats instead of the label array: counts[v] for the values
    1..255, counts[256] negative voxels, counts[257] voxels above 255.

    Sets bm.allLabels and bm.nol as pre_processing does and bm.kept, the label
    values of 1..255 that stay non-zero in the edited label volume, and
    bm.neg_kept (negative labels stay, as ignored regions).  Returns None when
    an `ignore` entry names a negative value present in the volume (the counts
    cannot tell which negative voxels it removes); the caller then works on
    the array itself.
    """
    counts = np.asarray(counts)
    present = np.flatnonzero(counts[1:256]) + 1
    has_neg, has_big = bool(counts[256]), bool(counts[257])
    bm.allLabels = present.astype(np.int64)
    if has_neg:
        print("Warning: Negative labels indicate ignored regions.")
    if has_big:
        print("Warning: Only label values <=255 are allowed. Label values greater than 255 will be removed.")
    bm.allLabels = np.append(0, bm.allLabels)
    labels_to_compute = str(bm.only).split(",")
    if not any(x in ["all", "All", "ALL"] for x in labels_to_compute):
        labels_to_remove = [k for k in bm.allLabels if str(k) not in labels_to_compute and k > 0]
        for k in labels_to_remove:
            bm.allLabels = bm.allLabels[bm.allLabels != k]
    neg_kept = has_neg
    labels_to_remove = str(bm.ignore).split(",")
    if not any(x in ["none", "None", "NONE"] for x in labels_to_remove):
        for k in labels_to_remove:
            try:
                k = int(k)
            except ValueError:
                continue
            if k < 0 and has_neg:
                return None
            bm.allLabels = bm.allLabels[bm.allLabels != k]
    bm.kept = np.asarray([v for v in bm.allLabels if v > 0], dtype=np.int64)
    bm.neg_kept = neg_kept
    bm.nol = len(bm.allLabels)
    if bm.nol < 2:
        return _error_(bm, "No labeled slices found.")
    bm.success = True
    return bm


def box_from_class_boxes(boxes, kept, neg_kept, shape):
    """predict_blocksize (biomedisa_helper.py:1105-1124) of the padded label
    volume, from the per-class boxes of bmx_label_stats (unpadded (Z,Y,X)
    coordinates, zmin > zmax when absent): the union of the boxes of the kept
    values and of the negative class, +1 for the pad, +-100 clamped."""
    sel = [int(v) for v in kept] + ([256] if neg_kept else [])
    b = np.asarray(boxes, dtype=np.int64)[sel] if sel else np.zeros((0, 6), np.int64)
    b = b[b[:, 0] <= b[:, 1]]
    zsh, ysh, xsh = (int(s) + 2 for s in shape)
    if b.shape[0] == 0:
        amin_z, amax_z, amin_y, amax_y, amin_x, amax_x = zsh, 0, ysh, 0, xsh, 0
    else:
        amin_z, amax_z = int(b[:, 0].min()) + 1, int(b[:, 1].max()) + 1
        amin_y, amax_y = int(b[:, 2].min()) + 1, int(b[:, 3].max()) + 1
        amin_x, amax_x = int(b[:, 4].min()) + 1, int(b[:, 5].max()) + 1

    def lo(v):
        return v - 100 if v - 100 > 0 else 0

    def hi(v, n):
        return v + 100 if v + 100 < n else n

    return lo(amin_z), hi(amax_z, zsh), lo(amin_y), hi(amax_y, ysh), lo(amin_x), hi(amax_x, xsh)


def class_lut(allLabels, neg_kept):
    """Label class -> label index (uint8[258]): the index of the value in
    allLabels (0 for values left out: background, removed or above 255),
    BMX_LABEL_NONE for kept negative labels; as label_index_volume."""
    lut = np.zeros(258, np.uint8)
    vals = np.asarray(allLabels, dtype=np.int64)
    lut[vals] = np.arange(len(vals), dtype=np.uint8)
    lut[257] = 0
    lut[256] = 0xFF if neg_kept else 0
    return lut


def dice_counted(label_dtype):
    """Per label value v, whether a voxel whose label and result are both v is
    a match of Dice_score: (ground_truth + result) > 0 wraps to 0 for v = 128
    when both are uint8 (biomedisa_helper.py:164)."""
    counted = np.ones(256, np.uint8)
    counted[0] = 0
    if np.dtype(label_dtype) == np.uint8:
        counted[128] = 0
    return counted


def dice_from_counts(a, b, c):
    """The global Dice of Dice_score from bmx_dice_counts (NaN when empty)."""
    return 2.0 * a / float(b + c) if b + c else float("nan")


def normalize_image(data):
    """Image dtype handling of interpolation.py:135-144.

    int8 -> uint8 (+128); uint8 kept; everything else -> float32 scaled to [0,255].
    """
    if data.dtype == np.int8:
        return (data.astype(np.int16) + 128).astype(np.uint8)
    if data.dtype == np.uint8:
        return data
    d = data.astype(np.float32)
    d -= np.amin(d)
    d /= np.amax(d)
    d *= 255.0
    return d


def label_index_volume(labelData, allLabels):
    """Label value -> label index (uint8): 0 = background, negative -> 0xFF.

    The index order is the order of ``allLabels``, which is the row order of the
    reference's hit volume (pyopencl_small.py:47-58).
    """
    lut = np.zeros(256, np.uint8)
    lut[np.asarray(allLabels, dtype=np.int64)] = np.arange(len(allLabels), dtype=np.uint8)
    if labelData.dtype == np.uint8:
        return lut[labelData]
    lab = labelData
    neg = lab < 0
    idx = lut[np.clip(lab, 0, 255).astype(np.intp, copy=False)]
    if neg.any():
        idx[neg] = 0xFF
    return idx


def predict_blocksize(labelData):
    """Bounding box of non-zero labels, +-100 voxels, clamped (biomedisa_helper.py:1105-1124).

    Returns (argmin_z, argmax_z, argmin_y, argmax_y, argmin_x, argmax_x) as slice bounds.
    """
    zsh, ysh, xsh = labelData.shape
    nzm = labelData != 0
    zs = np.flatnonzero(nzm.any(axis=(1, 2)))
    if zs.size == 0:
        amin_z, amax_z, amin_y, amax_y, amin_x, amax_x = zsh, 0, ysh, 0, xsh, 0
    else:
        ys = np.flatnonzero(nzm.any(axis=(0, 2)))
        xs = np.flatnonzero(nzm.any(axis=(0, 1)))
        amin_z, amax_z = int(zs[0]), int(zs[-1])
        amin_y, amax_y = int(ys[0]), int(ys[-1])
        amin_x, amax_x = int(xs[0]), int(xs[-1])

    def lo(v):
        return v - 100 if v - 100 > 0 else 0

    def hi(v, n):
        return v + 100 if v + 100 < n else n

    return lo(amin_z), hi(amax_z, zsh), lo(amin_y), hi(amax_y, ysh), lo(amin_x), hi(amax_x, xsh)


def adjacent_planes(flags):
    """True if two consecutive planes of one axis are labelled (interpolation.py:175-183)."""
    for f in flags:
        f = np.asarray(f, dtype=bool)
        if f.size > 1 and np.any(f[1:] & f[:-1]):
            return True
    return False


def get_labels(pre_final, labels):
    """Label index -> label value (biomedisa_helper.py:1209-1213)."""
    lut = np.zeros(max(256, len(labels)), dtype=pre_final.dtype)
    lut[:len(labels)] = labels
    lut[0] = 0
    return lut[pre_final]


def Dice_score(ground_truth, result, average_dice=False):
    """Dice of biomedisa_helper.py:155-165 (global, or mean over labels >= 1)."""
    if average_dice:
        dice = 0.0
        allLabels = unique(ground_truth)
        for l in allLabels[1:]:
            dice += 2 * np.logical_and(ground_truth == l, result == l).sum() / \
                float((ground_truth == l).sum() + (result == l).sum())
        dice /= float(len(allLabels) - 1)
    else:
        dice = 2 * np.logical_and(ground_truth == result, (ground_truth + result) > 0).sum() / \
            float((ground_truth > 0).sum() + (result > 0).sum())
    return dice


def per_label_dice(a, b, labels=None):
    """Dice of every label >= 1 present in either volume -> {label: dice}."""
    if labels is None:
        labels = np.union1d(unique(a), unique(b))
    out = {}
    for l in labels:
        if l == 0:
            continue
        ma, mb = a == l, b == l
        s = int(ma.sum()) + int(mb.sum())
        out[int(l)] = 1.0 if s == 0 else 2.0 * int(np.logical_and(ma, mb).sum()) / s
    return out
