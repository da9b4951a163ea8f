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
