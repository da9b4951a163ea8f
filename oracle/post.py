"""CPU restatement of the label post-processing (test infrastructure).

clean and fill follow biomedisa/features/remove_outlier.py:41-139 step by
step on NumPy arrays; the 6-connected component labelling is the oracle's own
breadth-first labelling (orc.label6), numbered like scipy.ndimage.label.
"""
import numpy as np

from . import orc


def _box(mask):
    """reduce_blocksize (remove_outlier.py:41-58): the box of the non-zero
    voxels grown by one voxel and clamped, as slices."""
    zsh, ysh, xsh = mask.shape
    nz = np.nonzero(mask)
    z0, z1 = int(nz[0].min()), int(nz[0].max())
    y0, y1 = int(nz[1].min()), int(nz[1].max())
    x0, x1 = int(nz[2].min()), int(nz[2].max())
    return (slice(max(z0 - 1, 0), min(z1 + 1, zsh - 1) + 1), slice(max(y0 - 1, 0), min(y1 + 1, ysh - 1) + 1),
            slice(max(x0 - 1, 0), min(x1 + 1, xsh - 1) + 1))


def clean(image, threshold=0.1):
    """remove_outlier.py:60-95."""
    img = np.copy(image, order="C")
    for k in np.unique(img)[1:]:
        label = img == k
        box = _box(label)
        labeled, _ = orc.label6(label[box])
        size = np.bincount(labeled.ravel())
        biggest = int(np.argmax(size[1:])) + 1
        label_size = size[biggest]
        # for l, m in enumerate(size[1:]): if m > threshold * label_size: keep l+1
        keep = np.concatenate([[0], (size[1:] > threshold * label_size).astype(np.uint8)])
        reduced = keep[labeled]
        mask = np.zeros(img.shape, np.uint8)
        mask[box] = reduced
        img[label] = 0
        img[mask == 1] = k
    return img


def fill(image, threshold=0.9):
    """remove_outlier.py:97-139 (labels in ascending order, each on the volume
    as changed by the ones before it)."""
    img = np.copy(image, order="C")
    for k in np.unique(img)[1:]:
        label = img == k
        if not label.any():
            continue
        box = _box(label)
        reduced = label[box].astype(np.uint8)
        label_size = int(reduced.sum())
        inverse = 1 - reduced
        labeled, _ = orc.label6(inverse)
        size = np.bincount(labeled.ravel())
        biggest = int(np.argmax(size))
        # 1 everywhere, 0 on the biggest entry and on components with
        # m > threshold * label_size (l + 1 != biggest)
        one = np.ones(size.size, np.uint8)
        one[biggest] = 0
        one[1:][size[1:] > threshold * label_size] = 0
        reduced = one[labeled]
        mask = np.zeros(img.shape, np.uint8)
        mask[box] = reduced
        img[label] = 0
        img[mask == 1] = k
    return img


def cleaned_filled(final, cleaned, filled):
    """rw_small.py:238: cleaned + (filled - final) in uint8 arithmetic."""
    return (cleaned + (filled - final)).astype(np.uint8)
