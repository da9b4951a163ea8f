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


def img_to_uint8(img):
    """biomedisa_helper.py:345-352."""
    if img.dtype != np.uint8:
        img = img.astype(np.float32)
        img -= np.amin(img)
        img /= np.amax(img)
        img *= 255.0
        img = img.astype(np.uint8)
    return img


def _acwe_box(labels):
    """reduce_blocksize of active_contour.py:91-110: bbox of the non-zero labels
    grown by 100 and clamped."""
    zsh, ysh, xsh = labels.shape
    nz = np.nonzero(labels)
    lo = lambda v: v - 100 if v - 100 > 0 else 0
    hi = lambda v, n: v + 100 if v + 100 < n else n
    return (slice(lo(int(nz[0].min())), hi(int(nz[0].max()), zsh)),
            slice(lo(int(nz[1].min())), hi(int(nz[1].max()), ysh)),
            slice(lo(int(nz[2].min())), hi(int(nz[2].max()), xsh)))


def curvop(start, steps, label):
    """curvop_numba.py:124-148 with its start/final copies."""
    start = np.ascontiguousarray(start, np.uint8).copy()
    final = start.copy()
    for k in range(steps):
        if k % 2 == 0:
            orc.erosion(start, final, label)
            start = final.copy()
            orc.dilation(final, start, label)
            final = start.copy()
        else:
            orc.dilation(start, final, label)
            start = final.copy()
            orc.erosion(final, start, label)
            final = start.copy()
    return start


def acwe_labels(data, labels, alpha=1.0, smooth=1, steps=3):
    """The padded, cropped inputs and the acwe loop of active_contour.py:146-169
    -> (image box, label box before, label box after, box, padded shape)."""
    data = img_to_uint8(data)
    data = np.pad(data, 1)
    labels = np.pad(labels, 1)
    box = _acwe_box(labels)
    img = np.ascontiguousarray(data[box])
    lab = np.ascontiguousarray(labels[box])
    all_labels = np.unique(labels)
    all_labels = all_labels[all_labels >= 0]
    if not np.any(all_labels == 0):
        all_labels = np.concatenate([[0], all_labels])
    cur = lab.copy()
    for n in range(steps):
        mean = np.zeros(256, dtype=np.float32)
        for k in all_labels:
            inside = cur == k
            if np.any(inside):
                mean[k] = np.mean(img[inside])
        cur = orc.evolution(mean, cur, img, alpha)
        for k in all_labels:
            cur = curvop(cur, smooth, k)
    return img, lab, cur, box, data.shape, all_labels


def acwe(data, labels, alpha=1.0, smooth=1, steps=3):
    """activeContour(data, labels, alpha, smooth, steps) with simple=False
    (active_contour.py:112-184): the refined labels in the input's shape."""
    _, _, cur, box, shape, _ = acwe_labels(data, labels, alpha, smooth, steps)
    final = np.zeros(shape, np.uint8)
    final[box] = cur
    return np.ascontiguousarray(final[1:-1, 1:-1, 1:-1])


def refinement_labels(img, lab, all_labels):
    """refinement of active_contour.py:193-225 as written, on the padded,
    cropped uint8 image and labels."""
    zsh, ysh, xsh = img.shape
    distance = np.zeros(img.shape, dtype=np.float32)
    distance[lab == 0] = 100
    distance = orc.geodis(distance, np.sqrt(2), np.sqrt(3), 1)
    distance[distance <= 10] = 0
    distance[distance > 10] = 1
    distance = distance.astype(np.uint8)
    result = np.zeros_like(lab)
    for l in all_labels[1:]:
        for k in range(zsh):
            im = img[k]
            mask = lab[k]
            d = distance[k]
            if np.any(np.logical_and(d == 0, mask == l)) and np.any(np.logical_and(d == 0, mask != l)):
                m1, s1 = np.mean(im[np.logical_and(d == 0, mask == l)]), np.std(im[np.logical_and(d == 0, mask == l)])
                m2, s2 = np.mean(im[np.logical_and(d == 0, mask != l)]), np.std(im[np.logical_and(d == 0, mask != l)])
                s1 = max(s1, 1)
                s2 = max(s2, 1)
                p1 = np.exp(-(im - m1) ** 2 / (2 * s1 ** 2)) / np.sqrt(2 * np.pi * s1 ** 2)
                p2 = np.exp(-(im - m2) ** 2 / (2 * s2 ** 2)) / np.sqrt(2 * np.pi * s2 ** 2)
                result[k] = np.logical_and(d == 0, p1 > p2) * l
    return result


def refinement(data, labels):
    """activeContour(data, labels, simple=True) (active_contour.py:112-184)."""
    data = img_to_uint8(data)
    data = np.pad(data, 1)
    labels = np.pad(labels, 1)
    box = _acwe_box(labels)
    img = np.ascontiguousarray(data[box])
    lab = np.ascontiguousarray(labels[box])
    all_labels = np.unique(labels)
    all_labels = all_labels[all_labels >= 0]
    if not np.any(all_labels == 0):
        all_labels = np.concatenate([[0], all_labels])
    res = refinement_labels(img, lab, all_labels)
    final = np.zeros(data.shape, np.uint8)
    final[box] = res
    return np.ascontiguousarray(final[1:-1, 1:-1, 1:-1])
