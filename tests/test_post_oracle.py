"""Pins of the oracle's clean/fill (remove_outlier.py:41-139): its component
labelling against scipy.ndimage.label, and hand-built cases whose results are
known without running either."""
import numpy as np
from scipy import ndimage

from oracle import orc
from oracle import post

S6 = [[[0, 0, 0], [0, 1, 0], [0, 0, 0]], [[0, 1, 0], [1, 1, 1], [0, 1, 0]], [[0, 0, 0], [0, 1, 0], [0, 0, 0]]]


def test_label6_matches_scipy_numbering():
    rng = np.random.default_rng(3)
    for t in range(20):
        shape = tuple(int(s) for s in rng.integers(1, 14, 3))
        mask = rng.random(shape) < rng.uniform(0.2, 0.7)
        got, n = orc.label6(mask)
        want, nw = ndimage.label(mask, structure=S6)
        assert n == nw
        np.testing.assert_array_equal(got, want)


def test_clean_removes_small_components_only():
    img = np.zeros((12, 14, 16), np.uint8)
    img[2:8, 2:8, 2:8] = 3            # 216 voxels
    img[10, 10, 10] = 3               # 1 voxel: 1 <= 0.1 * 216, removed
    img[9:11, 1:4, 12:15] = 3         # 18 voxels: kept (> 21.6? no: removed)
    img[1:4, 10:13, 10:13] = 5        # label 5: one component, kept
    img[5, 12, 2] = 5                 # 1 <= 2.7: removed
    out = post.clean(img, 0.1)
    want = img.copy()
    want[10, 10, 10] = 0
    want[9:11, 1:4, 12:15] = 0
    want[5, 12, 2] = 0
    np.testing.assert_array_equal(out, want)
    out2 = post.clean(img, 0.05)      # 18 > 10.8: kept now
    want2 = want.copy()
    want2[9:11, 1:4, 12:15] = 3
    np.testing.assert_array_equal(out2, want2)


def test_clean_diagonal_neighbours_are_separate():
    img = np.zeros((5, 5, 5), np.uint8)
    img[1, 1, 1] = img[2, 2, 2] = img[3, 3, 3] = 1   # three 1-voxel components (no face contact)
    img[1:4, 3, 1] = 1                               # a 3-voxel rod
    out = post.clean(img, 0.5)                        # keep components > 1.5 voxels
    want = np.zeros_like(img)
    want[1:4, 3, 1] = 1
    np.testing.assert_array_equal(out, want)


def _hollow_cube(n_box, lo, hi, k):
    """A one-voxel-thick cube shell of value k: outer [lo, hi)^3, cavity inside."""
    img = np.zeros((n_box,) * 3, np.uint8)
    img[lo:hi, lo:hi, lo:hi] = k
    img[lo + 1:hi - 1, lo + 1:hi - 1, lo + 1:hi - 1] = 0
    return img


def test_fill_fills_enclosed_holes():
    """Shell of 296 voxels (outer 8^3) in a 10^3 box whose outside has 488: the
    216-voxel cavity, another label's voxel included, becomes the shell's value
    (216 <= 0.9 * 296); the other label is then gone and is skipped."""
    img = _hollow_cube(12, 2, 10, 4)
    img[5, 5, 5] = 6
    out = post.fill(img, 0.9)
    want = img.copy()
    want[3:9, 3:9, 3:9] = 4
    np.testing.assert_array_equal(out, want)


def test_fill_keeps_large_holes():
    """Shell of 728 voxels (outer 12^3), 1000-voxel cavity, outside 1016."""
    img = _hollow_cube(14, 1, 13, 2)
    np.testing.assert_array_equal(post.fill(img, 0.9), img)     # 1000 > 655.2: kept
    want = img.copy()
    want[2:12, 2:12, 2:12] = 2
    np.testing.assert_array_equal(post.fill(img, 1.5), want)    # 1000 <= 1092: filled


def test_fill_quirk_when_the_label_outnumbers_its_outside():
    """When the voxels of k outnumber every component of "not k" in the box
    (a solid cube in a 1-voxel shell), np.argmax(size) picks index 0 - the
    label itself: the label is removed and its shell takes its value."""
    img = np.zeros((12, 12, 12), np.uint8)
    img[1:11, 1:11, 1:11] = 1          # 1000 voxels; box 12^3 has a 728-voxel shell
    out = post.fill(img, 0.9)
    want = np.ones_like(img)
    want[1:11, 1:11, 1:11] = 0
    np.testing.assert_array_equal(out, want)


def test_unique_skips_the_smallest_present_value():
    """unique(...)[1:]: without background the smallest label is left as is."""
    img = np.full((6, 6, 6), 2, np.uint8)
    img[0, 0, 0] = 3
    img[5, 5, 5] = 3
    out = post.clean(img, 0.6)        # label 3: two 1-voxel components, 1 > 0.6: both kept
    np.testing.assert_array_equal(out, img)
    out = post.clean(img, 1.0)        # 1 > 1.0 false: both removed to 0
    want = img.copy()
    want[0, 0, 0] = want[5, 5, 5] = 0
    np.testing.assert_array_equal(out, want)


def test_cleaned_filled_wraps_in_uint8():
    final = np.array([5, 5, 0, 7], np.uint8)
    cleaned = np.array([5, 0, 0, 7], np.uint8)
    filled = np.array([5, 9, 9, 7], np.uint8)
    np.testing.assert_array_equal(post.cleaned_filled(final, cleaned, filled), [5, 4, 9, 7])
