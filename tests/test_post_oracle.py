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


# --- active contour (curvop_numba.py) -------------------------------------

def test_evolution_takes_the_cheapest_neighbour_label():
    A = np.ones((5, 5, 5), np.uint8)
    data = np.full(A.shape, 100, np.uint8)
    mean = np.zeros(256, np.float32)
    mean[1], mean[2], mean[3] = 100.0, 12.0, 8.0
    data[2, 2, 2] = 10
    A[1, 2, 2] = 2                     # cost |12 - 10| = 2
    A[3, 3, 3] = 3                     # cost |8 - 10| = 2: a tie, later in order
    B = orc.evolution(mean, A, data, 1.0)
    assert B[2, 2, 2] == 2             # the first strict minimum in (dz, dy, dx) order
    A[1, 2, 2] = 0                     # label 0 costs alpha * |mean[0] - 10| = alpha * 10
    B = orc.evolution(mean, A, data, 1.0)
    assert B[2, 2, 2] == 3
    B = orc.evolution(mean, A, data, 0.1)   # 0.1 * 10 = 1 < 2
    assert B[2, 2, 2] == 0
    assert (B[0] == A[0]).all() and (B[:, 0] == A[:, 0]).all()   # only interior voxels change


def test_erosion_of_a_cube_leaves_a_cross_and_dilation_grows_a_voxel():
    """A 3^3 cube: only voxels with a whole planar 3x3 pattern of the label keep
    it - the centre and the six face centres; corners and edges take the other
    label.  Dilation turns the 26-neighbourhood of a voxel into the label."""
    start = np.zeros((7, 7, 7), np.uint8)
    start[2:5, 2:5, 2:5] = 1
    final = start.copy()
    orc.erosion(start, final, 1)
    want = np.zeros_like(start)
    want[3, 3, 3] = 1
    for d in range(3):
        for s in (-1, 1):
            c = [3, 3, 3]
            c[d] += s
            want[tuple(c)] = 1
    np.testing.assert_array_equal(final, want)
    one = np.zeros((5, 5, 5), np.uint8)
    one[2, 2, 2] = 4
    out = one.copy()
    orc.dilation(one, out, 4)
    want = np.zeros_like(one)
    want[1:4, 1:4, 1:4] = 4
    np.testing.assert_array_equal(out, want)


def test_curvop_rounds_alternate():
    """Even rounds erode then dilate, odd rounds dilate then erode: one round
    removes a lone voxel, two rounds of a 3^3 cube end where erosion of its
    opening leaves it."""
    v = np.zeros((9, 9, 9), np.uint8)
    v[4, 4, 4] = 1
    np.testing.assert_array_equal(post.curvop(v, 1, 1), np.zeros_like(v))
    cube = np.zeros((9, 9, 9), np.uint8)
    cube[3:6, 3:6, 3:6] = 2
    r1 = post.curvop(cube, 1, 2)
    e = cube.copy()
    orc.erosion(cube, e, 2)
    d = e.copy()
    orc.dilation(e, d, 2)
    np.testing.assert_array_equal(r1, d)
    r2 = post.curvop(cube, 2, 2)
    d2 = d.copy()
    orc.dilation(d, d2, 2)
    e2 = d2.copy()
    orc.erosion(d2, e2, 2)
    np.testing.assert_array_equal(r2, e2)


def test_acwe_means_are_float32_of_the_float64_mean():
    data = np.array([[[1, 2], [2, 2]]], np.uint8)
    labels = np.array([[[1, 1], [1, 0]]], np.uint8)
    img, lab, cur, box, shape, all_labels = post.acwe_labels(data, labels, 1.0, 0, 1)
    assert list(all_labels) == [0, 1]
    assert np.float32(np.mean(np.array([1, 2, 2], np.uint8))) == np.float32(5.0 / 3.0)


# --- refinement (active_contour.py:193-225) --------------------------------

def _chamfer(off):
    a, b, c = sorted((abs(int(v)) for v in off), reverse=True)
    return c * np.sqrt(3) + (b - c) * np.sqrt(2) + (a - b)


def test_geodis_is_the_chamfer_distance_from_a_source():
    """One forward and one backward scan give the 1/sqrt2/sqrt3 chamfer distance
    (largest offset minus middle, plus sqrt2 x middle minus smallest, plus sqrt3 x
    smallest) from a single source, to float32 rounding of the partial sums."""
    c = np.full((9, 9, 9), 100, np.float32)
    c[4, 4, 4] = 0
    orc.geodis(c, np.sqrt(2), np.sqrt(3), 1)
    for z in range(9):
        for y in range(1, 8):
            for x in range(1, 8):
                assert abs(c[z, y, x] - _chamfer((z - 4, y - 4, x - 4))) < 1e-5, (z, y, x)
    assert (c[:, 0, :] == 100).all() and (c[:, :, 8] == 100).all()   # the rim is never scanned


def test_geodis_two_sources_take_the_nearer():
    c = np.full((12, 12, 12), 100, np.float32)
    c[3, 3, 3] = c[8, 9, 7] = 0
    orc.geodis(c, np.sqrt(2), np.sqrt(3), 1)
    for z in range(1, 11):
        for y in range(1, 11):
            for x in range(1, 11):
                want = min(_chamfer((z - 3, y - 3, x - 3)), _chamfer((z - 8, y - 9, x - 7)))
                assert abs(c[z, y, x] - want) < 1e-5


def test_refinement_last_label_takes_the_plane():
    """Labels 1 and 2 side by side in one plane, both with near and far pixels:
    the plane's result is that of label 2 alone (its assignment overwrites the
    plane); pixels far from every label (distance > 10) stay 0."""
    img = np.zeros((3, 40, 40), np.uint8)
    lab = np.zeros((3, 40, 40), np.uint8)
    lab[1, 5:20, 5:15] = 1
    lab[1, 5:20, 15:25] = 2
    img[1] = 30
    img[1, 5:20, 5:15] = 100
    img[1, 5:20, 15:25] = 200
    res = post.refinement(img, lab)
    assert set(np.unique(res[1])) <= {0, 2}
    assert (res[1][lab[1] == 2] == 2).all()
    assert (res[1][lab[1] == 1] == 0).all()
    assert (res[1, 35:, 35:] == 0).all()
