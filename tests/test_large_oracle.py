"""Pins of the oracle's large-volume path (plane readers, adaptive walks, the
clamp of the first smoothed map) by hand-built cases and closed forms."""
import numpy as np

import oracle.pipeline as opipe
from oracle import orc


def test_xy_planes_hide_the_next_four():
    """read_labeled_slices_large: scanning z, a labelled plane hides the 4 after it;
    negative labels count as labelled (np.any)."""
    lab = np.zeros((30, 6, 7), np.int32)
    for z in (3, 5, 9, 12, 13, 20):
        lab[z, 2, 3] = 1
    lab[25, 1, 1] = -1
    fz, fy, fx = opipe.plane_flags_large(lab, allaxis=False)
    assert list(np.nonzero(fz)[0]) == [3, 9, 20, 25]
    assert not fy.any() and not fx.any()


def test_allaxis_planes_count_positive_labels_only():
    """read_labeled_slices_allx_large counts arr > 0; without negative labels its
    planes equal the small path's (the C plane detector of the oracle)."""
    rng = np.random.default_rng(5)
    for t in range(6):
        shape = tuple(int(s) for s in rng.integers(8, 20, 3))
        gt = rng.integers(0, 4, shape).astype(np.int32)
        lab = np.zeros_like(gt)
        lab[1::4] = gt[1::4]
        lab[:, 2::5] = gt[:, 2::5]
        lab[:, :, 3::6] = gt[:, :, 3::6]
        large = opipe.plane_flags_large(lab, allaxis=True)
        small = orc.plane_flags(lab.astype(np.uint8), allaxis=True)
        for a, b in zip(large, small):
            np.testing.assert_array_equal(a, b)
    # a y-plane holding only negative labels is labelled for the small path only
    lab = np.zeros((9, 9, 9), np.int32)
    lab[:, 4, :] = -1
    lab[2, 2, 2:5] = 1
    lab[2, 3, 2:5] = 1
    large = opipe.plane_flags_large(lab, allaxis=True)
    small = orc.plane_flags(np.where(lab < 0, 0xFF, lab).astype(np.uint8), allaxis=True)
    assert small[1][4] == 1 and large[1][4] == 0


def _one_plane_volume(shape, z0, other, lab_main=1, lab_other=2):
    Z, Y, X = shape
    labidx = np.zeros(shape, np.uint8)
    labidx[z0, :, :] = lab_main
    labidx[(z0,) + other] = lab_other
    flags = (np.eye(1, Z, z0, dtype=np.uint8)[0], np.zeros(Y, np.uint8), np.zeros(X, np.uint8))
    return labidx, flags


def test_adaptive_walks_closed_form():
    """One XY plane of label 1 with a single voxel of label 2: a label-1 seed at
    (z0, y, x) walks iff it is on the sub-lattice (x+z0, y+z0 even) or the
    label-2 voxel lies on its stride-5 grid within 100 voxels per axis."""
    shape, z0, (yo, xo) = (21, 40, 260), 10, (20, 30)
    labidx, flags = _one_plane_volume(shape, z0, (yo, xo))
    Z, Y, X = shape
    yy, xx = np.meshgrid(np.arange(1, Y - 1), np.arange(1, X - 1), indexing="ij")
    pos = (z0 * Y * X + yy * X + xx).ravel().astype(np.int64)
    got = orc.adaptive(labidx, flags, pos).reshape(yy.shape)
    lattice = ((xx + z0) % 2 == 0) & ((yy + z0) % 2 == 0)
    sees = ((yy - yo) % 5 == 0) & ((xx - xo) % 5 == 0) & (np.abs(xx - xo) <= 100) & (np.abs(yy - yo) <= 100)
    sees[yo - 1, xo - 1] = False   # the label-2 seed itself: own label
    want = lattice | sees
    want[yo - 1, xo - 1] = True    # ... but it is on the sub-lattice
    np.testing.assert_array_equal(got, want)
    assert 0 < got.mean() < 0.5


def test_adaptive_walks_ignore_off_plane_and_edges():
    """Voxels off the labelled planes are ignored, and so are non-interior probes:
    a label-2 voxel on the volume's edge is never seen."""
    shape, z0 = (15, 30, 30), 7
    labidx, flags = _one_plane_volume(shape, z0, (0, 10))      # edge row y = 0
    labidx[3, 10, 10] = 2                                       # off-plane
    Z, Y, X = shape
    yy, xx = np.meshgrid(np.arange(1, Y - 1), np.arange(1, X - 1), indexing="ij")
    pos = (z0 * Y * X + yy * X + xx).ravel().astype(np.int64)
    got = orc.adaptive(labidx, flags, pos).reshape(yy.shape)
    lattice = ((xx + z0) % 2 == 0) & ((yy + z0) % 2 == 0)
    np.testing.assert_array_equal(got, lattice)


def two_far_regions(seed=4):
    """Labels 1 (x < 20) and 2 (x >= 310) on the XY planes z = 4, 9, 14 of a
    20 x 30 x 330 volume, background between: background seeds 100 voxels
    away from both walk only on the sub-lattice."""
    rng = np.random.default_rng(seed)
    img = rng.integers(60, 90, (20, 30, 330)).astype(np.uint8)
    img[:, :, :20] += 80
    img[:, :, 310:] += 120
    labels = np.zeros(img.shape, np.uint8)
    for z in (4, 9, 14):
        labels[z, :, :] = 0
        labels[z, 2:28, 1:20] = 1
        labels[z, 2:28, 310:329] = 2
    return img, labels


def test_large_path_walks_the_adaptive_subset_and_clamps_smoothing():
    """The oracle pipeline's large path walks exactly the seeds passing the
    adaptive test, and its 'smooth' differs from the small path's only through
    the clamp of the first label at 0 (here it changes voxels)."""
    img, labels = two_far_regions()
    kw = dict(nbrw=2, sorw=60, smooth=3)
    small = opipe.smart_interpolation(img, labels, path="small", **kw)
    large = opipe.smart_interpolation(img, labels, path="large", **kw)
    p = large["_prep"]
    flags = opipe.plane_flags_large(p["lab"], False)
    pos, _, _ = orc.seeds(p["labidx"], p["img"], flags)
    keep = orc.adaptive(p["labidx"], flags, pos)
    assert large["_nseeds"] == int(keep.sum()) < small["_nseeds"]
    # smoothing of the large path's hits without the clamp
    phi = np.stack([orc.smooth(h.astype(np.float32), 3) for h in large["_hits_u32"]])
    unclamped = opipe._uncrop(orc.argmax(phi, p["values"]), p)
    assert (phi[0] < 0).any()
    phi[0][phi[0] < 0] = 0
    np.testing.assert_array_equal(large["smooth"], opipe._uncrop(orc.argmax(phi, p["values"]), p))
    assert (unclamped != large["smooth"]).any()


def test_path_rule():
    """interpolation.py:192-196 on box sizes."""
    p = {"img": np.zeros((10, 10, 10), np.uint8), "nol": 3}
    assert opipe.select_path(p) == "small"

    class Big:
        shape = (1026, 1026, 1026)
    assert opipe.select_path({"img": Big, "nol": 2}) == "large"
    assert opipe.select_path({"img": type("M", (), {"shape": (514, 514, 514)}), "nol": 20}) == "large"
    assert opipe.select_path({"img": type("M", (), {"shape": (514, 514, 514)}), "nol": 9}) == "small"
