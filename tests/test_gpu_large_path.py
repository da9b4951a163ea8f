"""The large-volume path on the GPU against the oracle: plane readers, the
adaptive-walk mask (three separable stride-5 passes against the reference's
brute-force scan), and the whole call with path='large'."""
import numpy as np
import pytest

import oracle.pipeline as opipe
from oracle import orc
from biomedisa_amd import _hip
from tests.test_large_oracle import two_far_regions

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _labels(rng, shape, every, allaxis, nlab, neg=0.0):
    gt = rng.integers(0, nlab + 1, size=shape).astype(np.int32)
    lab = np.zeros_like(gt)
    lab[1::every] = gt[1::every]
    if allaxis:
        lab[:, 2::every] = gt[:, 2::every]
        lab[:, :, 3::every] = gt[:, :, 3::every]
    lab[(rng.random(shape) < neg) & (lab > 0)] = -1
    return lab


def _index(lab):
    return np.where(lab < 0, 0xFF, lab).astype(np.uint8)


@pytest.mark.parametrize("allaxis", [False, True])
def test_plane_flags_large_bitexact(eng, allaxis):
    rng = np.random.default_rng(21)
    for t in range(4):
        shape = tuple(int(s) for s in rng.integers(12, 40, 3))
        every = int(rng.integers(2, 7))
        lab = _labels(rng, shape, every, allaxis, 3, neg=0.02)
        if not allaxis:
            lab[every + 2] = lab[1]   # a plane 1..4 after a labelled one is hidden
        _say("large plane flags allaxis=%s shape=%s every=%d" % (allaxis, shape, every))
        lb = eng.upload(_index(lab))
        fb = eng.plane_flags(lb, shape, allaxis, large=True)
        got = fb.download((sum(shape),), np.uint8)
        want = np.concatenate(opipe.plane_flags_large(lab, allaxis))
        np.testing.assert_array_equal(got, want)


def _mask_vs_oracle(eng, labidx, flags, dims):
    lb = eng.upload(np.asarray(labidx).astype(np.uint8))   # the device's uint8 indices (0xFF: ignored)
    fb = eng.upload(np.concatenate(flags).astype(np.uint8))
    mask = eng.adaptive_mask(lb, dims, fb).download(dims, np.uint8)
    pos, _, _ = orc.seeds(labidx, np.zeros(dims, np.uint8), flags)
    want = orc.adaptive(labidx, flags, pos)
    np.testing.assert_array_equal(mask.ravel()[pos].astype(bool), want)
    return want


def test_adaptive_mask_far_regions(eng):
    """Seeds 100 voxels away from other labels walk only on the sub-lattice."""
    img, labels = two_far_regions()
    p = opipe.prepare(img, labels)
    flags = opipe.plane_flags_large(p["lab"], False)
    _say("adaptive mask, far regions %s" % (p["labidx"].shape,))
    want = _mask_vs_oracle(eng, p["labidx"], flags, p["labidx"].shape)
    assert 0.3 < want.mean() < 1.0


@pytest.mark.parametrize("case", range(4))
def test_adaptive_mask_random_bitexact(eng, case):
    """Random label volumes (XY or allaxis planes, negative labels, up to 5
    labels, extents past 200 so that the +-100 window is cut)."""
    rng = np.random.default_rng(100 + case)
    shape = [int(s) for s in rng.integers(16, 40, 3)]
    shape[case % 3] = 210
    shape = tuple(shape)
    allaxis = case % 2 == 1
    nlab = int(rng.integers(1, 6))
    lab = np.zeros(shape, np.int32)
    # few, sparse labelled voxels so that many seeds see only their own label
    sparse = _labels(rng, shape, 7, allaxis, nlab, neg=0.05)
    keep = rng.random(shape) < 0.02
    lab[keep] = sparse[keep]
    for z in range(1, shape[0], 7):   # labelled planes mostly of label 1
        lab[z][lab[z] == 0] = 1
    labidx = _index(lab)
    flags = orc.plane_flags(labidx, allaxis)
    _say("adaptive mask random case %d shape %s allaxis %s labels %d" % (case, shape, allaxis, nlab))
    want = _mask_vs_oracle(eng, labidx, flags, shape)
    assert 0.0 < want.mean() < 1.0


def test_smart_interpolation_large_path_matches_oracle():
    """path='large': regular, smooth (first label clamped) and uncertainty
    bit-exact against the oracle's large path; fewer seeds walk than in the
    small path."""
    from biomedisa_amd.interpolation import smart_interpolation
    img, labels = two_far_regions()
    kw = dict(nbrw=2, sorw=60, smooth=3, uncertainty=True)
    _say("smart_interpolation path=large on %s" % (img.shape,))
    got = smart_interpolation(img, labels, path="large", **kw)
    want = opipe.smart_interpolation(img, labels, path="large", **kw)
    for k in ("regular", "smooth", "uncertainty"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    small = opipe.smart_interpolation(img, labels, path="small", **kw)
    assert want["_nseeds"] < small["_nseeds"]
    assert (small["smooth"] != want["smooth"]).any() or (small["regular"] != want["regular"]).any()


def test_auto_path_is_small_for_small_boxes():
    from biomedisa_amd.interpolation import smart_interpolation
    from biomedisa_amd.synth import ellipsoids_volume
    vol = ellipsoids_volume(shape=(30, 34, 38), seed=3, semi=(5, 10), every=4)
    a = smart_interpolation(vol.img, vol.labels, nbrw=2, sorw=80)
    b = smart_interpolation(vol.img, vol.labels, nbrw=2, sorw=80, path="small")
    np.testing.assert_array_equal(a["regular"], b["regular"])
