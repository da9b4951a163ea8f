"""clean / fill on the GPU (bmx_clean, bmx_fill: union-find components) against
the oracle's restatement of remove_outlier.py, bit-exact."""
import numpy as np
import pytest

import oracle.pipeline as opipe
from oracle import post
from biomedisa_amd import _hip
from tests.test_post_oracle import _hollow_cube

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _gpu(eng, name, vol, threshold):
    buf = eng.upload(np.ascontiguousarray(vol, dtype=np.uint8))
    out = eng.alloc(vol.size)
    getattr(eng, name)(buf, vol.shape, threshold, out)
    return out.download(vol.shape, np.uint8)


def _blobs(rng, shape, nlab, density):
    """Random label volume: smoothed noise thresholded into several labels, so
    there are large components, small outliers and holes."""
    from scipy import ndimage
    noise = ndimage.uniform_filter(rng.random(shape), size=3)
    q = np.quantile(noise, np.linspace(1 - density, 1, nlab + 1))
    vol = np.zeros(shape, np.uint8)
    for k in range(nlab):
        vol[(noise >= q[k]) & (noise < q[k + 1])] = k + 1
    speck = rng.random(shape) < 0.01
    vol[speck] = rng.integers(0, nlab + 1, int(speck.sum()))
    return vol


@pytest.mark.parametrize("case", range(6))
def test_clean_and_fill_random_bitexact(eng, case):
    rng = np.random.default_rng(300 + case)
    shape = tuple(int(s) for s in rng.integers(10, 48, 3))
    vol = _blobs(rng, shape, int(rng.integers(1, 6)), float(rng.uniform(0.3, 0.9)))
    for name, t in (("clean", 0.1), ("clean", 0.01), ("fill", 0.9), ("fill", 0.05)):
        _say("%s threshold %g case %d shape %s" % (name, t, case, shape))
        got = _gpu(eng, name, vol, t)
        want = getattr(post, name)(vol, t)
        np.testing.assert_array_equal(got, want, err_msg="%s %g" % (name, t))


def test_fill_hand_cases(eng):
    for vol, t in ((_hollow_cube(12, 2, 10, 4), 0.9), (_hollow_cube(14, 1, 13, 2), 0.9),
                   (_hollow_cube(14, 1, 13, 2), 1.5)):
        vol = vol.copy()
        vol[vol.shape[0] // 2, 5, 5] = 6
        np.testing.assert_array_equal(_gpu(eng, "fill", vol, t), post.fill(vol, t))
    cube = np.zeros((12, 12, 12), np.uint8)
    cube[1:11, 1:11, 1:11] = 1        # the label outnumbers its outside: removed, shell filled
    np.testing.assert_array_equal(_gpu(eng, "fill", cube, 0.9), post.fill(cube, 0.9))


def test_components_in_a_large_box(eng):
    """A 160 x 200 x 220 volume (7 M voxels) of many components: the union-find
    across all XCDs gives the oracle's components."""
    rng = np.random.default_rng(9)
    vol = _blobs(rng, (160, 200, 220), 3, 0.6)
    _say("clean/fill on %s" % (vol.shape,))
    np.testing.assert_array_equal(_gpu(eng, "clean", vol, 0.05), post.clean(vol, 0.05))
    np.testing.assert_array_equal(_gpu(eng, "fill", vol, 0.5), post.fill(vol, 0.5))


def test_smart_interpolation_clean_fill_keys():
    """smart_interpolation(clean=, fill=) returns the reference's post-processing
    keys equal to the oracle pipeline's."""
    from biomedisa_amd.interpolation import smart_interpolation
    from biomedisa_amd.synth import structures_volume
    vol = structures_volume(shape=(36, 40, 44), seed=12, n_struct=3, radius=(4, 10), every=4, allaxis=False)
    kw = dict(nbrw=3, sorw=150, smooth=2, clean=0.1, fill=0.9)
    got = smart_interpolation(vol.img, vol.labels, **kw)
    want = opipe.smart_interpolation(vol.img, vol.labels, **kw)
    for k in ("regular", "smooth", "cleaned", "smooth_cleaned", "filled", "cleaned_filled"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)


@pytest.mark.parametrize("case", range(4))
def test_acwe_bitexact(eng, case):
    """bmx_acwe (means, evolution, per-label erosion/dilation rounds) against the
    oracle's restatement of curvop_numba.py on random label volumes."""
    from biomedisa_amd.post import active_contour
    rng = np.random.default_rng(400 + case)
    shape = tuple(int(s) for s in rng.integers(14, 40, 3))
    labels = _blobs(rng, shape, int(rng.integers(1, 5)), float(rng.uniform(0.3, 0.8)))
    img = (labels.astype(np.float32) * 40 + rng.normal(0, 25, shape)).clip(0, 255).astype(np.uint8)
    if case == 3:
        img = img.astype(np.float32) * 3.7   # float input: img_to_uint8 on both sides
    alpha, smooth, steps = [(1.0, 1, 3), (0.5, 2, 2), (2.0, 1, 1), (1.0, 3, 2)][case]
    _say("acwe case %d shape %s alpha %g smooth %d steps %d" % (case, shape, alpha, smooth, steps))
    got = active_contour(img, labels, alpha, smooth, steps)
    want = post.acwe(img, labels, alpha, smooth, steps)
    assert (want != labels).any()
    np.testing.assert_array_equal(got, want)


def test_smart_interpolation_acwe_key():
    from biomedisa_amd.interpolation import smart_interpolation
    from biomedisa_amd.synth import structures_volume
    vol = structures_volume(shape=(30, 36, 40), seed=14, n_struct=2, radius=(4, 9), every=4, allaxis=False)
    kw = dict(nbrw=3, sorw=120, acwe=True, acwe_steps=2)
    got = smart_interpolation(vol.img, vol.labels, **kw)
    want = opipe.smart_interpolation(vol.img, vol.labels, **kw)
    np.testing.assert_array_equal(got["regular"], want["regular"])
    np.testing.assert_array_equal(got["acwe"], want["acwe"])


@pytest.mark.parametrize("case", range(4))
def test_refine_distance_bitexact(eng, case):
    """The wavefront order of bmx_refine_distance reproduces the sequential
    chamfer scans bit for bit (float32 distances), rim voxels included."""
    from oracle import orc
    rng = np.random.default_rng(500 + case)
    shape = tuple(int(s) for s in rng.integers(3, 36, 3))
    lab = (rng.random(shape) < [0.02, 0.2, 0.001, 0.5][case]).astype(np.uint8) * 3
    _say("refine distance case %d shape %s" % (case, shape))
    lb = eng.upload(np.where(lab == 0, 0, lab).astype(np.uint8))
    dist, d = eng.alloc(4 * lab.size), eng.alloc(lab.size)
    _hip.call("bmx_refine_distance", eng.ctx, lb.ptr, *shape, dist.ptr, d.ptr)
    got = dist.download(shape, np.float32)
    want = np.zeros(shape, np.float32)
    want[lab == 0] = 100
    orc.geodis(want, np.sqrt(2), np.sqrt(3), 1)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    np.testing.assert_array_equal(d.download(shape, np.uint8), (want > 10).astype(np.uint8))


@pytest.mark.parametrize("case", range(3))
def test_refinement_matches_oracle(eng, case):
    from biomedisa_amd.post import refinement
    rng = np.random.default_rng(600 + case)
    shape = tuple(int(s) for s in rng.integers(20, 50, 3))
    labels = _blobs(rng, shape, int(rng.integers(1, 5)), float(rng.uniform(0.2, 0.6)))
    img = (labels.astype(np.float32) * 45 + rng.normal(0, 30, shape)).clip(0, 255).astype(np.uint8)
    _say("refinement case %d shape %s" % (case, shape))
    got = refinement(img, labels)
    want = post.refinement(img, labels)
    assert want.any()
    np.testing.assert_array_equal(got, want)


def test_smart_interpolation_refined_key():
    from biomedisa_amd.interpolation import smart_interpolation
    from biomedisa_amd.synth import structures_volume
    vol = structures_volume(shape=(30, 36, 40), seed=15, n_struct=3, radius=(4, 9), every=4, allaxis=False)
    kw = dict(nbrw=3, sorw=120, acwe=True)
    got = smart_interpolation(vol.img, vol.labels, **kw)
    want = opipe.smart_interpolation(vol.img, vol.labels, **kw)
    for k in ("acwe", "refined"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
