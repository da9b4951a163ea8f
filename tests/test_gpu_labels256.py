"""Volumes that use all 256 label values (the reference's switch to short
labels with -1 as ignore value, pyopencl_small.py:90-95): the whole drop-in
call against the oracle, bit for bit, on the small- and the large-volume
driver, and the one combination a uint8 label index cannot hold (256 values
and a negative label) rejected with the reference's error convention."""
import numpy as np
import pytest

import oracle.pipeline as opipe
from biomedisa_amd import interpolation as I

pytestmark = pytest.mark.gpu


def volume_256(seed=0, shape=(30, 50, 50)):
    """Labels on z-planes 7, 14 and 21: each plane a 16 x 16 grid of 3 x 3
    patches holding the 256 values in a random order; the image is a smooth
    function of the value of the nearest patch plus noise."""
    rng = np.random.default_rng(seed)
    Z, Y, X = shape
    lab = np.zeros(shape, np.uint8)
    base = np.zeros((Y, X), np.int32)
    perm = rng.permutation(256).reshape(16, 16)
    for j in range(16):
        for i in range(16):
            base[1 + 3 * j:4 + 3 * j, 1 + 3 * i:4 + 3 * i] = perm[j, i]
    for z in (7, 14, 21):
        lab[z] = base
    img = (base[None, :, :] * 0.9 + rng.normal(0, 4, shape)).clip(0, 255).astype(np.uint8)
    return img, lab


@pytest.mark.parametrize("path", ["small", "large"])
def test_256_labels_match_oracle(path):
    img, lab = volume_256()
    assert np.unique(lab).size == 256
    kw = dict(nbrw=2, sorw=100, smooth=3, uncertainty=True)
    print("256 labels, path %s, %s" % (path, img.shape), flush=True)
    got = I.smart_interpolation(img, lab, path=path, **kw)
    want = opipe.smart_interpolation(img, lab, path=path, **kw)
    assert got is not None and want is not None
    for key in ("regular", "smooth", "uncertainty"):
        np.testing.assert_array_equal(got[key], want[key], err_msg=key)
    # label 255 is a label like the others: its patches keep their value
    assert np.count_nonzero(got["regular"][14] == 255) >= 9


def test_256_labels_with_negative_rejected(capsys):
    img, lab = volume_256(seed=1)
    lab = lab.astype(np.int16)
    lab[14, 40:45, 40:45] = -1
    capsys.readouterr()
    got = I.smart_interpolation(img, lab, nbrw=2, sorw=50)
    assert got is None
    assert I.NEG_WITH_256_LABELS in capsys.readouterr().out


def test_guided_method_with_256_labels_is_plain():
    """More labels than the proposal network predicts: method='guided' returns
    the plain result, bit for bit."""
    img, lab = volume_256(seed=2)
    kw = dict(nbrw=2, sorw=100, smooth=3)
    print("256 labels, guided method", flush=True)
    plain = I.smart_interpolation(img, lab, **kw)
    guided = I.smart_interpolation(img, lab, method="guided", **kw)
    assert plain is not None and guided is not None
    for key in ("regular", "smooth"):
        np.testing.assert_array_equal(guided[key], plain[key], err_msg=key)
