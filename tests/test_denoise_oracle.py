"""The oracle's smooth_img_3x3 (features/biomedisa_helper.py:309-323) against the
same mean computed another way: the 27-neighbour sum and the in-bounds count as
sums of shifted zero-padded arrays, divided in float64."""
import numpy as np
import pytest

from oracle import orc


def _mean3(img):
    f = img.astype(np.float64)
    Z, Y, X = img.shape
    p = np.pad(f, 1)
    c = np.pad(np.ones_like(f), 1)
    s = np.zeros_like(f)
    n = np.zeros_like(f)
    for k in range(3):
        for l in range(3):
            for m in range(3):
                s += p[k:k + Z, l:l + Y, m:m + X]
                n += c[k:k + Z, l:l + Y, m:m + X]
    return s / n


@pytest.mark.parametrize("shape", [(1, 1, 1), (1, 4, 5), (3, 1, 7), (9, 11, 13)])
def test_uint8_truncates_the_mean(shape):
    img = np.random.default_rng(sum(shape)).integers(0, 256, shape).astype(np.uint8)
    np.testing.assert_array_equal(orc.smooth_img(img), _mean3(img).astype(np.uint8))


def test_float32_integer_values_exact():
    img = np.random.default_rng(5).integers(0, 1000, (8, 9, 10)).astype(np.float32)
    np.testing.assert_array_equal(orc.smooth_img(img), _mean3(img).astype(np.float32))


def test_float32_random_close_and_corner_cases():
    img = (np.random.default_rng(6).random((7, 8, 9)) * 255).astype(np.float32)
    got = orc.smooth_img(img)
    np.testing.assert_allclose(got, _mean3(img).astype(np.float32), rtol=1e-6)
    # a corner voxel averages its 8 in-bounds neighbours, an interior one 27
    assert got[0, 0, 0] == np.float32(img[:2, :2, :2].astype(np.float64).sum() / 8)
    assert np.isclose(got[3, 4, 4], img[2:5, 3:6, 3:6].astype(np.float64).mean(), rtol=1e-6)


def test_constant_image_is_fixed():
    for img in (np.full((4, 5, 6), 77, np.uint8), np.full((4, 5, 6), 3.25, np.float32)):
        np.testing.assert_array_equal(orc.smooth_img(img), img)
