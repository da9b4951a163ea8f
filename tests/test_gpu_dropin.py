"""smart_interpolation at the edges of its input contract, against the oracle
pipeline: the label options only/ignore, label dtypes, negative and >255 label
values, image dtypes, and the inputs the reference rejects with an error
(biomedisa_helper.py:639-724, interpolation.py:135-183)."""
import numpy as np
import pytest

import oracle.pipeline as opipe
from biomedisa_amd.interpolation import smart_interpolation
from biomedisa_amd.synth import structures_volume

pytestmark = pytest.mark.gpu

KW = dict(nbrw=3, sorw=120)


@pytest.fixture(scope="module")
def vol():
    return structures_volume(shape=(30, 34, 38), seed=51, n_struct=4, radius=(4, 9), every=5, allaxis=False)


def _eq(got, want):
    assert got is not None and want is not None
    np.testing.assert_array_equal(got["regular"], want["regular"])


@pytest.mark.parametrize("opt", [dict(only="1,3"), dict(ignore="2"), dict(ignore="2,4"), dict(only="2")])
def test_only_and_ignore(vol, opt):
    """Labels left out by only/ignore become 0 (background seeds), as
    pre_processing does before the walk."""
    lab = vol.labels.copy()
    present = [int(v) for v in np.unique(lab) if v > 0]
    if "only" in opt:
        keep = {int(k) for k in opt["only"].split(",")}
        drop = [v for v in present if v not in keep]
    else:
        drop = [int(k) for k in opt["ignore"].split(",")]
    ref_lab = lab.copy()
    for v in drop:
        ref_lab[ref_lab == v] = 0
    _eq(smart_interpolation(vol.img, lab, **KW, **opt), opipe.smart_interpolation(vol.img, ref_lab, **KW))


@pytest.mark.parametrize("dtype", [np.int32, np.uint16, np.int64, np.uint32])
def test_label_dtypes(vol, dtype):
    want = opipe.smart_interpolation(vol.img, vol.labels, **KW)
    _eq(smart_interpolation(vol.img, vol.labels.astype(dtype), **KW), want)


def test_negative_and_large_label_values(vol):
    """Negative labels mark ignored regions; values above 255 are removed."""
    lab = vol.labels.astype(np.int32)
    planes = np.flatnonzero(lab.reshape(lab.shape[0], -1).any(1))
    lab[planes[0], :3, :3] = -1
    lab[planes[1], -4:, -4:] = 300
    ref = lab.copy()
    ref[ref > 255] = 0
    _eq(smart_interpolation(vol.img, lab, **KW), opipe.smart_interpolation(vol.img, ref, **KW))


@pytest.mark.parametrize("dtype", [np.int8, np.uint16, np.float64, np.float32])
def test_image_dtypes(vol, dtype):
    img = vol.img.astype(np.int16)
    if dtype == np.int8:
        img = (img - 128).astype(np.int8)
    else:
        img = (img * 3 + 7).astype(dtype)
    _eq(smart_interpolation(img, vol.labels, **KW), opipe.smart_interpolation(img, vol.labels, **KW))


def test_rejected_inputs(vol, capsys):
    assert smart_interpolation(vol.img, np.zeros_like(vol.labels), **KW) is None
    assert "Error" in capsys.readouterr().out
    assert smart_interpolation(vol.img, vol.labels[:-1], **KW) is None
    assert "Error" in capsys.readouterr().out
    # allaxis with labels on two adjacent planes
    lab = np.zeros_like(vol.labels)
    lab[10] = vol.labels[10]
    lab[11] = vol.labels[10]
    lab[10, 0, 0] = lab[11, 0, 0] = 1
    assert smart_interpolation(vol.img, lab, allaxis=True, **KW) is None
    assert "Error" in capsys.readouterr().out
