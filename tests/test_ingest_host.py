"""Host logic of the device pre-processing, on the CPU: the label set, box and
label-index table derived from label-class counts and boxes must equal what
pre_processing, predict_blocksize and label_index_volume give on the array
itself (biomedisa_helper.py:639-724, 1105-1124); the Dice rule of the sanity
check (rw_small.py:203-211) from counts; the out-of-memory stages."""
import numpy as np
import pytest

from biomedisa_amd import _hip
from biomedisa_amd import helpers as H
from biomedisa_amd.interpolation import (SMOOTH_OOM_WARNING, bad_result_warning, optional_stage)


def class_stats(lab):
    """Counts and boxes per label class, computed on the array (what
    bmx_label_stats returns): values 1..255, 256 negative, 257 above 255."""
    cls = np.where(lab < 0, 256, np.where(lab > 255, 257, lab)).astype(np.int64)
    counts = np.bincount(cls.ravel(), minlength=258).astype(np.uint64)
    counts[0] = 0
    boxes = np.tile(np.array([np.iinfo(np.int32).max, -1] * 3, np.int32), (258, 1))
    for c in np.flatnonzero(counts):
        z, y, x = np.nonzero(cls == c)
        boxes[c] = [z.min(), z.max(), y.min(), y.max(), x.min(), x.max()]
    return counts, boxes


def make_bm(lab, only="all", ignore="none"):
    bm = H.Biomedisa()
    bm.data = np.zeros(lab.shape, np.uint8)
    bm.labelData = lab
    bm.only, bm.ignore, bm.success = only, ignore, True
    return bm


def random_labels(seed, shape=(18, 21, 24)):
    rng = np.random.default_rng(seed)
    lab = np.zeros(shape, np.int32)
    for z in rng.choice(shape[0], size=4, replace=False):
        vals = rng.choice([0, 1, 2, 3, 7, 200, 255, 300, -1, -5], size=shape[1:],
                          p=[.4, .15, .1, .1, .05, .05, .05, .04, .03, .03])
        lab[z] = vals
    return lab


OPTS = [dict(), dict(only="2,7"), dict(only="3"), dict(ignore="1"), dict(ignore="2,255"), dict(ignore="0"),
        dict(ignore="300")]


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("opt", OPTS)
def test_counts_give_pre_processing(seed, opt, capsys):
    lab = random_labels(seed)
    counts, boxes = class_stats(lab)
    a = H.pre_processing(make_bm(lab.copy(), **opt))
    out_a = capsys.readouterr().out
    b = H.pre_processing_counts(make_bm(lab, **opt), counts)
    out_b = capsys.readouterr().out
    assert a.success == b.success
    assert out_a == out_b
    if not a.success:
        return
    np.testing.assert_array_equal(a.allLabels, b.allLabels)
    assert a.nol == b.nol
    # box of the edited, padded labels
    box = H.predict_blocksize(np.pad(a.labelData, 1))
    assert H.box_from_class_boxes(boxes, b.kept, b.neg_kept, lab.shape) == box
    # label index of the edited labels = class table of the raw labels
    want = H.label_index_volume(a.labelData, a.allLabels)
    cls = np.where(lab < 0, 256, np.where(lab > 255, 257, lab))
    np.testing.assert_array_equal(H.class_lut(b.allLabels, b.neg_kept)[cls], want)


def test_negative_ignore_needs_the_array():
    lab = random_labels(3)
    counts, _ = class_stats(lab)
    assert counts[256] > 0
    assert H.pre_processing_counts(make_bm(lab, ignore="-1"), counts) is None


def test_no_labels_is_an_error(capsys):
    lab = np.zeros((5, 6, 7), np.int32)
    lab[2, 1, 1] = -3
    bm = H.pre_processing_counts(make_bm(lab), class_stats(lab)[0])
    assert not bm.success
    assert "No labeled slices found" in capsys.readouterr().out


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.int32])
def test_dice_rule_from_counts(dtype):
    """Counts of matches / labelled / labelled with a result, with the
    uint8 wrap of (labels + result) for label 128, give Dice_score of the
    labels against result * (labels > 0)."""
    rng = np.random.default_rng(4)
    vals = np.array([0, 1, 2, 128, 200], dtype)
    gt = rng.choice(vals, size=(6, 7, 8))
    res = np.where(rng.random(gt.shape) < 0.7, gt, rng.choice(vals, size=gt.shape)).astype(np.uint8)
    mask = gt > 0
    want = H.Dice_score(gt, res * mask)
    counted = H.dice_counted(dtype)
    g = gt.astype(np.int64)
    a = int(np.sum(mask & (res == g) & (counted[np.clip(g, 0, 255)] != 0)))
    b = int(mask.sum())
    c = int(np.sum(mask & (res > 0)))
    assert H.dice_from_counts(a, b, c) == pytest.approx(want, rel=0, abs=0)
    if dtype == np.uint8:
        assert counted[128] == 0
    assert np.isnan(H.dice_from_counts(0, 0, 0))


def test_bad_result_warning(capsys):
    assert bad_result_warning(0.29)
    assert "Bad result" in capsys.readouterr().out
    assert not bad_result_warning(0.31)
    assert not bad_result_warning(float("nan"))
    assert capsys.readouterr().out == ""


def test_optional_stage_degrades_on_oom(capsys):
    def oom():
        raise _hip.BmxError(_hip.BMX_E_OOM, "out of memory")

    assert optional_stage(oom, SMOOTH_OOM_WARNING) is None
    assert capsys.readouterr().out.strip() == SMOOTH_OOM_WARNING

    def other():
        raise _hip.BmxError(_hip.BMX_E_DEVICE, "device")

    with pytest.raises(_hip.BmxError):
        optional_stage(other, SMOOTH_OOM_WARNING)
    assert optional_stage(lambda: 5, SMOOTH_OOM_WARNING) == 5
