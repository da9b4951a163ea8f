"""Host-side helpers of the drop-in against the reference's documented behaviour."""
import numpy as np
import pytest

from biomedisa_amd import helpers as H
import oracle.pipeline as opipe


def _bm(data, labels, only="all", ignore="none"):
    bm = H.Biomedisa()
    bm.data, bm.labelData, bm.only, bm.ignore = data, labels, only, ignore
    bm.success = True
    return bm


def test_pre_processing_label_set():
    img = np.zeros((4, 5, 6), np.uint8)
    lab = np.zeros((4, 5, 6), np.int32)
    lab[1, 1, 1], lab[1, 2, 2], lab[2, 3, 3], lab[2, 0, 0] = 7, 3, -2, 300
    bm = H.pre_processing(_bm(img, lab))
    assert bm.success and list(bm.allLabels) == [0, 3, 7] and bm.nol == 3
    assert bm.labelData[2, 0, 0] == 0 and bm.labelData[2, 3, 3] == -2


def test_pre_processing_only_and_ignore():
    img = np.zeros((3, 3, 3), np.uint8)
    lab = np.zeros((3, 3, 3), np.uint8)
    lab[0, 0, :] = [1, 2, 3]
    bm = H.pre_processing(_bm(img, lab.copy(), only="1,3"))
    assert list(bm.allLabels) == [0, 1, 3] and bm.labelData[0, 0, 1] == 0
    bm = H.pre_processing(_bm(img, lab.copy(), ignore="2"))
    assert list(bm.allLabels) == [0, 1, 3]


def test_pre_processing_errors(capsys):
    img = np.zeros((3, 3, 3), np.uint8)
    assert not H.pre_processing(_bm(img, np.zeros((3, 3, 3), np.uint8))).success
    assert "No labeled slices found." in capsys.readouterr().out
    assert not H.pre_processing(_bm(img, np.zeros((3, 3), np.uint8))).success
    assert not H.pre_processing(_bm(img, np.zeros((3, 3, 4), np.uint8))).success


@pytest.mark.parametrize("trial", range(5))
def test_predict_blocksize_matches_slice_loop(trial):
    rng = np.random.default_rng(trial)
    shape = tuple(int(v) for v in rng.integers(20, 260, 3))
    lab = np.zeros(shape, np.int16)
    for _ in range(3):
        p = tuple(int(rng.integers(1, s - 1)) for s in shape)
        lab[p] = rng.integers(1, 5)
    assert H.predict_blocksize(lab) == opipe._bbox(lab)


def test_label_index_volume():
    lab = np.array([[[0, 5, 9, -1, 5]]], np.int32)
    idx = H.label_index_volume(lab, np.array([0, 5, 9]))
    assert idx.tolist() == [[[0, 1, 2, 255, 1]]]
    lab8 = np.array([[[0, 5, 9]]], np.uint8)
    assert H.label_index_volume(lab8, np.array([0, 5, 9])).tolist() == [[[0, 1, 2]]]


def test_get_labels_and_dice():
    pre = np.array([0, 1, 2, 2], np.uint8)
    assert H.get_labels(pre, np.array([0, 4, 9])).tolist() == [0, 4, 9, 9]
    a = np.array([0, 1, 1, 2, 2])
    b = np.array([0, 1, 2, 2, 0])
    assert H.Dice_score(a, b) == pytest.approx(2 * 2 / (4 + 3))
    assert H.Dice_score(a, b, average_dice=True) == pytest.approx((2 / 3 + 1 / 2) / 2)
    assert H.per_label_dice(a, b) == {1: pytest.approx(2 / 3), 2: pytest.approx(1 / 2)}


def test_adjacent_planes():
    assert H.adjacent_planes([np.array([0, 1, 1, 0]), np.zeros(3), np.zeros(3)])
    assert not H.adjacent_planes([np.array([1, 0, 1, 0]), np.array([0, 1]), np.array([1])])


def test_normalize_image():
    i8 = np.array([-128, 0, 127], np.int8)
    assert H.normalize_image(i8).tolist() == [0, 128, 255]
    u16 = np.array([10, 20, 30], np.uint16)
    f = H.normalize_image(u16)
    assert f.dtype == np.float32 and f.tolist() == [0.0, 127.5, 255.0]


def test_batch_volumes_workload_shape():
    """The batched workload: seeded, shapes within [96, (256, 320, 320)], 2-12
    structures, XY-plane labels only."""
    from biomedisa_amd.synth import batch_volumes
    a = batch_volumes(seed=3, n=3)
    b = batch_volumes(seed=3, n=3)
    for va, vb in zip(a, b):
        np.testing.assert_array_equal(va.img, vb.img)
        assert va.img.dtype == np.uint8 and not va.allaxis
        assert all(96 <= s <= hi for s, hi in zip(va.img.shape, (256, 320, 320)))
        assert 2 <= int(va.gt.max()) <= 12
        zs = np.nonzero(va.labels.reshape(va.labels.shape[0], -1).any(axis=1))[0]
        assert len(zs) >= 2 and any(np.all(zs % k == 0) for k in range(6, 13))


@pytest.mark.parametrize("kw", [dict(far_walks=3), dict(mid_walks=3), dict(mid_walks=0), dict(nbrw=5, mid_walks=2)])
def test_guided_parameter_checks(kw):
    """run_pipeline_guided rejects walk counts that do not divide nbrw before it
    touches the device (the volume here has no engine at all)."""
    from biomedisa_amd.interpolation import PipelineError, run_pipeline_guided

    class NoDevice:
        eng, dims, V = None, (4, 4, 4), 64

    args = dict(nbrw=10, sorw=10)
    args.update(kw)
    with pytest.raises(PipelineError, match="must divide nbrw"):
        run_pipeline_guided(NoDevice(), 2, np.arange(2, dtype=np.uint8), None, **args)


def test_guided_label_limit_falls_back_to_plain(monkeypatch):
    """More labels than the network's head predicts: the guided method runs the
    plain pipeline with the same walk arguments and marks the result as a
    fallback, before any device work of its own."""
    from biomedisa_amd import interpolation as I
    from biomedisa_amd import surrogate as S

    class NoDevice:
        eng, dims, V = None, (4, 4, 4), 64

    calls = []

    def plain(vol, nlabels, values, nbrw, sorw, allaxis, smooth, **kw):
        calls.append((nlabels, nbrw, sorw, allaxis, smooth, kw["key"]))
        return {"nseeds": 7, "walked_seeds": 7}

    monkeypatch.setattr(I, "run_pipeline", plain)
    out = I.run_pipeline_guided(NoDevice(), S.L_MAX + 1, np.arange(S.L_MAX + 1, dtype=np.uint8), None,
                                nbrw=10, sorw=50, allaxis=True, smooth=3, key=9)
    assert calls == [(S.L_MAX + 1, 10, 50, True, 3, 9)]
    assert out["fallback"] is True and out["walked_seeds"] == 7 and out["ring_seeds"] == 0


def test_guided_defaults_for_xy_labels():
    """XY-only labels get the stricter guided settings, with ring walk counts
    that divide nbrw; all-axis labels keep the pipeline's own defaults."""
    from biomedisa_amd.interpolation import guided_defaults
    assert guided_defaults(10, True) == {}
    assert guided_defaults(10, False) == {"thr": 0.95, "brad": 6, "far_walks": 2, "mid_walks": 5}
    for nbrw in range(1, 40):
        g = guided_defaults(nbrw, False)
        assert nbrw % g["far_walks"] == 0 and nbrw % g["mid_walks"] == 0
        assert 1 <= g["far_walks"] <= g["mid_walks"] <= max(nbrw // 2, 1)
