"""Parity of the HIP path (libbmx_hip through its C ABI) against the CPU oracle.

Integer and byte results (plane flags, seeds, hit counts, labels) must be
bit-exact; so must the float32 smoothing (both sides evaluate the same IEEE
operations in the same order, no contraction).
"""
import numpy as np
import pytest

import oracle
from oracle import orc
import oracle.pipeline as opipe
from biomedisa_amd import _hip
from biomedisa_amd.synth import ellipsoids_volume, structures_volume

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _random_labels(shape, rng, every, allaxis, nlab=3, neg_frac=0.01):
    gt = rng.integers(0, nlab + 1, size=shape).astype(np.uint8)
    lab = np.zeros_like(gt)
    lab[1::every] = gt[1::every]
    if allaxis:
        lab[:, 2::every] = gt[:, 2::every]
        lab[:, :, 3::every] = gt[:, :, 3::every]
    lab[(rng.random(shape) < neg_frac) & (lab > 0)] = oracle.IGN
    return lab


def _gpu_seeds(eng, lab, img, allaxis):
    dims = lab.shape
    lb = eng.upload(lab)
    ib = eng.upload(img)
    dt = _hip.IMG_U8 if img.dtype == np.uint8 else _hip.IMG_F32
    fb = eng.plane_flags(lb, dims, allaxis)
    flags = fb.download((sum(dims),), np.uint8)
    sb, n = eng.seeds(lb, ib, dt, dims, fb)
    seeds = sb.download((n,), _hip.SEED_DTYPE) if n else np.zeros(0, _hip.SEED_DTYPE)
    return flags, seeds, (lb, ib, dt, sb, n)


@pytest.mark.parametrize("allaxis", [False, True])
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_plane_flags_and_seeds_bitexact(eng, allaxis, dtype):
    rng = np.random.default_rng(21 + allaxis)
    shape = (17, 23, 29)
    _say("flags/seeds shape=%s allaxis=%s dtype=%s" % (shape, allaxis, np.dtype(dtype).name))
    lab = _random_labels(shape, rng, 4, allaxis)
    img = (rng.integers(0, 256, shape).astype(np.uint8) if dtype == np.uint8
           else (rng.random(shape) * 255).astype(np.float32))
    flags, seeds, _ = _gpu_seeds(eng, lab, img, allaxis)
    ofl = orc.plane_flags(lab, allaxis)
    np.testing.assert_array_equal(flags, np.concatenate(ofl))
    pos, l, var = orc.seeds(lab, img, ofl)
    np.testing.assert_array_equal(seeds["pos"], pos)
    np.testing.assert_array_equal(seeds["label"], l)
    np.testing.assert_array_equal(seeds["var"].view(np.uint32), var.view(np.uint32))


@pytest.mark.parametrize("case", [
    dict(shape=(20, 22, 24), dtype=np.uint8, nbrw=3, sorw=203, rng="philox", allaxis=False),
    dict(shape=(19, 21, 18), dtype=np.float32, nbrw=2, sorw=150, rng="philox", allaxis=True),
    dict(shape=(22, 19, 26), dtype=np.uint8, nbrw=10, sorw=101, rng="philox", allaxis=True),
    dict(shape=(18, 20, 17), dtype=np.uint8, nbrw=8, sorw=64, rng="philox", allaxis=False),
    dict(shape=(17, 18, 19), dtype=np.float32, nbrw=10, sorw=55, rng="philox", allaxis=False),
    dict(shape=(16, 18, 20), dtype=np.uint8, nbrw=2, sorw=97, rng="reference", allaxis=False),
    dict(shape=(15, 16, 17), dtype=np.float32, nbrw=2, sorw=64, rng="reference", allaxis=True),
])
def test_walk_hits_bitexact(eng, case):
    rng = np.random.default_rng(31)
    shape = case["shape"]
    _say("walk %s" % case)
    lab = _random_labels(shape, rng, 5, case["allaxis"], nlab=4)
    base = rng.integers(40, 200, shape).astype(np.float32)
    img = base.astype(np.uint8) if case["dtype"] == np.uint8 else (base + rng.random(shape)).astype(np.float32)
    _, seeds, (lb, ib, dt, sb, n) = _gpu_seeds(eng, lab, img, case["allaxis"])
    nl = 5
    hits = eng.alloc(nl * lab.size * 4)
    hits.zero()
    key = 0x1234_5678_9ABC_DEF0
    rngc = _hip.RNG_PHILOX if case["rng"] == "philox" else _hip.RNG_REF
    acc = eng.walk(ib, dt, shape, sb, n, nl, case["nbrw"], case["sorw"], rngc, key, hits)
    h = hits.download((nl,) + shape, np.uint32)
    pos, l, var = seeds["pos"].astype(np.int64), seeds["label"].astype(np.uint8), seeds["var"]
    if case["rng"] == "philox":
        oh, oacc = orc.philox_walk(img, pos, l, var, nl, case["nbrw"], case["sorw"], key)
    else:
        oh, oacc = orc.ref_walk(img, pos, l, var, nl, case["nbrw"], case["sorw"])
    assert acc == oacc
    assert int(h.sum(dtype=np.int64)) == acc
    np.testing.assert_array_equal(h, oh)


@pytest.mark.parametrize("seg,blog", [(8, 1), (12, 2), (128, 3)])
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_sorted_walker_bitexact(eng, seg, blog, dtype):
    """Lockstep segments + brick sort give the oracle's hits exactly (sorw not a
    multiple of the segment, so the last segment is partial)."""
    rng = np.random.default_rng(41)
    shape = (23, 27, 25)
    _say("sorted walker seg=%d brick_log2=%d dtype=%s" % (seg, blog, np.dtype(dtype).name))
    lab = _random_labels(shape, rng, 4, True, nlab=3)
    base = rng.integers(30, 220, shape).astype(np.float32)
    img = base.astype(np.uint8) if dtype == np.uint8 else (base + rng.random(shape)).astype(np.float32)
    _, seeds, (lb, ib, dt, sb, n) = _gpu_seeds(eng, lab, img, True)
    nl, nbrw, sorw, key = 4, 3, 101, 77
    hits = eng.alloc(nl * lab.size * 4)
    hits.zero()
    acc = eng.walk(ib, dt, shape, sb, n, nl, nbrw, sorw, _hip.RNG_PHILOX, key, hits,
                   algorithm=_hip.WALK_SORTED, segment_steps=seg, brick_log2=blog)
    h = hits.download((nl,) + shape, np.uint32)
    oh, oacc = orc.philox_walk(img, seeds["pos"].astype(np.int64), seeds["label"].astype(np.uint8),
                               seeds["var"], nl, nbrw, sorw, key)
    assert acc == oacc
    np.testing.assert_array_equal(h, oh)
    phase, kern, launches = eng.last_walk_stats()
    assert launches == (sorw + seg - 1) // seg and 0 < kern <= phase


@pytest.mark.parametrize("cap", [8, 1024])
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
@pytest.mark.parametrize("shape", [(23, 27, 25), (40, 37, 45)])
def test_tiled_walker_bitexact(eng, cap, dtype, shape):
    """Tile-owner walker (LDS-privatised hits, rounds of bucket sorting) gives the
    oracle's hits exactly: several tiles, partial edge tiles, walkers leaving and
    re-entering regions, a round cap that forces many rounds, sorw % 4 != 0."""
    rng = np.random.default_rng(43)
    _say("tiled walker cap=%d dtype=%s shape=%s" % (cap, np.dtype(dtype).name, shape))
    lab = _random_labels(shape, rng, 4, True, nlab=3)
    base = rng.integers(30, 220, shape).astype(np.float32)
    img = base.astype(np.uint8) if dtype == np.uint8 else (base + rng.random(shape)).astype(np.float32)
    _, seeds, (lb, ib, dt, sb, n) = _gpu_seeds(eng, lab, img, True)
    nl, nbrw, sorw, key = 4, 3, 203, 78
    hits = eng.alloc(nl * lab.size * 4)
    hits.zero()
    acc = eng.walk(ib, dt, shape, sb, n, nl, nbrw, sorw, _hip.RNG_PHILOX, key, hits,
                   algorithm=_hip.WALK_TILED, segment_steps=cap)
    h = hits.download((nl,) + shape, np.uint32)
    oh, oacc = orc.philox_walk(img, seeds["pos"].astype(np.int64), seeds["label"].astype(np.uint8),
                               seeds["var"], nl, nbrw, sorw, key)
    assert acc == oacc
    np.testing.assert_array_equal(h, oh)
    phase, kern, launches = eng.last_walk_stats()
    assert launches >= 1 and 0 < kern <= phase


@pytest.mark.parametrize("graph_rounds", [-1, 0])
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_tiled_walker_crowded_tiles_bitexact(eng, graph_rounds, dtype):
    """Every walker starts on one labelled plane, so a few tiles hold them all:
    the rounds list such a (tile, label) key as several buckets of at most
    1024 walkers, walked by different workgroups into the same hits.  Hits
    equal the oracle's with both round drivers."""
    rng = np.random.default_rng(47)
    shape = (40, 36, 38)
    _say("tiled walker, crowded tiles, graph_rounds=%d dtype=%s" % (graph_rounds, np.dtype(dtype).name))
    lab = np.zeros(shape, np.uint8)
    lab[20, 6:30, 6:32] = rng.integers(1, 3, (24, 26))
    base = rng.integers(30, 220, shape).astype(np.float32)
    img = base.astype(np.uint8) if dtype == np.uint8 else (base + rng.random(shape)).astype(np.float32)
    _, seeds, (lb, ib, dt, sb, n) = _gpu_seeds(eng, lab, img, False)
    nl, nbrw, sorw, key = 3, 40, 203, 79
    # the plane's walkers over its 3 x 3 tiles and 3 labels: keys of more than 1024 walkers
    assert n * nbrw > 27 * 1024
    hits = eng.alloc(nl * lab.size * 4)
    hits.zero()
    acc = eng.walk(ib, dt, shape, sb, n, nl, nbrw, sorw, _hip.RNG_PHILOX, key, hits,
                   algorithm=_hip.WALK_TILED, graph_rounds=graph_rounds)
    h = hits.download((nl,) + shape, np.uint32)
    oh, oacc = orc.philox_walk(img, seeds["pos"].astype(np.int64), seeds["label"].astype(np.uint8),
                               seeds["var"], nl, nbrw, sorw, key)
    assert acc == oacc
    np.testing.assert_array_equal(h, oh)


def test_walk_accumulates_and_handles_empty(eng):
    shape = (8, 9, 10)
    img = eng.upload(np.full(shape, 9, np.uint8))
    hits = eng.alloc(2 * 720 * 4)
    hits.zero()
    assert eng.walk(img, _hip.IMG_U8, shape, None, 0, 2, 10, 100, _hip.RNG_PHILOX, 1, hits) == 0
    assert hits.download((2 * 720,), np.uint32).sum() == 0


def test_abi_rejects_bad_arguments(eng):
    shape = (8, 9, 10)
    img = eng.upload(np.zeros(shape, np.uint8))
    hits = eng.alloc(720 * 4)
    with pytest.raises(_hip.BmxError, match="nlabels"):
        eng.walk(img, _hip.IMG_U8, shape, None, 0, 0, 1, 1, _hip.RNG_PHILOX, 1, hits)
    with pytest.raises(_hip.BmxError, match="dims"):
        eng.walk(img, _hip.IMG_U8, (2, 9, 10), None, 0, 1, 1, 1, _hip.RNG_PHILOX, 1, hits)
    with pytest.raises(_hip.BmxError, match="rng"):
        eng.walk(img, _hip.IMG_U8, shape, None, 0, 1, 1, 1, 7, 1, hits)


@pytest.mark.parametrize("V", [4 * 1000, 4 * 1000 + 3])
def test_argmax_bitexact(eng, V):
    rng = np.random.default_rng(V)
    nl = 6
    h = rng.integers(0, 5, (nl, V)).astype(np.uint32)
    h[:, :7] = 0
    values = np.array([0, 2, 4, 8, 16, 250], np.uint8)
    hb = eng.upload(h)
    out = eng.alloc(V)
    eng.argmax(hb, nl, V, values, out)
    got = out.download((V,), np.uint8)
    np.testing.assert_array_equal(got, orc.argmax(h.reshape(nl, 1, 1, V), values).ravel())


@pytest.mark.parametrize("iters", [1, 2, 5])
def test_smooth_bitexact(eng, iters):
    rng = np.random.default_rng(iters)
    shape = (37, 11, 70)          # x spans two 64-wide tiles, z spans one chunk boundary? no: < 64
    phi = (rng.random(shape) * 50).astype(np.float32)
    a = eng.upload(phi)
    b = eng.alloc(phi.nbytes)
    eng.smooth(a, b, shape, iters)
    got = a.download(shape, np.float32)
    ref = orc.smooth(phi, iters)
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_smooth_bitexact_tall(eng):
    shape = (150, 9, 7)           # z spans three 64-plane chunks
    phi = (np.random.default_rng(2).random(shape) * 9).astype(np.float32)
    a = eng.upload(phi)
    b = eng.alloc(phi.nbytes)
    eng.smooth(a, b, shape, 3)
    np.testing.assert_array_equal(a.download(shape, np.float32).view(np.uint32),
                                  orc.smooth(phi, 3).view(np.uint32))


@pytest.mark.parametrize("iters", [3, 6, 7])
def test_smooth_argmax_small_boxes_bitexact(eng, iters):
    """Labels whose hits sit in small clusters: each is smoothed in its own box
    (hits box grown by iters, clipped to the volume), two rounds per pass and a
    single pass for an odd count, against the oracle's whole-volume rounds."""
    rng = np.random.default_rng(30 + iters)
    shape = (40, 45, 140)
    nl = 4
    h = np.zeros((nl,) + shape, np.uint32)
    for l, (z, y, x) in enumerate([(20, 22, 70), (1, 2, 3), (38, 40, 135), (10, 30, 100)]):
        sl = tuple(slice(max(0, c - 4), c + 5) for c in (z, y, x))
        h[l][sl] = rng.integers(0, 30, h[l][sl].shape)
    values = np.array([0, 2, 4, 9], np.uint8)
    hb = eng.upload(h)
    out = eng.alloc(h[0].size)
    eng.smooth_argmax(hb, nl, shape, iters, values, out)
    phi = np.stack([orc.smooth(x.astype(np.float32), iters) for x in h])
    np.testing.assert_array_equal(out.download(shape, np.uint8), orc.argmax(phi, values))


@pytest.mark.parametrize("iters", [20, 21])
def test_smooth_argmax_skipped_blocks_bitexact(eng, iters):
    """Boxes spanning many pair-pass blocks (64 x 16 x 32 voxels), most of them
    still +0 for the first passes: those blocks are skipped, and write zeros
    where the reused buffers hold the dense first label's values.  Labels
    against the oracle's whole-volume rounds, even and odd round counts."""
    rng = np.random.default_rng(60 + iters)
    shape = (100, 70, 300)
    nl = 4
    h = np.zeros((nl,) + shape, np.uint32)
    h[0] = rng.integers(0, 25, shape)
    for l, (z, y, x) in enumerate([(50, 35, 150), (5, 60, 20), (90, 10, 280)], start=1):
        sl = tuple(slice(max(0, c - 3), c + 4) for c in (z, y, x))
        h[l][sl] = rng.integers(20, 400, h[l][sl].shape)
    values = np.array([0, 3, 5, 7], np.uint8)
    hb = eng.upload(h)
    out = eng.alloc(h[0].size)
    eng.smooth_argmax(hb, nl, shape, iters, values, out)
    phi = np.stack([orc.smooth(x.astype(np.float32), iters) for x in h])
    np.testing.assert_array_equal(out.download(shape, np.uint8), orc.argmax(phi, values))


@pytest.mark.parametrize("clamp", [False, True])
def test_smooth_argmax_label_sets_bitexact(eng, clamp):
    """Seven labels smoothed side by side: the largest box belongs to label 3
    (its update runs last), labels 1 and 5 hold the same hits (the tie goes to
    label 1), label 4 has no hits, and the other labels take turns on the
    reused buffer sets.  With clamp, label 0's negative values become 0 before
    the maximum, as the large-volume driver does."""
    rng = np.random.default_rng(71)
    shape = (30, 40, 90)
    nl = 7
    h = np.zeros((nl,) + shape, np.uint32)
    h[3] = rng.integers(0, 9, shape)
    for l, (z, y, x) in [(0, (15, 20, 45)), (1, (3, 5, 10)), (2, (27, 35, 80)), (6, (10, 30, 60))]:
        sl = tuple(slice(max(0, c - 3), c + 4) for c in (z, y, x))
        h[l][sl] = rng.integers(5, 300, h[l][sl].shape)
    h[5] = h[1]
    values = np.array([1, 2, 3, 4, 5, 6, 7], np.uint8)
    hb = eng.upload(h)
    out = eng.alloc(h[0].size)
    for iters in (5, 4):
        eng.smooth_argmax(hb, nl, shape, iters, values, out, clamp_first=clamp)
        phi = np.stack([orc.smooth(x.astype(np.float32), iters) for x in h])
        if clamp:
            phi[0][phi[0] < 0] = 0
        np.testing.assert_array_equal(out.download(shape, np.uint8), orc.argmax(phi, values))


@pytest.mark.parametrize("iters", [1, 2, 7])
def test_smooth_argmax_label_batches_bitexact(eng, iters):
    """Twelve labels: the eleven beside the largest box (label 6) share the two
    side streams, each stream reusing its buffers label after label, with an
    empty label (2) and boxes of different sizes; even, odd and single-round
    counts."""
    rng = np.random.default_rng(90 + iters)
    shape = (26, 37, 150)
    nl = 12
    h = np.zeros((nl,) + shape, np.uint32)
    h[6] = rng.integers(0, 5, shape)
    centres = rng.integers((0, 0, 0), shape, (nl, 3))
    for l in range(nl):
        if l in (6, 2):
            continue
        r = 2 + l % 4
        sl = tuple(slice(max(0, c - r), c + r + 1) for c in centres[l])
        h[l][sl] = rng.integers(1, 200, h[l][sl].shape)
    values = np.arange(10, 10 + nl, dtype=np.uint8)
    hb = eng.upload(h)
    out = eng.alloc(h[0].size)
    eng.smooth_argmax(hb, nl, shape, iters, values, out)
    phi = np.stack([orc.smooth(x.astype(np.float32), iters) for x in h])
    np.testing.assert_array_equal(out.download(shape, np.uint8), orc.argmax(phi, values))


@pytest.mark.parametrize("iters", [3, 4])
def test_smooth_argmax_single_label_bitexact(eng, iters):
    """One label runs on the context stream alone, into its own argmax."""
    rng = np.random.default_rng(120 + iters)
    shape = (20, 33, 70)
    h = np.zeros((1,) + shape, np.uint32)
    h[0][5:12, 8:20, 30:50] = rng.integers(0, 50, (7, 12, 20))
    values = np.array([4], np.uint8)
    hb = eng.upload(h)
    out = eng.alloc(h[0].size)
    eng.smooth_argmax(hb, 1, shape, iters, values, out)
    phi = np.stack([orc.smooth(x.astype(np.float32), iters) for x in h])
    np.testing.assert_array_equal(out.download(shape, np.uint8), orc.argmax(phi, values))


@pytest.mark.parametrize("big", [0, 1])
def test_smooth_argmax_two_labels_bitexact(eng, big):
    """Two labels: one on the context stream, the other alone on a side
    stream; the largest box on either label, and equal hits on a patch where
    the two boxes overlap (the tie goes to label 0)."""
    rng = np.random.default_rng(140 + big)
    shape = (24, 30, 80)
    h = np.zeros((2,) + shape, np.uint32)
    h[big] = rng.integers(0, 12, shape)
    h[1 - big][8:14, 10:16, 30:40] = rng.integers(1, 90, (6, 6, 10))
    h[0][20:22, 2:4, 70:74] = 7
    h[1][20:22, 2:4, 70:74] = 7
    values = np.array([3, 9], np.uint8)
    hb = eng.upload(h)
    out = eng.alloc(h[0].size)
    eng.smooth_argmax(hb, 2, shape, 5, values, out)
    phi = np.stack([orc.smooth(x.astype(np.float32), 5) for x in h])
    np.testing.assert_array_equal(out.download(shape, np.uint8), orc.argmax(phi, values))


@pytest.mark.parametrize("iters", [2, 3, 4])
def test_smooth_pairs_bitexact_wide(eng, iters):
    """A volume wider than one tile in x and y and taller than one z chunk."""
    shape = (70, 35, 130)
    phi = (np.random.default_rng(40 + iters).random(shape) * 20).astype(np.float32)
    a = eng.upload(phi)
    b = eng.alloc(phi.nbytes)
    eng.smooth(a, b, shape, iters)
    np.testing.assert_array_equal(a.download(shape, np.float32).view(np.uint32),
                                  orc.smooth(phi, iters).view(np.uint32))


def test_smooth_argmax_and_uncertainty_bitexact(eng):
    rng = np.random.default_rng(4)
    shape = (21, 30, 66)
    nl = 4
    h = rng.integers(0, 40, (nl,) + shape).astype(np.uint32)
    values = np.array([0, 3, 5, 7], np.uint8)
    hb = eng.upload(h)
    out = eng.alloc(h[0].size)
    eng.smooth_argmax(hb, nl, shape, 4, values, out)
    phi = np.stack([orc.smooth(x.astype(np.float32), 4) for x in h])
    np.testing.assert_array_equal(out.download(shape, np.uint8), orc.argmax(phi, values))
    uq = eng.alloc(h[0].size)
    eng.uncertainty(hb, nl, shape, uq)
    u = orc.uncertainty(h.astype(np.float32))
    ref = (u * np.float32(255.0)).astype(np.uint8)
    ref[0], ref[-1], ref[:, 0], ref[:, -1], ref[:, :, 0], ref[:, :, -1] = 0, 0, 0, 0, 0, 0
    np.testing.assert_array_equal(uq.download(shape, np.uint8), ref)


E2E = [
    dict(name="xy_u8", kw=dict(nbrw=4, sorw=300, smooth=2, uncertainty=True, return_hits=True)),
    dict(name="xy_u8_table_walker", kw=dict(nbrw=10, sorw=200, smooth=1)),
    dict(name="allaxis_u8", allaxis=True, kw=dict(nbrw=3, sorw=250, allaxis=True, smooth=1)),
    dict(name="xy_u16_float_path", u16=True, kw=dict(nbrw=3, sorw=200)),
    dict(name="xy_reference_stream", kw=dict(nbrw=2, sorw=150, rng="reference")),
]


@pytest.mark.parametrize("case", E2E, ids=[c["name"] for c in E2E])
def test_smart_interpolation_matches_oracle(case):
    from biomedisa_amd import smart_interpolation
    _say("e2e %s" % case["name"])
    allaxis = case.get("allaxis", False)
    if case.get("u16"):
        vol = structures_volume(shape=(36, 40, 44), seed=5, n_struct=4, radius=(4, 9), every=5,
                                allaxis=False, dtype=np.uint16, max_value=4095,
                                mean_range=(400, 3600), sigma=150, bias=200)
    else:
        vol = structures_volume(shape=(40, 44, 48), seed=3, n_struct=4, radius=(4, 10),
                                every=6, allaxis=allaxis)
    labels = vol.labels.astype(np.int32)
    labels[0, 0, 0] = -1                     # an ignored voxel
    res = smart_interpolation(vol.img, labels, **case["kw"])
    ref = opipe.smart_interpolation(vol.img, labels, **case["kw"])
    for k in ("regular", "smooth", "uncertainty", "hits"):
        if k in ref:
            assert k in res
            np.testing.assert_array_equal(res[k], ref[k], err_msg=k)
    assert set(res) == {k for k in ref if not k.startswith("_")}


def test_adjacent_allaxis_planes_rejected(capsys):
    from biomedisa_amd import smart_interpolation
    img = np.full((12, 12, 12), 50, np.uint8)
    lab = np.zeros((12, 12, 12), np.uint8)
    lab[4, 2:9, 2:9] = 1
    lab[5, 2:9, 2:9] = 1
    assert smart_interpolation(img, lab, nbrw=1, sorw=10, allaxis=True) is None
    assert "At least one empty slice between labels required." in capsys.readouterr().out


def test_config1_shape_dice_vs_reference_stream():
    """128^3 phantom of the CPU-reference workload, shortened walks: the Philox walker
    and the reference stream agree to per-label Dice >= 0.98."""
    from biomedisa_amd import smart_interpolation
    from biomedisa_amd.helpers import per_label_dice
    vol = ellipsoids_volume(seed=1)
    _say("config1 shape, nbrw=10 sorw=1000")
    a = smart_interpolation(vol.img, vol.labels, nbrw=10, sorw=1000)["regular"]
    b = smart_interpolation(vol.img, vol.labels, nbrw=10, sorw=1000, rng="reference")["regular"]
    d = per_label_dice(a, b)
    assert min(d.values()) >= 0.98, d


def test_config1_dice_vs_reference_stream_real_parameters():
    """The CPU-reference workload at its own parameters (128^3, 3 ellipsoids,
    labels every 8th XY plane, nbrw=10, sorw=4000, about 1e10 steps): the
    Philox walker against the reference's float MRG stream (as its compilers
    contract it), per-label Dice >= 0.98."""
    from biomedisa_amd import smart_interpolation
    from biomedisa_amd.helpers import per_label_dice
    vol = ellipsoids_volume(seed=1)
    _say("config1, nbrw=10 sorw=4000, philox vs reference stream")
    a = smart_interpolation(vol.img, vol.labels, nbrw=10, sorw=4000)["regular"]
    b = smart_interpolation(vol.img, vol.labels, nbrw=10, sorw=4000, rng="reference")["regular"]
    d = per_label_dice(a, b)
    print("per-label Dice philox vs reference stream:", d, flush=True)
    assert min(d.values()) >= 0.98, d


def test_hit_weight_scales_hits(eng):
    """hit_weight k adds k per accepted move: walk(nbrw=1, k) == k * walk(nbrw=1, 1),
    and only the tiled Philox walker accepts it."""
    from biomedisa_amd.synth import ellipsoids_volume
    vol = ellipsoids_volume(shape=(36, 40, 44), seed=31, semi=(6, 12), every=6)
    img = np.ascontiguousarray(np.pad(vol.img, 1))
    lab = np.ascontiguousarray(np.pad(vol.labels, 1))
    dims, V = img.shape, img.size
    ib, lb = eng.upload(img), eng.upload(lab)
    flags = eng.plane_flags(lb, dims, False)
    seeds, n = eng.seeds(lb, ib, _hip.IMG_U8, dims, flags)
    out = []
    for w in (1, 10):
        h = eng.alloc(4 * V * 4)
        h.zero()
        acc = eng.walk(ib, _hip.IMG_U8, dims, seeds, n, 4, 1, 300, _hip.RNG_PHILOX, 99, h,
                       algorithm=_hip.WALK_TILED, hit_weight=w)
        out.append((h.download((4,) + dims, np.uint32), acc))
    np.testing.assert_array_equal(out[1][0], 10 * out[0][0])
    assert out[0][1] == out[1][1]
    with pytest.raises(_hip.BmxError):
        eng.walk(ib, _hip.IMG_U8, dims, seeds, n, 4, 1, 300, _hip.RNG_PHILOX, 99, h,
                 algorithm=_hip.WALK_DIRECT, hit_weight=2)


def test_seed_overrides_equal_separate_walks(eng):
    """Per-seed walk counts and weights (bmx_seed_walks + seed_overrides) give the
    hits of the two seed groups walked separately: group A all nbrw walks, group B
    its first walk counted nbrw times."""
    from biomedisa_amd.synth import structures_volume
    vol = structures_volume(shape=(40, 36, 44), seed=32, n_struct=3, radius=(4, 10), every=5, allaxis=True)
    img = np.ascontiguousarray(np.pad(vol.img, 1))
    lab = np.ascontiguousarray(np.pad(vol.labels, 1))
    dims, V = img.shape, img.size
    ib, lb = eng.upload(img), eng.upload(lab)
    flags = eng.plane_flags(lb, dims, True)
    seeds, n = eng.seeds(lb, ib, _hip.IMG_U8, dims, flags)
    nl, nbrw, sorw, key = 4, 5, 200, 7
    na = n // 2
    tail = _hip.BufferView(seeds, na * _hip.SEED_DTYPE.itemsize)
    _hip.call("bmx_seed_walks", eng.ctx, tail.ptr, n - na, 1, nbrw)
    both = eng.alloc(nl * V * 4)
    both.zero()
    eng.walk(ib, _hip.IMG_U8, dims, seeds, n, nl, nbrw, sorw, _hip.RNG_PHILOX, key, both, seed_overrides=True)
    _hip.call("bmx_seed_walks", eng.ctx, tail.ptr, n - na, 0, 0)
    ha, hb = eng.alloc(nl * V * 4), eng.alloc(nl * V * 4)
    ha.zero()
    hb.zero()
    eng.walk(ib, _hip.IMG_U8, dims, seeds, na, nl, nbrw, sorw, _hip.RNG_PHILOX, key, ha,
             algorithm=_hip.WALK_TILED)
    eng.walk(ib, _hip.IMG_U8, dims, tail, n - na, nl, 1, sorw, _hip.RNG_PHILOX, key, hb,
             algorithm=_hip.WALK_TILED)
    shape = (nl,) + dims
    ref = ha.download(shape, np.uint32) + nbrw * hb.download(shape, np.uint32)
    np.testing.assert_array_equal(both.download(shape, np.uint32), ref)


def test_auto_walker_with_overrides_on_few_seeds(eng):
    """Fewer than 2^16 walkers with per-seed overrides: the automatic choice takes
    the tiled walker (the direct one has no weights) and the hits equal the
    explicitly tiled walk."""
    from biomedisa_amd.synth import structures_volume
    vol = structures_volume(shape=(24, 26, 28), seed=33, n_struct=2, radius=(3, 6), every=8, allaxis=False)
    img = np.ascontiguousarray(np.pad(vol.img, 1))
    lab = np.ascontiguousarray(np.pad(vol.labels, 1))
    dims, V = img.shape, img.size
    ib, lb = eng.upload(img), eng.upload(lab)
    flags = eng.plane_flags(lb, dims, False)
    seeds, n = eng.seeds(lb, ib, _hip.IMG_U8, dims, flags)
    nl, nbrw, sorw, key = 3, 4, 150, 11
    m = min(n, 300)
    assert m * nbrw < 1 << 16
    tail = _hip.BufferView(seeds, (m // 2) * _hip.SEED_DTYPE.itemsize)
    _hip.call("bmx_seed_walks", eng.ctx, tail.ptr, m - m // 2, 2, 2)
    h_auto, h_tiled = eng.alloc(nl * V * 4), eng.alloc(nl * V * 4)
    h_auto.zero()
    h_tiled.zero()
    a1 = eng.walk(ib, _hip.IMG_U8, dims, seeds, m, nl, nbrw, sorw, _hip.RNG_PHILOX, key, h_auto, seed_overrides=True)
    a2 = eng.walk(ib, _hip.IMG_U8, dims, seeds, m, nl, nbrw, sorw, _hip.RNG_PHILOX, key, h_tiled,
                  seed_overrides=True, algorithm=_hip.WALK_TILED)
    assert a1 == a2 > 0
    np.testing.assert_array_equal(h_auto.download((nl,) + dims, np.uint32), h_tiled.download((nl,) + dims, np.uint32))


@pytest.mark.parametrize("w", [16, 20])
def test_tile_counter_overflow_path(eng, w):
    """Heavy weights and many walks from two seeds in a small uniform volume push
    single voxels far past 65535 hits within one round; totals must equal w times
    the unweighted walk."""
    dims = (14, 14, 14)
    img = np.full(dims, 100, np.uint8)
    lab = np.zeros(dims, np.uint8)
    lab[7, 6, 6] = 1
    lab[7, 7, 7] = 2
    lab[7, 6, 7] = 255
    flags = np.zeros(sum(dims), np.uint8)
    flags[7] = 1
    ib, lb, fb = eng.upload(img), eng.upload(lab), eng.upload(flags)
    seeds, n = eng.seeds(lb, ib, _hip.IMG_U8, dims, fb)
    V = int(np.prod(dims))
    out = []
    for weight in (1, w):
        h = eng.alloc(3 * V * 4)
        h.zero()
        eng.walk(ib, _hip.IMG_U8, dims, seeds, n, 3, 3000, 400, _hip.RNG_PHILOX, 5, h,
                 algorithm=_hip.WALK_TILED, hit_weight=weight)
        out.append(h.download((3,) + dims, np.uint32).astype(np.uint64))
    assert out[1].max() > 3 * 65536
    np.testing.assert_array_equal(out[1], w * out[0])
