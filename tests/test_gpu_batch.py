"""Batched walk (bmx_walk_batch), tile shapes and graph-driven rounds.

Every variant must give the oracle's hit counts exactly: the tile shape, the
round cap and the round driver change only the order of work, and a volume
walked inside a batch follows the same Philox streams as walked alone.
"""
import os

import numpy as np
import pytest

from oracle import orc
from biomedisa_amd import _hip
from biomedisa_amd.synth import structures_volume

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _volume(shape, seed, n_struct, every, allaxis, dtype=np.uint8):
    vol = structures_volume(shape=shape, seed=seed, n_struct=n_struct, radius=(3, max(4, min(shape) // 3)),
                            every=every, allaxis=allaxis)
    img = np.ascontiguousarray(np.pad(vol.img, 1))
    if dtype == np.float32:
        img = np.minimum(img.astype(np.float32) + np.random.default_rng(seed).random(img.shape, dtype=np.float32),
                         np.float32(255))
    lab = np.ascontiguousarray(np.pad(vol.labels, 1))
    return img, lab, n_struct + 1


def _single(eng, img, lab, nl, allaxis, nbrw, sorw, key, **opts):
    dims = img.shape
    dt = _hip.IMG_U8 if img.dtype == np.uint8 else _hip.IMG_F32
    ib, lb = eng.upload(img), eng.upload(lab)
    fb = eng.plane_flags(lb, dims, allaxis)
    sb, n = eng.seeds(lb, ib, dt, dims, fb)
    h = eng.alloc(nl * img.size * 4)
    h.zero()
    acc = eng.walk(ib, dt, dims, sb, n, nl, nbrw, sorw, _hip.RNG_PHILOX, key, h, algorithm=_hip.WALK_TILED,
                   **opts)
    seeds = sb.download((n,), _hip.SEED_DTYPE) if n else np.zeros(0, _hip.SEED_DTYPE)
    return h.download((nl,) + dims, np.uint32), acc, seeds


@pytest.mark.parametrize("graph_rounds", [-1, 0, 4])
@pytest.mark.parametrize("shape_idx", range(_hip.TILE_SHAPES))
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_tile_shapes_and_graph_rounds_bitexact(eng, dtype, shape_idx, graph_rounds):
    """Each compiled tile shape, host- or graph-driven rounds, a small round cap
    (many rounds, so the graph is replayed several times) and sorw % 4 != 0: the
    oracle's hits and accepted-move count."""
    _say("tile shape %d graph_rounds %d dtype %s" % (shape_idx, graph_rounds, np.dtype(dtype).name))
    img, lab, nl = _volume((30, 41, 37), 61, 3, 5, True, dtype)
    nbrw, sorw, key = 3, 203, 991
    h, acc, seeds = _single(eng, img, lab, nl, True, nbrw, sorw, key, segment_steps=8, tile_shape=shape_idx,
                            graph_rounds=graph_rounds)
    oh, oacc = orc.philox_walk(img, seeds["pos"].astype(np.int64), seeds["label"].astype(np.uint8),
                               seeds["var"], nl, nbrw, sorw, key)
    assert acc == oacc
    np.testing.assert_array_equal(h, oh)


def _hit16_shapes(dtype):
    """Tile shapes with 16-bit LDS hit counters (read from the library; none
    while it is not built, so that collection never fails)."""
    if not os.path.exists(_hip.LIB_PATH):
        return []
    dt = _hip.IMG_U8 if dtype == np.uint8 else _hip.IMG_F32
    return [s for s in range(_hip.TILE_SHAPES) if _hip.tile_shape_info(dt, s)["hit_bytes"] == 2]


@pytest.mark.parametrize("weight", [40000, 65535, 65536])
@pytest.mark.parametrize("dtype,shape_idx", [(d, s) for d in (np.uint8, np.float32) for s in _hit16_shapes(d)])
def test_hit16_counter_wraps_exact(eng, dtype, shape_idx, weight):
    """Tile shapes with 16-bit LDS hit counters, under hit weights that wrap a
    counter on most adds (65535: the high half of a word sits at 0xFFFF when a
    carry from its low half arrives): every hit equals weight x the oracle's
    count.  A weight of 2^16 takes the 32-bit sibling shape; same bar."""
    _say("16-bit counters: shape %d weight %d dtype %s" % (shape_idx, weight, np.dtype(dtype).name))
    img, lab, nl = _volume((33, 40, 36), 65, 3, 6, True, dtype)
    nbrw, sorw, key = 4, 301, 4242
    for graph_rounds in (0, -1):
        h, acc, seeds = _single(eng, img, lab, nl, True, nbrw, sorw, key, segment_steps=32, tile_shape=shape_idx,
                                graph_rounds=graph_rounds, hit_weight=weight)
        oh, oacc = orc.philox_walk(img, seeds["pos"].astype(np.int64), seeds["label"].astype(np.uint8),
                                   seeds["var"], nl, nbrw, sorw, key)
        assert acc == oacc
        assert int(oh.max()) * weight < 2 ** 32
        np.testing.assert_array_equal(h, (oh.astype(np.uint64) * weight).astype(np.uint32))


@pytest.mark.parametrize("shape_idx", _hit16_shapes(np.uint8))
def test_hit16_seed_weights_exact(eng, shape_idx):
    """Per-seed walk counts and hit weights (seed overrides) on the 16-bit
    shapes: weights up to 65535 per seed, then hit_weight 2 on top (the 32-bit
    sibling shape): weight x the oracle's per-seed hits."""
    img, lab, nl = _volume((30, 34, 38), 66, 3, 5, True, np.uint8)
    nbrw, sorw, key = 4, 257, 77
    dims = img.shape
    ib, lb = eng.upload(img), eng.upload(lab)
    fb = eng.plane_flags(lb, dims, True)
    sb, n = eng.seeds(lb, ib, _hip.IMG_U8, dims, fb)
    seeds = sb.download((n,), _hip.SEED_DTYPE)
    rng = np.random.default_rng(5)
    walks = rng.integers(1, nbrw + 1, size=n).astype(np.uint32)
    wts = rng.choice(np.array([1, 3, 40000, 65535], np.uint32), size=n)
    seeds["reserved"] = walks | (wts << 16)
    sb2 = eng.upload(seeds)
    for hit_weight in (1, 2):
        _say("16-bit counters, seed overrides: shape %d hit_weight %d" % (shape_idx, hit_weight))
        h = eng.alloc(nl * img.size * 4)
        h.zero()
        acc = eng.walk(ib, _hip.IMG_U8, dims, sb2, n, nl, nbrw, sorw, _hip.RNG_PHILOX, key, h,
                       algorithm=_hip.WALK_TILED, tile_shape=shape_idx, seed_overrides=True, hit_weight=hit_weight,
                       segment_steps=32)
        hits = h.download((nl,) + dims, np.uint32)
        expect = np.zeros((nl,) + dims, np.uint64)
        oacc = 0
        for w in np.unique(wts):
            for k in range(1, nbrw + 1):
                sel = (wts == w) & (walks == k)
                if not sel.any():
                    continue
                s = seeds[sel]
                oh, a = orc.philox_walk(img, s["pos"].astype(np.int64), s["label"].astype(np.uint8), s["var"], nl,
                                        k, sorw, key)
                expect += oh.astype(np.uint64) * int(w) * hit_weight
                oacc += a
        assert acc == oacc
        assert int(expect.max()) < 2 ** 32
        np.testing.assert_array_equal(hits, expect.astype(np.uint32))
        h.free()


def _batch_items(dtype):
    shapes = [(22, 30, 26), (40, 24, 33), (17, 17, 17), (35, 44, 29), (26, 21, 38)]
    items = []
    for k, shp in enumerate(shapes):
        img, lab, nl = _volume(shp, 70 + k, 2 + (k * 3) % 7, 4 + k % 3, k % 2 == 0, dtype)
        items.append((img, lab, nl, k % 2 == 0))
    # a volume without labelled planes: no seeds, hits stay zero
    img, lab, nl = _volume((19, 23, 21), 99, 2, 5, False, dtype)
    items.insert(2, (img, np.zeros_like(lab), nl, False))
    return items


@pytest.mark.parametrize("opts", [dict(), dict(tile_shape=3, segment_steps=12, graph_rounds=-1),
                                  dict(graph_rounds=6, tile_shape=2),
                                  dict(tile_shape=1, graph_rounds=2, segment_steps=64)],
                         ids=["default", "shape3-cap12-host", "shape2-graph6", "shape1-graph2-cap64"])
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_walk_batch_equals_single_volume_walks(eng, dtype, opts):
    """Six volumes of uneven shape and label count (one without seeds) in one
    batched walk: each volume's hit block equals its single-volume walk, and the
    accepted moves add up."""
    from biomedisa_amd.batch import DeviceBatch
    _say("walk_batch dtype %s opts %s" % (np.dtype(dtype).name, opts))
    items = _batch_items(dtype)
    nbrw, sorw, key = 4, 150, 12345
    batch = DeviceBatch(eng, [(img, lab, nl, ax) for img, lab, nl, ax in items])
    assert batch.nseeds[2] == 0
    hits = eng.alloc(batch.hits_elems * 4)
    hits.zero()
    acc = eng.walk_batch(batch.img, batch.img_dtype, batch.table, batch.seeds, nbrw, sorw, key, hits, **opts)
    total = 0
    for k, (img, lab, nl, ax) in enumerate(items):
        want, a, _ = _single(eng, img, lab, nl, ax, nbrw, sorw, key)
        got = _hip.BufferView(hits, batch.hits_offsets[k] * 4, nl * img.size * 4).download((nl,) + img.shape,
                                                                                          np.uint32)
        np.testing.assert_array_equal(got, want, err_msg="volume %d" % k)
        total += a
    assert acc == total
    # the same batch again, accumulating: every hit doubles
    acc2 = eng.walk_batch(batch.img, batch.img_dtype, batch.table, batch.seeds, nbrw, sorw, key, hits, **opts)
    assert acc2 == acc
    k = 3
    img, lab, nl, ax = items[k]
    want, _, _ = _single(eng, img, lab, nl, ax, nbrw, sorw, key)
    got = _hip.BufferView(hits, batch.hits_offsets[k] * 4, nl * img.size * 4).download((nl,) + img.shape,
                                                                                      np.uint32)
    np.testing.assert_array_equal(got, 2 * want)
    hits.free()
    batch.free()


def test_walk_batch_rejects_bad_tables(eng):
    from biomedisa_amd.batch import DeviceBatch
    items = _batch_items(np.uint8)[:2]
    batch = DeviceBatch(eng, items)
    hits = eng.alloc(batch.hits_elems * 4)
    bad = batch.table.copy()
    bad[1]["offset"] = 0   # overlaps volume 0
    with pytest.raises(_hip.BmxError):
        eng.walk_batch(batch.img, batch.img_dtype, bad, batch.seeds, 2, 50, 1, hits)
    bad = batch.table.copy()
    bad[1]["seed_offset"] += 1   # seeds not contiguous
    with pytest.raises(_hip.BmxError):
        eng.walk_batch(batch.img, batch.img_dtype, bad, batch.seeds, 2, 50, 1, hits)
    with pytest.raises(_hip.BmxError):
        eng.walk_batch(batch.img, batch.img_dtype, batch.table, batch.seeds, 2, 50, 1, hits,
                       tile_shape=_hip.TILE_SHAPES)
    hits.free()
    batch.free()


def test_smart_interpolation_batch_equals_per_volume_calls():
    """The public batch API returns, per volume, what smart_interpolation returns
    for that volume alone (regular and smoothed labels), and None for a rejected
    input."""
    from biomedisa_amd.batch import smart_interpolation_batch
    from biomedisa_amd.interpolation import smart_interpolation
    vols = [structures_volume(shape=s, seed=80 + k, n_struct=2 + k, radius=(3, 9), every=6, allaxis=False)
            for k, s in enumerate([(30, 34, 28), (24, 40, 36), (36, 26, 31)])]
    datas = [v.img for v in vols] + [vols[0].img]
    labels = [v.labels for v in vols] + [np.zeros_like(vols[0].labels)]
    _say("smart_interpolation_batch on %d volumes" % len(datas))
    res = smart_interpolation_batch(datas, labels, nbrw=4, sorw=300, smooth=2, tune=False)
    assert res[3] is None
    for k in range(3):
        one = smart_interpolation(datas[k], labels[k], nbrw=4, sorw=300, smooth=2)
        np.testing.assert_array_equal(res[k]["regular"], one["regular"])
        np.testing.assert_array_equal(res[k]["smooth"], one["smooth"])


def test_autotune_picks_valid_options(eng):
    """The tuner returns options within the compiled range whose walk equals the
    default walk, and caches them per workload class."""
    from biomedisa_amd.autotune import tune_walk
    from biomedisa_amd.batch import DeviceBatch
    items = _batch_items(np.uint8)
    batch = DeviceBatch(eng, items)
    _say("autotune on a %d-volume batch" % len(items))
    opts = tune_walk(batch, 4, 150, 5, cache=False)
    assert set(opts) == {"tile_shape", "segment_steps", "graph_rounds"}
    assert 0 <= opts["tile_shape"] < _hip.TILE_SHAPES
    assert len(tune_walk.last_table) == _hip.TILE_SHAPES + 5
    hits = eng.alloc(batch.hits_elems * 4)
    outs = []
    for o in (opts, {}):
        hits.zero()
        eng.walk_batch(batch.img, batch.img_dtype, batch.table, batch.seeds, 4, 150, 5, hits, **o)
        outs.append(hits.download((batch.hits_elems,), np.uint32))
    np.testing.assert_array_equal(outs[0], outs[1])
    hits.free()
    batch.free()
