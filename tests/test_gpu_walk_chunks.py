"""A single volume with more walkers than one tiled walk holds (2^32 - 1) is
walked in chunks of whole seeds.  BMX_WALK_CHUNK lowers the chunk size so that
the chunked walk can be checked on small volumes: against the oracle's hits and
move count, and against the unchunked walk with per-seed walk counts and
weights.  The walk at full size (more than 2^32 walkers) is opt-in."""
import os

import numpy as np
import pytest

from oracle import orc
from biomedisa_amd import _hip
from biomedisa_amd.synth import structures_volume

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _volume(eng, shape, seed, every, allaxis):
    vol = structures_volume(shape=shape, seed=seed, n_struct=3, radius=(4, 10), every=every, allaxis=allaxis)
    img = np.ascontiguousarray(np.pad(vol.img, 1))
    lab = np.ascontiguousarray(np.pad(vol.labels, 1))
    ib, lb = eng.upload(img), eng.upload(lab)
    flags = eng.plane_flags(lb, img.shape, allaxis)
    seeds, n = eng.seeds(lb, ib, _hip.IMG_U8, img.shape, flags)
    return img, ib, seeds, n


@pytest.mark.parametrize("graph_rounds", [0, -1])
@pytest.mark.parametrize("chunk", [997, 4000])
def test_chunked_walk_equals_oracle(eng, monkeypatch, graph_rounds, chunk):
    img, ib, seeds, n = _volume(eng, (36, 40, 38), 41, 6, True)
    nl, nbrw, sorw, key = 4, 3, 150, 19
    assert n * nbrw > 2 * chunk, "the walk must split into several chunks"
    _say("chunked walk: %d seeds, chunk %d walkers, graph_rounds %d" % (n, chunk, graph_rounds))
    monkeypatch.setenv("BMX_WALK_CHUNK", str(chunk))
    hits = eng.alloc(nl * img.size * 4)
    hits.zero()
    acc = eng.walk(ib, _hip.IMG_U8, img.shape, seeds, n, nl, nbrw, sorw, _hip.RNG_PHILOX, key, hits,
                   algorithm=_hip.WALK_TILED, graph_rounds=graph_rounds)
    _, kern, launches = eng.last_walk_stats()
    sd = seeds.download((n,), _hip.SEED_DTYPE)
    oh, oacc = orc.philox_walk(img, sd["pos"].astype(np.int64), sd["label"].astype(np.uint8), sd["var"],
                               nl, nbrw, sorw, key)
    assert acc == oacc
    np.testing.assert_array_equal(hits.download((nl,) + img.shape, np.uint32), oh)
    # every chunk runs at least one round
    assert launches >= -(-n * nbrw // chunk) and kern > 0


def test_chunked_walk_with_overrides_equals_unchunked(eng, monkeypatch):
    """Per-seed walk counts and weights (the guided method's ring seeds) give the
    same hits chunked and unchunked."""
    img, ib, seeds, n = _volume(eng, (40, 36, 44), 42, 5, True)
    nl, nbrw, sorw, key = 4, 6, 180, 23
    tail = _hip.BufferView(seeds, (n // 3) * _hip.SEED_DTYPE.itemsize)
    _hip.call("bmx_seed_walks", eng.ctx, tail.ptr, n - n // 3, 2, 3)
    shape = (nl,) + img.shape
    out = []
    for chunk in (None, 1201):
        if chunk is None:
            monkeypatch.delenv("BMX_WALK_CHUNK", raising=False)
        else:
            monkeypatch.setenv("BMX_WALK_CHUNK", str(chunk))
        _say("overrides walk, chunk %s" % chunk)
        h = eng.alloc(nl * img.size * 4)
        h.zero()
        a = eng.walk(ib, _hip.IMG_U8, img.shape, seeds, n, nl, nbrw, sorw, _hip.RNG_PHILOX, key, h,
                     seed_overrides=True)
        out.append((a, h.download(shape, np.uint32)))
    assert out[0][0] == out[1][0] > 0
    np.testing.assert_array_equal(out[0][1], out[1][1])


@pytest.mark.skipif(os.environ.get("BMX_LARGE_TESTS") != "1",
                    reason="opt-in: about 120 GB of device memory (BMX_LARGE_TESTS=1)")
def test_walk_above_2_32_walkers(eng):
    """1024^3 uint8 with every other XY plane labelled (two labels): 5.3e8 seeds
    x 10 walks = 5.3e9 walkers, above 2^32, walked in two chunks.  The hits
    equal those of two walks over the halves of the seeds, each below 2^32
    walkers, compared slab by slab."""
    Z = Y = X = 1024
    dims = (Z, Y, X)
    V = Z * Y * X
    rng = np.random.default_rng(5)
    img = rng.integers(90, 110, size=dims, dtype=np.uint8)
    lab = np.zeros(dims, np.uint8)
    lab[1:-1:2, 1:-1, 1:X // 2] = 1
    ib, lb = eng.upload(img), eng.upload(lab)
    del img, lab
    flags = eng.plane_flags(lb, dims, False)
    seeds, n = eng.seeds(lb, ib, _hip.IMG_U8, dims, flags)
    nl, nbrw, sorw, key = 2, 10, 4, 29
    _say("above 2^32: %d seeds, %d walkers" % (n, n * nbrw))
    assert n * nbrw >= 1 << 32 and (n // 2) * nbrw < 1 << 32
    whole = eng.alloc(nl * V * 4)
    whole.zero()
    a = eng.walk(ib, _hip.IMG_U8, dims, seeds, n, nl, nbrw, sorw, _hip.RNG_PHILOX, key, whole,
                 algorithm=_hip.WALK_TILED)
    _say("whole walk done: %d moves" % a)
    halves = eng.alloc(nl * V * 4)
    halves.zero()
    h = n // 2
    a1 = eng.walk(ib, _hip.IMG_U8, dims, seeds, h, nl, nbrw, sorw, _hip.RNG_PHILOX, key, halves,
                  algorithm=_hip.WALK_TILED)
    rest = _hip.BufferView(seeds, h * _hip.SEED_DTYPE.itemsize)
    a2 = eng.walk(ib, _hip.IMG_U8, dims, rest, n - h, nl, nbrw, sorw, _hip.RNG_PHILOX, key, halves,
                  algorithm=_hip.WALK_TILED)
    _say("half walks done")
    assert a == a1 + a2 > 0
    slab = 1 << 28
    for off in range(0, nl * V, slab):
        w = _hip.BufferView(whole, off * 4, slab * 4).download((slab,), np.uint32)
        v = _hip.BufferView(halves, off * 4, slab * 4).download((slab,), np.uint32)
        np.testing.assert_array_equal(w, v)
