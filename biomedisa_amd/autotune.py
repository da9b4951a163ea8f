"""Autotuner of the tiled walker's launch shape.

Three choices change only the speed of a tiled walk, never its hits (every
walker follows its own Philox stream whatever the order of work):

    tile_shape     core side, halo and workgroup size of a tile (BMX_TILE_SHAPES
                   compiled variants; small tiles fit more workgroups per CU
                   and suit small or sparsely seeded volumes)
    segment_steps  the most steps a walker takes per round (more steps per
                   round: fewer re-bucketing passes, longer tails)
    graph_rounds   -1 = rounds driven by the host (a workgroup per bucket),
                   k = k rounds per replay of one captured HIP graph with a
                   persistent walk kernel (0 = 8)

tune_walk times the candidates on a sample of the workload itself - every
seed, one walk each (1/nbrw of the work, the same seed geometry) - and keeps
the fastest.  The search is greedy: shapes at the default cap, then caps at
the best shape, then the round driver.  Results are cached in-process per
workload class (image type, volume count, voxels and seeds to within a factor
of 2, nbrw, sorw).
"""
import math
import time

from . import _hip

_CACHE = {}
CAPS = (64, 128, 256, 512)
GRAPH_ROUNDS = 8


def _class_key(batch, nbrw, sorw):
    nseeds = int(sum(batch.nseeds))
    return (batch.img_dtype, len(batch.dims), int(math.log2(max(1, batch.total_voxels))),
            int(math.log2(max(1, nseeds))), int(nbrw), int(sorw))


def _time_walk(batch, hits, walks, sorw, key, opts):
    eng = batch.eng
    hits.zero()
    eng.sync()
    t0 = time.perf_counter()
    eng.walk_batch(batch.img, batch.img_dtype, batch.table, batch.seeds, walks, sorw, key, hits, **opts)
    eng.sync()
    return time.perf_counter() - t0


def tune_walk(batch, nbrw, sorw, key, sample_walks=1, log=None, cache=True):
    """Fastest walk options for `batch` (a DeviceBatch): a dict of
    tile_shape / segment_steps / graph_rounds for Engine.walk_batch."""
    ck = _class_key(batch, nbrw, sorw)
    if cache and ck in _CACHE:
        return dict(_CACHE[ck])
    eng = batch.eng
    walks = max(1, min(int(sample_walks), int(nbrw)))
    hits = eng.alloc(max(1, batch.hits_elems) * 4)
    table = []
    try:
        def trial(opts):
            sec = _time_walk(batch, hits, walks, sorw, key, opts)
            table.append((dict(opts), sec))
            if log:
                log("[autotune] %s: %.1f ms" % (opts, 1e3 * sec))
            return sec

        # warm-up of the default shape (first-launch costs stay out of the comparison)
        _time_walk(batch, hits, walks, sorw, key, {})
        best = min(({"tile_shape": s, "segment_steps": 256, "graph_rounds": GRAPH_ROUNDS}
                    for s in range(_hip.TILE_SHAPES)), key=trial)
        cands = [dict(best, segment_steps=c) for c in CAPS if c != 256]
        best_sec = min(sec for o, sec in table if o == best)
        for o in cands:
            sec = trial(o)
            if sec < best_sec:
                best, best_sec = o, sec
        for g in (dict(best, graph_rounds=-1), dict(best, graph_rounds=2 * GRAPH_ROUNDS)):
            sec = trial(g)
            if sec < best_sec:
                best, best_sec = g, sec
    finally:
        hits.free()
    if cache:
        _CACHE[ck] = dict(best)
    tune_walk.last_table = table
    return dict(best)


tune_walk.last_table = []
