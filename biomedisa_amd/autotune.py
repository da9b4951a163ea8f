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

tune_walk times the candidates on a sample of the workload itself and keeps
the fastest.  The sample is the first volumes of the batch that hold about
1/nbrw of its seeds, walked with all nbrw walks: 1/nbrw of the work at the
walk's own walker density per tile (which decides how full the tile buckets
are, and so which tile shape wins).  A batch whose first volume holds more
than that (or a single volume) is sampled with every seed walking once
instead.  The search is greedy: shapes at the default cap, then caps at
the best shape, then the round driver.  Results are cached in-process per
workload class (image type, volume count, voxels and seeds to within a factor
of 2, nbrw, sorw).
"""
import math
import time

import numpy as np

from . import _hip

_CACHE = {}
CAPS = (64, 128, 256, 512)
GRAPH_ROUNDS = 8


def _class_key(batch, nbrw, sorw):
    nseeds = int(sum(batch.nseeds))
    return (batch.img_dtype, len(batch.dims), int(math.log2(max(1, batch.total_voxels))),
            int(math.log2(max(1, nseeds))), int(nbrw), int(sorw))


def _time_walk(batch, table, hits, walks, sorw, key, opts):
    eng = batch.eng
    hits.zero()
    eng.sync()
    t0 = time.perf_counter()
    eng.walk_batch(batch.img, batch.img_dtype, table, batch.seeds, walks, sorw, key, hits, **opts)
    eng.sync()
    return time.perf_counter() - t0


def sample_table(batch, nbrw):
    """The fewest leading rows of batch.table whose volumes hold at least
    1/nbrw of the seeds (their seeds are the first of the seed list, as the
    batch walk requires), or None when those hold more than 2/nbrw of the seeds
    or no proper subset does."""
    counts = np.asarray(batch.nseeds, dtype=np.int64)
    total = int(counts.sum())
    if len(counts) < 2 or total == 0 or nbrw <= 1:
        return None
    cum = np.cumsum(counts)
    m = int(np.searchsorted(cum, total / nbrw)) + 1
    if m >= len(counts) or cum[m - 1] > 2 * total / nbrw:
        return None
    return batch.table[:m]


def tune_walk(batch, nbrw, sorw, key, sample_walks=1, log=None, cache=True):
    """Fastest walk options for `batch` (a DeviceBatch): a dict of
    tile_shape / segment_steps / graph_rounds for Engine.walk_batch."""
    ck = _class_key(batch, nbrw, sorw)
    if cache and ck in _CACHE:
        return dict(_CACHE[ck])
    eng = batch.eng
    sub = sample_table(batch, nbrw)
    if sub is not None:
        vols, walks = sub, int(nbrw)
    else:
        vols, walks = batch.table, max(1, min(int(sample_walks), int(nbrw)))
    hits = eng.alloc(max(1, batch.hits_elems) * 4)
    table = []
    try:
        def trial(opts):
            sec = _time_walk(batch, vols, hits, walks, sorw, key, opts)
            table.append((dict(opts), sec))
            if log:
                log("[autotune] %s: %.1f ms" % (opts, 1e3 * sec))
            return sec

        # warm-up of the default shape (first-launch costs stay out of the comparison)
        _time_walk(batch, vols, hits, walks, sorw, key, {})
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
