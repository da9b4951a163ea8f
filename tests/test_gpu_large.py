"""Large volume: 1024^3 uint16 (12-bit) with 16 structures, labels on every 16th
plane of all axes; plain and surrogate-guided on the same GPU with all uint32 hit
volumes resident in HBM, compared by per-label Dice and time."""
import gc
import os
import time

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("BMX_LARGE") != "1",
                                 reason="the 1024^3 workload takes about five minutes; run with BMX_LARGE=1 "
                                        "(bench.py --workload large measures it)")]


def test_large_uint16_guided_vs_plain(eng):
    from biomedisa_amd import helpers as H
    from biomedisa_amd import surrogate as S
    from biomedisa_amd.interpolation import DeviceVolume, run_pipeline, run_pipeline_guided
    from biomedisa_amd.synth import large_volume
    t0 = time.perf_counter()
    vol = large_volume(seed=5, size=1024, log=lambda m: print(m, flush=True))
    bm = H.Biomedisa()
    bm.data, bm.labelData, bm.only, bm.ignore, bm.success = vol.img, vol.labels, "all", "none", True
    bm = H.pre_processing(bm)
    nol = bm.nol
    assert 10 <= nol <= 17
    img = np.pad(H.normalize_image(bm.data), 1)
    lab = np.pad(bm.labelData, 1)
    z0, z1, y0, y1, x0, x1 = H.predict_blocksize(lab)
    img = np.ascontiguousarray(img[z0:z1, y0:y1, x0:x1])
    labidx = H.label_index_volume(np.ascontiguousarray(lab[z0:z1, y0:y1, x0:x1]), bm.allLabels)
    values = bm.allLabels.astype(np.uint8)
    assert img.dtype == np.float32
    del vol, bm
    gc.collect()
    dv = DeviceVolume(eng, img, labidx)
    print("prepared %s in %.0fs" % (dv.dims, time.perf_counter() - t0), flush=True)
    del img, labidx
    t = time.perf_counter()
    p = run_pipeline(dv, nol, values, nbrw=10, sorw=4000, allaxis=True)
    eng.sync()
    tp = time.perf_counter() - t
    preg = p["regular"].download(dv.dims, np.uint8)
    nseeds = p["nseeds"]
    del p
    gc.collect()
    net = S.net(0)
    t = time.perf_counter()
    g = run_pipeline_guided(dv, nol, values, net, nbrw=10, sorw=4000, allaxis=True)
    eng.sync()
    tg = time.perf_counter() - t
    greg = g["regular"].download(dv.dims, np.uint8)
    d = H.per_label_dice(greg, preg)
    print("plain %.1fs (%d seeds), guided %.1fs (%.3f near + %.3f ring), speedup %.2f, min per-label Dice %.4f"
          % (tp, nseeds, tg, g["walked_seeds"] / nseeds, g["ring_seeds"] / nseeds, tp / tg, min(d.values())),
          flush=True)
    print("per-label Dice guided vs plain:", {k: round(v, 4) for k, v in d.items()}, flush=True)
    assert len(d) == nol - 1
    assert min(d.values()) >= 0.98
    assert tg < tp
