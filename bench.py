"""Smart-interpolation throughput on one MI355X.

Workload (default): a 512^3 uint8 volume of 8 overlapping ellipsoids/tubes
(radius 5-60, means spread over 30-230, noise sigma 12, bias +-15), labels 0..8
on every 10th plane of all three axes (allaxis), nbrw=10, sorw=4000, curvature
smoothing with smooth=100.  A step is one full device pipeline over that
volume: plane detection, seed extraction, the walk, the argmax merge and the
smoothing + merge, with image and label volume already resident in HBM.

Both pipelines run on the same resident volume: the plain walker over every
seed, and the surrogate-guided pipeline (network proposal, uncertain set,
walks of the seeds near it, merge).  `value` is the guided throughput; the
config carries the plain throughput, the speedup and the per-label Dice of
guided against plain.  `roofline` is the walker kernel of the plain pipeline:
algorithmic bytes (6 neighbour reads per step + one 8-byte hit RMW per
accepted move) over the summed HIP-event time of its launches; `cpu_baseline`
times the C restatement of the reference kernel (oracle ref_walk, float MRG
stream) on a bounded sample of the same seeds and extrapolates.

--workload batch measures the batched plain method instead: 64 uint8 volumes
(shapes uniform in 96^3 .. 256x320x320, 2-12 structures, noise 8-20, labels on
every 6th-12th XY plane), nbrw=10, sorw=4000, no smoothing, all walked in one
bmx_walk_batch call with autotuned tile shape, round cap and round driver.
`value` is batch voxels/s; the config carries the same volumes interpolated
one call at a time on the same GPU and checks that both give identical labels.

--workload large measures the large volume: 1024^3 uint16 (12-bit), 16
structures, noise sigma 40, bias +-200, labels on every 16th plane of all axes,
nbrw=10, sorw=4000, no smoothing; plain and surrogate-guided on the same GPU
with every uint32 hit volume resident in HBM.  `value` is the guided voxels/s;
the config carries the plain time and the per-label Dice of guided vs plain.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "k_tile_walk_512_traffic.md")


def measured_traffic(args, launches):
    """HBM bytes per walker launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload (None for other sizes or walk parameters).  The
    summary gives bytes per launch over its own launch count; the walk's total
    does not depend on how many empty rounds close its last graph replay (40 or
    48 launches), so it is spread over this run's `launches`."""
    if (args.size, args.nbrw, args.sorw) != (512, 10, 4000) or not os.path.exists(TRAFFIC_PROFILE):
        return None
    txt = open(TRAFFIC_PROFILE).read()
    i = txt.find("```json")
    if i < 0:
        return None
    prof = json.loads(txt[i + 7:txt.index("```", i + 7)])
    return float(prof["traffic_bytes_per_launch"]) * int(prof["launches"]) / max(int(launches), 1)


COUNTERS_PROFILE = os.path.join(HERE, "profiles", "k_tile_walk_512_counters.md")


def measured_valu(args):
    """VALU issue figures of the plain walker from the committed rocprofv3 counter
    summary of this workload (None for other sizes or when absent)."""
    if (args.size, args.nbrw, args.sorw) != (512, 10, 4000) or not os.path.exists(COUNTERS_PROFILE):
        return None
    txt = open(COUNTERS_PROFILE).read()
    i = txt.find("```json")
    if i < 0:
        return None
    return json.loads(txt[i + 7:txt.index("```", i + 7)]).get("valu")


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def prepare(args, vol=None):
    """The benchmark volume and its host-prepared walk inputs (cropped padded
    image, label-index volume), as the CPU baseline and the post workload use."""
    from biomedisa_amd.synth import structures_volume
    from biomedisa_amd import helpers as H
    if vol is None:
        t = time.perf_counter()
        n = args.size
        vol = structures_volume(shape=(n, n, n), seed=2, n_struct=8, radius=(5, 60), every=10,
                                allaxis=True)
        log("volume %d^3 generated in %.1fs" % (n, time.perf_counter() - t))
    labels = vol.labels
    bm = H.Biomedisa()
    bm.data, bm.labelData, bm.only, bm.ignore, bm.success = vol.img, labels, "all", "none", True
    bm = H.pre_processing(bm)
    img = np.pad(H.normalize_image(bm.data), 1)
    lab = np.pad(bm.labelData, 1)
    z0, z1, y0, y1, x0, x1 = H.predict_blocksize(lab)
    img_c = np.ascontiguousarray(img[z0:z1, y0:y1, x0:x1])
    labidx = H.label_index_volume(np.ascontiguousarray(lab[z0:z1, y0:y1, x0:x1]), bm.allLabels)
    return vol, img_c, labidx, bm.allLabels.astype(np.uint8), bm.nol


def cpu_baseline(img_c, seeds, nbrw, sorw, nlabels, budget_s):
    """Time oracle.ref_walk on a growing random subset of the seeds; extrapolate."""
    from oracle import orc
    rng = np.random.default_rng(0)
    n_total = seeds.size
    take = 64
    while True:
        idx = np.sort(rng.choice(n_total, size=min(take, n_total), replace=False))
        s = seeds[idx]
        t = time.perf_counter()
        _, acc = orc.ref_walk(img_c, s["pos"].astype(np.int64), s["label"].astype(np.uint8),
                              s["var"], nlabels, nbrw, sorw)
        dt = time.perf_counter() - t
        if dt > budget_s / 2 or take >= n_total:
            break
        take = int(take * max(2.0, min(8.0, (budget_s / 1.5) / max(dt, 1e-3))))
    steps_sample = idx.size * nbrw * sorw
    rate = steps_sample / dt
    return rate, idx.size, dt, orc.num_threads()


def time_pipeline(step, eng, warmup, steps, label):
    for w in range(warmup):
        t = time.perf_counter()
        out = step({})
        log("%s warmup %d: %.2fs" % (label, w, time.perf_counter() - t))
    eng.sync()
    tms = []
    t0 = time.perf_counter()
    for k in range(steps):
        tm = {}
        ts = time.perf_counter()
        out = step(tm)
        tms.append(tm)
        log("%s step %d: %.2fs  %s" % (label, k, time.perf_counter() - ts,
                                       " ".join("%s=%.3g" % kv for kv in sorted(tm.items())
                                                if isinstance(kv[1], (int, float)))))
    eng.sync()
    return out, tms, (time.perf_counter() - t0) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--nbrw", type=int, default=10)
    ap.add_argument("--sorw", type=int, default=4000)
    ap.add_argument("--smooth", type=int, default=100)
    ap.add_argument("--method", choices=["both", "plain", "guided"], default="both")
    ap.add_argument("--thr", type=float, default=0.9)
    ap.add_argument("--brad", type=int, default=4)
    ap.add_argument("--srad", type=int, default=6)
    ap.add_argument("--far", type=int, default=48)
    ap.add_argument("--far-walks", type=int, default=1)
    ap.add_argument("--mid", type=int, default=16, help="ring seeds within this radius walk --mid-walks; -1: none")
    ap.add_argument("--mid-walks", type=int, default=2)
    ap.add_argument("--repair", type=int, default=2, help="repair radius of the guided method; -1 disables")
    ap.add_argument("--repair-min", type=int, default=10)
    ap.add_argument("--repair-far", type=int, default=24)
    ap.add_argument("--repair-srad", type=int, default=3)
    ap.add_argument("--repair-far-walks", type=int, default=-1, help="-1: as --far-walks")
    ap.add_argument("--no-robust", action="store_true",
                    help="skip the second guided measurement with the robust preset")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-numpy", action="store_true",
                    help="skip the drop-in calls with NumPy arrays in and out (H2D and D2H included)")
    ap.add_argument("--no-reference-stream", action="store_true",
                    help="skip the Dice of the Philox walker against the reference stream at this size")
    ap.add_argument("--workload", choices=["512", "batch", "large", "post", "net"], default="512")
    ap.add_argument("--batch-n", type=int, default=64)
    ap.add_argument("--calibrate-valu", action="store_true",
                    help="also launch bmx_valu_calibrate once (256 Ki threads x 2000 x 128 float32 FMAs), the "
                         "known-count reference for VALU counters")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also launch bmx_f32_to_bf16 over 64 Mi floats (reads 256 MiB, writes "
                         "128 MiB) once, as the known-traffic reference for HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures a single MI355X (--gpus 1); sharding the walk over several GPUs "
              "is a design, not implemented", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "large":
        return bench_large(args)
    if args.workload == "post":
        return bench_post(args)
    if args.workload == "net":
        return bench_net(args)

    from biomedisa_amd import _hip, surrogate
    from biomedisa_amd.interpolation import (GUIDED_ROBUST, DeviceArray, smart_interpolation,
                                             smart_interpolation_device)

    from biomedisa_amd.synth import structures_volume
    t = time.perf_counter()
    n = args.size
    vol = structures_volume(shape=(n, n, n), seed=2, n_struct=8, radius=(5, 60), every=10, allaxis=True)
    log("volume %d^3 generated in %.1fs" % (n, time.perf_counter() - t))
    eng = _hip.engine(0)
    if args.weights:
        surrogate._nets[(0, surrogate.WEIGHTS_PATH)] = surrogate.net(0, args.weights)
    net = surrogate.net(0)
    # the caller's arrays, resident in HBM before any timed call
    img_d, lab_d = DeviceArray.upload(eng, vol.img), DeviceArray.upload(eng, vol.labels)
    eng.sync()
    n_vox = n ** 3
    common = dict(nbrw=args.nbrw, sorw=args.sorw, allaxis=True, smooth=args.smooth)
    guide = dict(thr=args.thr, brad=args.brad, srad=args.srad, far=args.far, far_walks=args.far_walks,
                 **guided_kw(args))
    robust = dict(GUIDED_ROBUST, thr=args.thr, brad=args.brad, far=args.far, repair_min=args.repair_min)
    methods = [("plain", "plain", None), ("guided", "guided", guide)]
    if args.method == "plain":
        methods = methods[:1]
    elif args.method == "guided":
        methods = methods[1:]
    if args.method != "plain" and not args.no_robust:
        methods.append(("robust", "guided", robust))

    def resident_step(method, g):
        def step(tm):
            r = smart_interpolation_device(img_d, lab_d, method=method, guide=g, timings=tm, **common)
            eng.sync()
            return r
        return step

    res = {}
    for name, method, g in methods:
        out, tms, sec = time_pipeline(resident_step(method, g), eng, args.warmup, args.steps, name)
        res[name] = dict(sec=sec, tms=tms, regular=out["regular"].download(), smooth=out["smooth"].download())
        del out
    if args.calibrate_copy or args.calibrate_valu:
        calibrate(eng, args)
    # the drop-in call itself, NumPy arrays in and out (H2D and D2H included)
    for name, method, g in ([] if args.no_numpy else methods[:2]):
        tm = {}
        t = time.perf_counter()
        r = smart_interpolation(vol.img, vol.labels, method=method, guide=g, timings=tm, **common)
        res[name]["numpy_sec"] = time.perf_counter() - t
        res[name]["numpy_identical"] = bool(np.array_equal(r["regular"], res[name]["regular"])
                                            and np.array_equal(r["smooth"], res[name]["smooth"]))
        log("%s, NumPy in and out: %.2fs (identical to the resident call: %s)"
            % (name, res[name]["numpy_sec"], res[name]["numpy_identical"]))
        del r
    img_d.free()

    main_key = "guided" if "guided" in res else "plain"
    r = res[main_key]
    config = {"workload": "%d^3 uint8, 8 ellipsoids/tubes, labels every 10th plane allaxis, "
                          "nbrw=%d sorw=%d smooth=%d" % (args.size, args.nbrw, args.sorw, args.smooth),
              "method": main_key,
              "timed_span": "smart_interpolation_device: the caller's uint8 image and label arrays resident in "
                            "HBM -> 'regular' and 'smooth' (Z,Y,X) volumes in HBM; label set and box, pad and "
                            "crop, seeds, walk, merge, smoothing, result placement and the Dice check inside",
              "seconds": r["sec"]}
    for name in res:
        tm = res[name]["tms"][-1]
        st = tm.get("stats", {})
        config["%s_seconds" % name] = res[name]["sec"]
        config["%s_phases" % name] = {k: v for k, v in tm.items() if k != "stats"}
        if "numpy_sec" in res[name]:
            config["%s_numpy_in_out_seconds" % name] = res[name]["numpy_sec"]
            config["%s_numpy_identical" % name] = res[name]["numpy_identical"]
        if name != "plain" and st.get("nseeds"):
            config["%s_walked_seed_fraction" % name] = st.get("walked_seeds", 0) / st["nseeds"]
            config["%s_ring_seed_fraction" % name] = st.get("ring_seeds", 0) / st["nseeds"]
            config["%s_repair_voxels" % name] = st.get("repair_voxels", 0)
            config["%s_fallback" % name] = st.get("fallback", False)
    config["guided_options"] = dict(guide, temperature=net.T)
    inner = (slice(1, -1),) * 3
    for name in res:
        config["min_dice_vs_ground_truth_%s" % name] = min(dice_by_confusion(
            np.ascontiguousarray(res[name]["regular"]), vol.gt).values())
    if "plain" in res:
        p = res["plain"]
        config.update(plain_voxels_per_s=n_vox / p["sec"])
        for name in [k for k in res if k != "plain"]:
            g = res[name]
            d = dice_by_confusion(g["regular"], p["regular"])
            ds = dice_by_confusion(g["smooth"], p["smooth"])
            pre = "" if name == "guided" else name + "_"
            config[pre + "speedup_vs_plain"] = p["sec"] / g["sec"]
            if "numpy_sec" in g and "numpy_sec" in p:
                config[pre + "speedup_vs_plain_numpy_in_out"] = p["numpy_sec"] / g["numpy_sec"]
            config[pre + "min_dice_guided_vs_plain"] = min(d.values())
            config[pre + "min_dice_smooth_guided_vs_plain"] = min(ds.values())
            config[pre + "min_dice_guided_vs_plain_regular_and_smooth"] = min(min(d.values()), min(ds.values()))
            config[pre + "dice_guided_vs_plain"] = {str(k): round(v, 5) for k, v in d.items()}
            config[pre + "dice_smooth_guided_vs_plain"] = {str(k): round(v, 5) for k, v in ds.items()}
    del inner

    line = {
        "metric": "segmented voxels/s at 512^3, 8 labels; surrogate-guided vs plain at Dice>=0.99",
        "value": n_vox / r["sec"],
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * r["sec"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": config,
    }
    # roofline of the dominant kernel: the walker of the plain pipeline (every seed walks)
    rk = res["plain"] if "plain" in res else res[main_key]
    tm = rk["tms"][-1]
    st = tm["stats"]
    if "plain" in res:
        steps_total = st["nseeds"] * args.nbrw * args.sorw
    else:
        steps_total = None
    launches = max(int(tm["walk_launches"]), 1)
    kernel_s = tm["walk_kernel_ms"] / 1000.0
    if steps_total is not None:
        walk_bytes = steps_total * 6 * 1 + st["accepted"] * 8
        achieved = walk_bytes / kernel_s / 1e9
        line["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args, launches),
                            "traffic_source": "profiles/k_tile_walk_512_traffic.md (rocprofv3 FETCH_SIZE + "
                                              "WRITE_SIZE, scaled by a copy-kernel calibration)",
                            "kernel": "k_tile_walk<uint8_t, 0, true> (persistent; rounds replayed in a HIP graph, "
                                      "kernel time = summed launch spans on the device wall clock)",
                            "walker_steps": steps_total, "steps_per_s": steps_total / kernel_s,
                            "accepted_moves": st["accepted"], "launches": launches,
                            "kernel_ms_total": tm["walk_kernel_ms"],
                            "launch_avg_ms": tm["walk_kernel_ms"] / launches,
                            "algorithmic_bytes_per_launch": walk_bytes / launches}
        valu = measured_valu(args)
        if valu is not None:
            line["roofline"]["valu"] = valu
    if not args.no_reference_stream and "plain" in res:
        dref, ref_s = reference_stream_dice(eng, vol, res["plain"]["regular"], args)
        config["min_dice_philox_vs_reference_stream"] = min(dref.values())
        config["dice_philox_vs_reference_stream"] = {str(k): round(v, 5) for k, v in dref.items()}
        config["reference_stream_walk_seconds"] = ref_s
    if not args.no_cpu_baseline:
        _, img_c, labidx, _, nol = prepare(args, vol)
        from biomedisa_amd.interpolation import DeviceVolume
        dv = DeviceVolume(eng, img_c, labidx)
        seeds_buf, ns = eng.seeds(dv.lab, dv.img, dv.img_dtype, dv.dims, eng.plane_flags(dv.lab, dv.dims, True))
        seeds = seeds_buf.download((ns,), _hip.SEED_DTYPE)
        rate, nsample, dt, cores = cpu_baseline(img_c, seeds, args.nbrw, args.sorw, nol, args.cpu_budget)
        walk_s = ns * args.nbrw * args.sorw / rate
        line["cpu_baseline"] = {
            "value": n_vox / walk_s, "unit": "voxels/s", "cores": cores, "kind": "port",
            "sample": "oracle ref_walk (C/OpenMP restatement of pyopencl_small.py randomWalk, "
                      "float MRG stream) on %d random seeds of this workload (%.1fs, %.3g steps/s); "
                      "full plain walk extrapolated to %.0fs; smoothing/merge not counted"
                      % (nsample, dt, rate, walk_s)}
    print(json.dumps(line), flush=True)


def calibrate(eng, args):
    """Known-work kernels for scaling counters read in the same run: a copy of
    64 Mi floats to bf16 (256 MiB read, 128 MiB written) and an FMA-chain
    kernel of 262144 threads x 2000 x 128 float32 FMAs."""
    import ctypes
    from biomedisa_amd import _hip
    if args.calibrate_copy:
        n_cal = 64 << 20
        src, dst = eng.alloc(n_cal * 4), eng.alloc(n_cal * 2)
        src.zero()
        _hip.call("bmx_f32_to_bf16", eng.ctx, src.ptr, dst.ptr, n_cal)
        eng.sync()
        log("calibration copy: %d floats -> bf16" % n_cal)
    if args.calibrate_valu:
        threads, iters = 1 << 18, 2000
        out = eng.alloc(threads * 4)
        ms = ctypes.c_float()
        _hip.call("bmx_valu_calibrate", eng.ctx, out.ptr, threads, iters, ctypes.byref(ms))
        fma = threads * iters * 128
        log("calibration FMA chains: %.4g FMAs in %.3f ms = %.2f TFLOP/s, %.4g wave-level instructions"
            % (fma, ms.value, 2 * fma / (ms.value * 1e-3) / 1e12, fma / 64))


def reference_stream_dice(eng, vol, plain_regular, args, chunk=1 << 22):
    """Per-label Dice between the Philox walker's 'regular' result and the
    reference stream's (the float MRG recurrence of pyopencl_small.py:213-282,
    one thread per seed, all walks in sequence) on the same volume; the seeds
    are walked in chunks (each seed's stream is its own, so chunks change
    nothing) so that no launch runs long."""
    from biomedisa_amd import _hip
    from biomedisa_amd.interpolation import _prepare_device
    prep = _prepare_device(eng, vol.img, vol.labels, args.nbrw, args.sorw, False, None, True, "none", "all", 0,
                           False, False)
    dv = prep.vol
    flags = eng.plane_flags(dv.lab, dv.dims, True)
    seeds, ns = eng.seeds(dv.lab, dv.img, dv.img_dtype, dv.dims, flags)
    hits = eng.alloc(prep.bm.nol * dv.V * 4)
    hits.zero()
    t = time.perf_counter()
    for c0 in range(0, ns, chunk):
        c1 = min(ns, c0 + chunk)
        view = _hip.BufferView(seeds, c0 * _hip.SEED_DTYPE.itemsize, (c1 - c0) * _hip.SEED_DTYPE.itemsize)
        eng.walk(dv.img, dv.img_dtype, dv.dims, view, c1 - c0, prep.bm.nol, args.nbrw, args.sorw, _hip.RNG_REF, 0,
                 hits)
        eng.sync()
        log("reference stream: %d / %d seeds walked (%.1fs)" % (c1, ns, time.perf_counter() - t))
    walk_s = time.perf_counter() - t
    reg = eng.alloc(dv.V)
    eng.argmax(hits, prep.bm.nol, dv.V, prep.values, reg)
    hits.free()
    full = eng.alloc(int(np.prod(prep.shape)))
    eng.place(reg, prep.shape, prep.box, full)
    ref_regular = full.download(prep.shape, np.uint8)
    return dice_by_confusion(ref_regular, plain_regular), walk_s


def guided_kw(args):
    """Inner-ring and repair options of the guided pipeline from the command line."""
    return dict(mid=None if args.mid < 0 else args.mid, mid_walks=args.mid_walks,
                repair=None if args.repair < 0 else args.repair, repair_min=args.repair_min,
                repair_far=args.repair_far, repair_srad=args.repair_srad,
                repair_far_walks=None if args.repair_far_walks < 0 else args.repair_far_walks)


def bench_batch(args):
    """The batched plain method: one bmx_walk_batch over every volume."""
    from biomedisa_amd.synth import batch_volumes
    t = time.perf_counter()
    vols = batch_volumes(seed=3, n=args.batch_n, workers=8)
    log("generated %d volumes in %.1fs" % (len(vols), time.perf_counter() - t))
    from biomedisa_amd import _hip
    from biomedisa_amd.autotune import tune_walk
    from biomedisa_amd.batch import DeviceBatch, run_batch
    from biomedisa_amd.interpolation import DeviceVolume, _prepare, run_pipeline
    preps = [_prepare(v.img, v.labels, args.nbrw, args.sorw, False, None, False, "none", "all", 0, False,
                      False) for v in vols]
    assert all(p is not None for p in preps)
    eng = _hip.engine(0)
    batch = DeviceBatch(eng, [(p.img, p.labidx, p.bm.nol, False) for p in preps])
    values = [p.values for p in preps]
    n_vox = sum(int(np.prod(v.img.shape)) for v in vols)
    log("batch: %d volumes, %d voxels (%d in label boxes), %d seeds"
        % (len(vols), n_vox, batch.total_voxels, sum(batch.nseeds)))
    t = time.perf_counter()
    opts = tune_walk(batch, args.nbrw, args.sorw, 0x5EED0B10D1CE0001, log=log)
    tune_s = time.perf_counter() - t
    log("autotune: %s in %.1fs" % (opts, tune_s))
    hits = eng.alloc(batch.hits_elems * 4)

    def step(tm, o=opts):
        return run_batch(batch, values, args.nbrw, args.sorw, 0, walk_opts=o, timings=tm, hits=hits)
    out, tms, sec = time_pipeline(step, eng, args.warmup, args.steps, "batch")
    labels = [o["regular"].download(p.img.shape, np.uint8) for o, p in zip(out, preps)]
    _, tms_def, sec_def = time_pipeline(lambda tm: step(tm, {}), eng, 0, 1, "batch default shape")
    # the same volumes one pipeline call at a time (what looping over smart_interpolation runs)
    t = time.perf_counter()
    same = True
    for p, lab in zip(preps, labels):
        dv = DeviceVolume(eng, p.img, p.labidx)
        o = run_pipeline(dv, p.bm.nol, p.values, nbrw=args.nbrw, sorw=args.sorw, allaxis=False, smooth=0)
        same &= bool(np.array_equal(o["regular"].download(p.img.shape, np.uint8), lab))
        o["regular"].free()
        dv.img.free()
        dv.lab.free()
    eng.sync()
    loop_s = time.perf_counter() - t
    log("per-volume loop: %.2fs, identical labels: %s" % (loop_s, same))
    tm = tms[-1]
    steps_total = sum(batch.nseeds) * args.nbrw * args.sorw
    walk_bytes = steps_total * 6 + tm["accepted"] * 8
    launches = max(int(tm["walk_launches"]), 1)
    kernel_s = tm["walk_kernel_ms"] / 1000.0
    achieved = walk_bytes / (tm["walk_phase_ms"] / 1000.0) / 1e9
    line = {
        "metric": "segmented voxels/s, 64-volume batch, plain walker",
        "value": n_vox / sec,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * sec,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "%d uint8 volumes 96^3..256x320x320, 2-12 structures, labels every 6-12th XY "
                               "plane, nbrw=%d sorw=%d smooth=0" % (len(vols), args.nbrw, args.sorw),
                   "voxels": n_vox, "seeds": int(sum(batch.nseeds)), "walk_options": opts,
                   "autotune_seconds": tune_s, "autotune_table": [[o, s] for o, s in tune_walk.last_table],
                   "default_options_seconds": sec_def, "per_volume_loop_seconds": loop_s,
                   "batched_vs_loop": loop_s / sec, "identical_to_per_volume_calls": same,
                   "timings": tm},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_tile_walk<uint8_t> (walk phase, graph-driven rounds time as one)",
                     "walker_steps": steps_total, "accepted_moves": tm["accepted"], "launches": launches,
                     "walk_kernel_ms": tm["walk_kernel_ms"], "walk_phase_ms": tm["walk_phase_ms"]},
    }
    hits.free()
    print(json.dumps(line), flush=True)


def bench_post(args):
    """Post-processing of the 512^3 plain result: clean (0.1), fill (0.9), acwe
    (alpha 1, smooth 1, 3 steps) and refined on the device, each timed once after
    a warm-up; the oracle's CPU restatements timed on a 256^3 corner of the same
    volumes."""
    from biomedisa_amd import _hip
    from biomedisa_amd import post as P
    from biomedisa_amd.interpolation import DeviceVolume, run_pipeline, post_process
    vol, img_c, labidx, values, nol = prepare(args)
    eng = _hip.engine(0)
    dv = DeviceVolume(eng, img_c, labidx)
    out = run_pipeline(dv, nol, values, nbrw=args.nbrw, sorw=args.sorw, allaxis=True, smooth=0)
    regular = out["regular"].download(dv.dims, np.uint8)[1:-1, 1:-1, 1:-1]
    data = img_c[1:-1, 1:-1, 1:-1]
    log("plain result %s" % (regular.shape,))
    steps = {
        "clean": lambda: post_process({"regular": regular}, clean=0.1)["cleaned"],
        "fill": lambda: post_process({"regular": regular}, fill=0.9)["filled"],
        "acwe": lambda: P.active_contour(data, regular, 1.0, 1, 3),
        "refined": lambda: P.refinement(data, regular),
    }
    gpu = {}
    for name, f in steps.items():
        f()
        eng.sync()
        t = time.perf_counter()
        r = f()
        eng.sync()
        gpu[name] = time.perf_counter() - t
        log("%s on the GPU: %.3fs (%d voxels changed)" % (name, gpu[name], int((r != regular).sum())))
    import oracle.post as OP
    crop = (slice(0, 256),) * 3
    rc, dc = np.ascontiguousarray(regular[crop]), np.ascontiguousarray(data[crop])
    cpu_steps = {"clean": lambda: OP.clean(rc, 0.1), "fill": lambda: OP.fill(rc, 0.9),
                 "acwe": lambda: OP.acwe(dc, rc, 1.0, 1, 3), "refined": lambda: OP.refinement(dc, rc)}
    cpu = {}
    for name, f in cpu_steps.items():
        t = time.perf_counter()
        f()
        cpu[name] = time.perf_counter() - t
        log("%s, oracle on 256^3: %.2fs" % (name, cpu[name]))
    n_vox = int(np.prod(regular.shape))
    line = {"metric": "post-processing voxels/s at 512^3 (clean, fill, acwe, refined)",
            "value": n_vox / sum(gpu.values()), "unit": "voxels/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": 1000.0 * sum(gpu.values()), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u8", "data": "synthetic",
            "config": {"workload": "512^3 plain result of the benchmark volume", "gpu_seconds": gpu,
                       "oracle_seconds_256cube": cpu,
                       "oracle_voxels_per_s": {k: 256 ** 3 / v for k, v in cpu.items()},
                       "gpu_voxels_per_s": {k: n_vox / v for k, v in gpu.items()}}}
    from oracle import orc
    line["cpu_baseline"] = {"value": 256 ** 3 / sum(cpu.values()), "unit": "voxels/s",
                            "cores": int(orc.lib().orc_num_threads()), "kind": "port",
                            "sample": "the oracle's C/NumPy restatements of clean, fill, acwe and refined "
                                      "on the 256^3 corner of the same volumes"}
    print(json.dumps(line), flush=True)


def bench_net(args):
    """The learned proposal's network on its own: the forward over the 512^3
    benchmark's proposal grid (the cropped 514^3 volume padded to 520^3, one
    slab, 32 input channels of random bf16 values), and training steps (forward,
    softmax cross-entropy, backward, Adam) at batch 8 x 72^3, each timed after a
    warm-up.  FLOPs count the 3x3x3 convolutions and the head (Net.flops_forward);
    a training step does the forward, the weight gradient of every convolution
    and the input gradient of every convolution but the first, i.e. about 3x."""
    from biomedisa_amd import _hip, surrogate
    eng = _hip.engine(0)
    net = surrogate.net(0)
    rng = np.random.default_rng(0)
    peak = 2500.0   # TFLOP/s, dense bf16 MFMA of one MI355X

    def rand_bf16(n):
        return (rng.standard_normal(n, dtype=np.float32).view(np.uint32) >> 16).astype(np.uint16)

    g = (520, 520, 520)
    x = eng.upload(rand_bf16(int(np.prod(g)) * surrogate.CIN))
    net.forward(x, 1, *g)
    eng.sync()
    fwd = []
    for _ in range(max(args.steps, 3)):
        t = time.perf_counter()
        net.forward(x, 1, *g)
        eng.sync()
        fwd.append(time.perf_counter() - t)
    fwd_s = min(fwd)
    f_fwd = net.flops_forward(1, *g)
    log("forward %s: %.2f ms, %.1f TFLOP/s" % (g, 1e3 * fwd_s, f_fwd / fwd_s / 1e12))
    del x
    B, tg = 8, (72, 72, 72)
    ng = int(np.prod(tg))
    tnet = surrogate.Net(eng, seed=1)
    xb = eng.upload(rand_bf16(B * ng * surrogate.CIN))
    ib = eng.upload(rng.integers(0, surrogate.L_MAX, size=B * ng).astype(np.uint8))
    mb = eng.upload(np.ones(B * ng, np.uint8))
    lossb = eng.alloc(4)

    def train_step():
        tnet.zero_grad()
        lossb.zero()
        tnet.loss_and_grad(xb, None, mb, B, *tg, lossb, target_idx=ib)
        tnet.adam(1e-3)

    for _ in range(2):
        train_step()
    eng.sync()
    it = max(args.steps, 10)
    t = time.perf_counter()
    for _ in range(it):
        train_step()
    eng.sync()
    tr_s = (time.perf_counter() - t) / it
    v0 = B * ng
    conv = sum(2.0 * 27 * ci * co * v0 / {"enc0": 1, "enc1": 8, "enc2": 64, "bott": 512, "dec2": 64,
                                           "dec1": 8, "dec0": 1}[nm] for nm, ci, co in surrogate.LAYERS)
    dgrad = conv - 2.0 * 27 * 32 * 16 * v0
    f_tr = tnet.flops_forward(B, *tg) + conv + dgrad + 2 * 2.0 * 16 * surrogate.L_MAX * v0
    log("training step %d x %s: %.2f ms, %.1f TFLOP/s" % (B, tg, 1e3 * tr_s, f_tr / tr_s / 1e12))
    tf = f_fwd / fwd_s / 1e12
    line = {"metric": "surrogate network forward TFLOP/s on the 512^3 proposal grid",
            "value": tf, "unit": "TFLOP/s", "n_gpus": 1, "steps": len(fwd), "warmup": 1,
            "ms_per_step": 1e3 * fwd_s, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16", "data": "synthetic",
            "config": {"workload": "3-D conv encoder-decoder (16/32/64/128 channels, 32 inputs, 30 label "
                                   "slots), forward on 1 x 520^3; training at batch 8 x 72^3",
                       "forward_flop": f_fwd, "forward_seconds": fwd_s, "forward_frac_of_bf16_peak": tf / peak,
                       "train_step_flop": f_tr, "train_step_seconds": tr_s,
                       "train_tflops": f_tr / tr_s / 1e12, "train_frac_of_bf16_peak": f_tr / tr_s / 1e12 / peak},
            "roofline": {"bound": "mfma", "achieved": tf, "peak": peak, "unit": "TFLOP/s", "frac": tf / peak,
                         "traffic": None}}
    print(json.dumps(line), flush=True)


def dice_by_confusion(a, b, nbins=256):
    """Per-label Dice of two uint8 label volumes from one confusion matrix,
    accumulated over z-slabs ({label >= 1 present in either: dice})."""
    cm = np.zeros((nbins, nbins), dtype=np.int64)
    for z in range(0, a.shape[0], 64):
        idx = a[z:z + 64].ravel().astype(np.int32) * nbins + b[z:z + 64].ravel()
        cm += np.bincount(idx, minlength=nbins * nbins).reshape(nbins, nbins)
    out = {}
    for l in range(1, nbins):
        s = int(cm[l, :].sum() + cm[:, l].sum())
        if s:
            out[l] = 2.0 * int(cm[l, l]) / s
    return out


def bench_large(args):
    """1024^3 uint16: plain vs surrogate-guided, one step each."""
    import gc
    from biomedisa_amd import _hip, surrogate
    from biomedisa_amd import helpers as H
    from biomedisa_amd.interpolation import DeviceVolume, run_pipeline, run_pipeline_guided
    from biomedisa_amd.synth import large_volume
    t = time.perf_counter()
    vol = large_volume(seed=5, size=1024, log=log)
    log("volume generated in %.0fs" % (time.perf_counter() - t))
    bm = H.Biomedisa()
    bm.data, bm.labelData, bm.only, bm.ignore, bm.success = vol.img, vol.labels, "all", "none", True
    bm = H.pre_processing(bm)
    nol = bm.nol
    img = np.pad(H.normalize_image(bm.data), 1)
    lab = np.pad(bm.labelData, 1)
    z0, z1, y0, y1, x0, x1 = H.predict_blocksize(lab)
    img = np.ascontiguousarray(img[z0:z1, y0:y1, x0:x1])
    labidx = H.label_index_volume(np.ascontiguousarray(lab[z0:z1, y0:y1, x0:x1]), bm.allLabels)
    values = bm.allLabels.astype(np.uint8)
    gt = vol.gt
    del vol, bm, lab
    gc.collect()
    eng = _hip.engine(0)
    dv = DeviceVolume(eng, img, labidx)
    del img, labidx
    gc.collect()
    n_vox = 1024 ** 3
    log("uploaded %s float32 image, %d labels" % (dv.dims, nol))
    res = {}
    for method in ("plain", "guided", "plain_large_path"):
        tm = {}
        t = time.perf_counter()
        if method == "plain":
            out = run_pipeline(dv, nol, values, nbrw=args.nbrw, sorw=args.sorw, allaxis=True, smooth=0, timings=tm)
        elif method == "plain_large_path":
            # the reference's own driver at this size (adaptive walks)
            out = run_pipeline(dv, nol, values, nbrw=args.nbrw, sorw=args.sorw, allaxis=True, smooth=0, timings=tm,
                               path="large")
        else:
            net = surrogate.net(0)
            out = run_pipeline_guided(dv, nol, values, net, nbrw=args.nbrw, sorw=args.sorw, allaxis=True, smooth=0,
                                      thr=args.thr, brad=args.brad, srad=args.srad, far=args.far,
                                      far_walks=args.far_walks, timings=tm, **guided_kw(args))
        eng.sync()
        sec = time.perf_counter() - t
        log("%s: %.1fs %s" % (method, sec, {k: round(v, 3) for k, v in tm.items()}))
        res[method] = dict(sec=sec, tm=tm, nseeds=out["nseeds"], walked=out.get("walked_seeds"),
                           ring=out.get("ring_seeds"), mid_ring=out.get("mid_ring_seeds"),
                           repair=out.get("repair_seeds"), accepted=out["accepted"],
                           regular=out["regular"].download(dv.dims, np.uint8))
        out["regular"].free()
        del out
        gc.collect()
    p, g, pl = res["plain"], res["guided"], res["plain_large_path"]
    d = dice_by_confusion(g["regular"], p["regular"])
    d_large = dice_by_confusion(pl["regular"], p["regular"])
    inner = (slice(1, -1),) * 3
    d_gt = dice_by_confusion(np.ascontiguousarray(p["regular"][inner]), gt)
    log("Dice computed")
    line = {
        "metric": "segmented voxels/s at 1024^3 uint16, 16 structures; surrogate-guided vs plain",
        "value": n_vox / g["sec"],
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": 1000.0 * g["sec"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "1024^3 uint16 (12-bit, normalised to float32), 16 ellipsoids/tubes, noise 40, "
                               "bias 200, labels every 16th plane allaxis, nbrw=%d sorw=%d smooth=0"
                               % (args.nbrw, args.sorw),
                   "labels": nol, "seeds": p["nseeds"], "plain_seconds": p["sec"], "guided_seconds": g["sec"],
                   "speedup_vs_plain": p["sec"] / g["sec"], "walked_seed_fraction": g["walked"] / g["nseeds"],
                   "ring_seed_fraction": g["ring"] / g["nseeds"],
                   "mid_ring_seed_fraction": g["mid_ring"] / g["nseeds"],
                   "repair_seed_fraction": (g["repair"] or 0) / g["nseeds"],
                   "min_dice_guided_vs_plain": min(d.values()),
                   "dice_guided_vs_plain": {str(k): round(v, 5) for k, v in d.items()},
                   "min_dice_vs_ground_truth_plain": min(d_gt.values()),
                   "plain_timings": p["tm"], "guided_timings": g["tm"],
                   "large_path": {"seconds": pl["sec"], "walked_seed_fraction": pl["walked"] / pl["nseeds"],
                                  "min_dice_vs_plain": min(d_large.values()),
                                  "speedup_vs_plain": p["sec"] / pl["sec"], "timings": pl["tm"]},
                   "hit_volumes_resident_GB": nol * int(np.prod(dv.dims)) * 4 / 1e9},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
