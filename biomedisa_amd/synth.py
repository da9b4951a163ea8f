"""Seeded synthetic volumes for the smart-interpolation workloads.

Every volume comes from numpy.random.Generator(Philox(seed)): a background,
ellipsoids and tubes with their own mean intensities (later structures
overwrite earlier ones; the ground-truth label is the structure number), a
linear bias field along a random unit vector, Gaussian noise, clipped and
rounded.  Sparse labels keep the ground truth on every k-th plane.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class Volume:
    img: np.ndarray      # (Z,Y,X) uint8 / uint16
    gt: np.ndarray       # (Z,Y,X) uint8 ground truth, 0 = background
    labels: np.ndarray   # (Z,Y,X) uint8 sparse labels (gt on labelled planes)
    allaxis: bool


def _rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def _rotation(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q *= np.sign(np.diag(r))
    return q


def _bbox(center, radius, shape):
    lo = [max(0, int(np.floor(c - radius))) for c in center]
    hi = [min(n, int(np.ceil(c + radius)) + 1) for c, n in zip(center, shape)]
    return lo, hi


def _ellipsoid(shape, rng, semi_lo, semi_hi):
    axes = rng.uniform(semi_lo, semi_hi, 3)
    rmax = axes.max()
    center = [rng.uniform(rmax + 1, n - rmax - 2) if n > 2 * rmax + 3 else n / 2 for n in shape]
    R = _rotation(rng)
    lo, hi = _bbox(center, rmax, shape)
    zz, yy, xx = np.meshgrid(*[np.arange(l, h, dtype=np.float32) - c
                               for l, h, c in zip(lo, hi, center)], indexing="ij")
    p = np.stack([zz, yy, xx], -1) @ R.astype(np.float32)
    inside = ((p / axes.astype(np.float32)) ** 2).sum(-1) <= 1.0
    return (slice(lo[0], hi[0]), slice(lo[1], hi[1]), slice(lo[2], hi[2])), inside


def _tube(shape, rng, r_lo, r_hi, nseg=3):
    radius = rng.uniform(r_lo, r_hi)
    n = np.array(shape, dtype=np.float64)
    pts = [rng.uniform(radius + 1, n - radius - 2)]
    for _ in range(nseg):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        length = rng.uniform(0.2, 0.5) * n.min()
        nxt = np.clip(pts[-1] + d * length, radius + 1, n - radius - 2)
        pts.append(nxt)
    pts = np.array(pts)
    lo = np.maximum(0, np.floor(pts.min(0) - radius)).astype(int)
    hi = np.minimum(n, np.ceil(pts.max(0) + radius) + 1).astype(int)
    zz, yy, xx = np.meshgrid(*[np.arange(l, h, dtype=np.float32) for l, h in zip(lo, hi)],
                             indexing="ij")
    P = np.stack([zz, yy, xx], -1)
    dmin = np.full(P.shape[:3], np.inf, dtype=np.float32)
    for a, b in zip(pts[:-1], pts[1:]):
        a = a.astype(np.float32)
        ab = (b - a).astype(np.float32)
        t = np.clip(((P - a) @ ab) / max(float(ab @ ab), 1e-6), 0.0, 1.0)
        d = np.linalg.norm(P - a - t[..., None] * ab, axis=-1)
        np.minimum(dmin, d, out=dmin)
    return (slice(lo[0], hi[0]), slice(lo[1], hi[1]), slice(lo[2], hi[2])), dmin <= radius


def _finish(img, shape, rng, bias, sigma, max_value, dtype):
    u = rng.standard_normal(3)
    u /= np.linalg.norm(u)
    c = [(s - 1) / 2.0 for s in shape]
    ramp = (u[0] * (np.arange(shape[0], dtype=np.float32) - c[0]))[:, None, None] \
        + (u[1] * (np.arange(shape[1], dtype=np.float32) - c[1]))[None, :, None] \
        + (u[2] * (np.arange(shape[2], dtype=np.float32) - c[2]))[None, None, :]
    ramp /= max(float(np.abs(ramp).max()), 1e-6)
    img += np.float32(bias) * ramp
    del ramp
    img += np.float32(sigma) * rng.standard_normal(shape, dtype=np.float32)
    np.clip(img, 0, max_value, out=img)
    return np.rint(img).astype(dtype)


def sparse_labels(gt, every, allaxis, offset=0):
    """Ground truth on planes k % every == offset (z only, or all three axes)."""
    lab = np.zeros_like(gt)
    lab[offset::every] = gt[offset::every]
    if allaxis:
        lab[:, offset::every] = gt[:, offset::every]
        lab[:, :, offset::every] = gt[:, :, offset::every]
    return lab


def ellipsoids_volume(shape=(128, 128, 128), seed=1, means=(60, 120, 190), background=20,
                      semi=(15, 40), sigma=15.0, bias=10.0, every=8, allaxis=False):
    """Ellipsoid phantom of the 128^3 CPU-reference workload (3 objects, z-planes every 8th)."""
    rng = _rng(seed)
    img = np.full(shape, background, dtype=np.float32)
    gt = np.zeros(shape, dtype=np.uint8)
    for k, m in enumerate(means, 1):
        sl, inside = _ellipsoid(shape, rng, *semi)
        img[sl][inside] = m
        gt[sl][inside] = k
    img = _finish(img, shape, rng, bias, sigma, 255, np.uint8)
    return Volume(img, gt, sparse_labels(gt, every, allaxis), allaxis)


def structures_volume(shape=(512, 512, 512), seed=2, n_struct=8, radius=(5, 60), mean_range=(30, 230),
                      background=None, sigma=12.0, bias=15.0, every=10, allaxis=True, dtype=np.uint8,
                      max_value=255, log=None):
    """Overlapping ellipsoids and tubes (half each), distinct means spread over mean_range.
    log: optional callable, told of each structure drawn (large volumes take minutes)."""
    rng = _rng(seed)
    lo, hi = mean_range
    means = np.linspace(lo, hi, n_struct + 1)
    order = rng.permutation(n_struct + 1)
    bg = means[order[0]] if background is None else background
    struct_means = means[order[1:]]
    img = np.full(shape, bg, dtype=np.float32)
    gt = np.zeros(shape, dtype=np.uint8)
    r_lo, r_hi = radius
    for k in range(n_struct):
        if k % 2 == 0:
            sl, inside = _ellipsoid(shape, rng, max(r_lo, 3), r_hi)
        else:
            sl, inside = _tube(shape, rng, r_lo, min(r_hi, 0.25 * min(shape)))
        img[sl][inside] = struct_means[k]
        gt[sl][inside] = k + 1
        if log:
            log("structure %d/%d drawn" % (k + 1, n_struct))
    img = _finish(img, shape, rng, bias, sigma, max_value, dtype)
    if log:
        log("bias field and noise added")
    return Volume(img, gt, sparse_labels(gt, every, allaxis), allaxis)



def large_volume(seed=5, size=1024, every=16, log=None):
    """Large-volume workload: size^3 uint16 with 12-bit values, 16 overlapping
    ellipsoids and tubes (radius 10-120 voxels at 1024^3, scaled with size), means
    spread over 480..3680, noise sigma 40, bias field +-200, labels on every
    `every`-th plane of all three axes."""
    f = size / 1024.0
    return structures_volume(shape=(size, size, size), seed=seed, n_struct=16,
                             radius=(max(3, int(round(10 * f))), max(6, int(round(120 * f)))),
                             mean_range=(480, 3680), sigma=40.0, bias=200.0, every=every, allaxis=True,
                             dtype=np.uint16, max_value=4095, log=log)


def _batch_member(args):
    shape, seed, n_struct, sigma, every = args
    return structures_volume(shape=shape, seed=seed, n_struct=n_struct, radius=(5, max(8, min(shape) // 4)),
                             mean_range=(30, 230), sigma=sigma, bias=15.0, every=every, allaxis=False)


def batch_volumes(seed=3, n=64, size_lo=96, size_hi=(256, 320, 320), workers=1):
    """Batched-interpolation workload: n uint8 volumes, shape per axis uniform in
    [size_lo, size_hi] (z up to 256, y and x up to 320), each with 2-12
    structures (ellipsoids and tubes, radius 5 .. a quarter of the smallest
    side), noise sigma uniform in [8, 20] and labels on every k-th XY plane,
    k uniform in [6, 12].  workers > 1 generates them in that many processes
    (same volumes)."""
    rng = _rng(seed)
    jobs = []
    for i in range(n):
        shape = tuple(int(rng.integers(size_lo, hi + 1)) for hi in size_hi)
        n_struct = int(rng.integers(2, 13))
        sigma = float(rng.uniform(8, 20))
        every = int(rng.integers(6, 13))
        jobs.append((shape, int(seed) * 1000 + i, n_struct, sigma, every))
    if workers <= 1:
        return [_batch_member(j) for j in jobs]
    import multiprocessing as mp
    with mp.get_context("spawn").Pool(workers) as pool:
        return pool.map(_batch_member, jobs)
