"""Pins of the CPU oracle (oracle/bmx_oracle.c) against independent facts.

Each test checks the oracle against something that does not come from the
oracle itself: published known-answer vectors, a float64 exp, brute-force
restatements of the reference's NumPy code, closed forms of the transition
law and of the curvature stencil, conservation laws.
"""
import numpy as np
import pytest

import oracle
from oracle import orc


# --- Philox4x32-10: Random123 known-answer vectors (kat_vectors, philox4x32_10) ---
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    out = orc.philox4x32_10(np.array([ctr], np.uint32), np.array(key, np.uint32))[0]
    assert tuple(int(v) for v in out) == expect


def test_philox_planted_mistake_fails():
    """A Philox with one round too few must miss the vectors (the KAT has teeth)."""
    def philox9(c, k):
        c = list(c); k = list(k)
        for _ in range(9):
            p0 = 0xD2511F53 * c[0]; p1 = 0xCD9E8D57 * c[2]
            c = [((p1 >> 32) ^ c[1] ^ k[0]) & 0xffffffff, p1 & 0xffffffff,
                 ((p0 >> 32) ^ c[3] ^ k[1]) & 0xffffffff, p0 & 0xffffffff]
            k = [(k[0] + 0x9E3779B9) & 0xffffffff, (k[1] + 0xBB67AE85) & 0xffffffff]
        return tuple(c)
    assert philox9(KAT[2][0], KAT[2][1]) != KAT[2][2]


# --- exp -----------------------------------------------------------------------
def test_expf_within_2_2_ulp():
    """Degree-5 Cody-Waite exp: within 2.2 ulp of float64 exp over the weights'
    range, including every -t^2/2 of uint8 differences; a planted coefficient
    error of 1e-5 relative in c2 must exceed the bound."""
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-86.5, 0, 2_000_000), np.linspace(-86.5, 0, 100_001),
                        -np.arange(0, 65026, dtype=np.float64) / 2.0]).astype(np.float32)
    x = x[x >= -86.5]
    y = orc.expf(x).astype(np.float64)
    r = np.exp(x.astype(np.float64))
    ulp = np.spacing(r.astype(np.float32)).astype(np.float64)
    err = np.abs(y - r) / ulp
    assert np.max(err) <= 2.2
    assert np.mean(err) < 0.75
    # the same method with c2 perturbed (float64 evaluation of the reduced polynomial)
    t = x.astype(np.float64) * 1.4426950408889634
    n = np.rint(t)
    rr = x - n * np.log(2)
    c = [1.0, 1.0, 0.499992311 * (1 + 1e-5), 0.166671142, 0.0418901145, 0.00831252523]
    p = sum(ck * rr ** k for k, ck in enumerate(c))
    bad = np.abs(p * 2.0 ** n - r) / ulp
    assert np.max(bad) > 2.2


def test_expf_special_values():
    y = orc.expf(np.array([0.0, -0.0, -1000.0, -104.0, -88.0], np.float32))
    assert y[0] == 1.0 and y[1] == 1.0
    assert y[2] == 0.0 and y[3] == 0.0               # below the smallest denormal
    assert 0.0 < y[4] < np.finfo(np.float32).tiny     # gradual underflow (ldexp)
    np.testing.assert_allclose(y[4], np.exp(-88.0), rtol=1e-3)


# --- labelled planes -------------------------------------------------------------
def _ref_allx_flags_numpy(arr):
    """read_indices_allx (biomedisa_helper.py:1047-1066), vectorised as the reference writes it."""
    gradient = np.zeros(arr.shape, dtype=np.int8)
    ones = np.zeros_like(gradient)
    ones[arr != 0] = 1
    tmp = np.abs(ones[:, :-1] - ones[:, 1:])
    gradient[:, :-1] += tmp
    gradient[:, 1:] += tmp
    ones[gradient == 2] = 0
    gradient.fill(0)
    tmp = np.abs(ones[:, :, :-1] - ones[:, :, 1:])
    gradient[:, :, :-1] += tmp
    gradient[:, :, 1:] += tmp
    ones[gradient == 2] = 0
    return np.array([np.any(s) for s in ones], np.uint8)


def _ref_flags(lab, allaxis):
    if not allaxis:
        return (np.array([np.any(s) for s in lab], np.uint8), np.zeros(lab.shape[1], np.uint8),
                np.zeros(lab.shape[2], np.uint8))
    fz = _ref_allx_flags_numpy(lab)
    t = np.ascontiguousarray(np.swapaxes(lab, 0, 1))
    fy = _ref_allx_flags_numpy(t)
    t = np.ascontiguousarray(np.swapaxes(t, 0, 2))
    fx = _ref_allx_flags_numpy(t)
    return fz, fy, fx


def _random_sparse(shape, rng, every, allaxis, nlab=3, thick=False):
    Z, Y, X = shape
    gt = rng.integers(0, nlab + 1, size=shape).astype(np.uint8)
    lab = np.zeros_like(gt)
    off = [int(v) for v in rng.integers(0, every, 3)]
    lab[off[0]::every] = gt[off[0]::every]
    if allaxis:
        lab[:, off[1]::every] = gt[:, off[1]::every]
        lab[:, :, off[2]::every] = gt[:, :, off[2]::every]
    return lab


@pytest.mark.parametrize("allaxis", [False, True])
@pytest.mark.parametrize("trial", range(6))
def test_plane_flags_vs_reference_numpy(allaxis, trial):
    rng = np.random.default_rng(100 + trial)
    shape = tuple(int(v) for v in rng.integers(6, 20, 3))
    lab = _random_sparse(shape, rng, int(rng.integers(3, 6)), allaxis)
    # sprinkle negative labels (ignored regions, non-zero for plane detection)
    neg = rng.random(shape) < 0.01
    labidx = lab.copy()
    labidx[neg & (lab > 0)] = oracle.IGN
    got = orc.plane_flags(labidx, allaxis)
    ref = _ref_flags((labidx != 0).astype(np.int32), allaxis)
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)


def test_plane_flags_thin_line_rule_hand_cases():
    lab = np.zeros((7, 7, 7), np.uint8)
    lab[:, 3, :] = 1                 # labelled XZ plane y=3: a thin line in every XY / YZ plane
    fz, fy, fx = orc.plane_flags(lab, True)
    assert fy.tolist() == [0, 0, 0, 1, 0, 0, 0]
    assert not fz.any() and not fx.any()
    lab[2, 1:3, 1:3] = 2             # a 2x2 blob in plane z=2 survives
    fz, fy, fx = orc.plane_flags(lab, True)
    assert fz.tolist() == [0, 0, 1, 0, 0, 0, 0]


# --- seeds and variance ------------------------------------------------------------
def _bruteforce_seeds(labidx, img, flags):
    """Rebuild the reference's slices volume (pyopencl_small.py:97-107) and scan it
    with _calc_var (:140-185) written out in float32 scalars, dev += tmp * tmp
    contracted to one fused multiply-add (exact rationals, one rounding) as
    the reference's compilers build it."""
    from tests.exact_f32 import fma
    Z, Y, X = labidx.shape
    fz, fy, fx = flags
    # -1: no label (off the labelled planes, or ignored), as the short-label path's ignore value
    ign = 0xFFFF if labidx.dtype == np.uint16 else oracle.IGN
    lab = np.where(labidx == ign, -1, labidx.astype(np.int32))
    sl = np.full(labidx.shape, -1, np.int32)
    for z in np.flatnonzero(fz):
        sl[z] = lab[z]
    for y in np.flatnonzero(fy):
        sl[:, y] = lab[:, y]
    for x in np.flatnonzero(fx):
        sl[:, :, x] = lab[:, :, x]
    f = np.float32
    raw = img.astype(np.float32)
    out = []
    for z in range(1, Z - 1):
        for y in range(1, Y - 1):
            for x in range(1, X - 1):
                seg = sl[z, y, x]
                if seg == -1:
                    continue
                B = raw[z, y, x]
                dev, summe = f(0), f(0)
                for n in (-1, 0, 1):
                    for o in (-1, 0, 1):
                        if sl[z, y + n, x + o] == seg:
                            t = f(B - raw[z, y + n, x + o]); dev = fma(t, t, dev); summe = f(summe + 1)
                for n in (-1, 1):
                    for o in (-1, 0, 1):
                        if summe < 9 and sl[z + n, y, x + o] == seg:
                            t = f(B - raw[z + n, y, x + o]); dev = fma(t, t, dev); summe = f(summe + 1)
                for n in (-1, 1):
                    for o in (-1, 1):
                        if summe < 9 and sl[z + n, y + o, x] == seg:
                            t = f(B - raw[z + n, y + o, x]); dev = fma(t, t, dev); summe = f(summe + 1)
                var = f(dev / summe)
                if var < 1:
                    var = f(1)
                out.append((z * Y * X + y * X + x, seg, var))
    pos = np.array([o[0] for o in out], np.int64)
    lab = np.array([o[1] for o in out], np.uint8)
    var = np.array([o[2] for o in out], np.float32)
    return pos, lab, var


@pytest.mark.parametrize("allaxis", [False, True])
@pytest.mark.parametrize("dtype", [np.uint8, np.float32])
def test_seeds_vs_bruteforce(allaxis, dtype):
    rng = np.random.default_rng(7 if allaxis else 8)
    shape = (9, 10, 11)
    lab = _random_sparse(shape, rng, 3, allaxis)
    lab[rng.random(shape) < 0.02] = oracle.IGN
    if dtype == np.uint8:
        img = rng.integers(0, 256, shape).astype(np.uint8)
    else:
        img = (rng.random(shape) * 255).astype(np.float32)
    flags = orc.plane_flags(lab, allaxis)
    got = orc.seeds(lab, img, flags)
    ref = _bruteforce_seeds(lab, img, flags)
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)


@pytest.mark.parametrize("allaxis", [False, True])
def test_seeds_256_labels_vs_bruteforce(allaxis):
    """All 256 label values 0..255 (uint16 indices, no ignored index), as the
    reference's short-label path holds them (pyopencl_small.py:90-95): label
    255 seeds like any other and its variance counts only its own voxels."""
    rng = np.random.default_rng(11 + allaxis)
    shape = (20, 22, 24)
    lab = np.zeros(shape, np.uint16)
    planes = [(2, None, None), (6, None, None)] + ([(None, 11, None)] if allaxis else [])
    for z, y, x in planes:
        # every value on interior voxels of the plane
        idx = (z if z is not None else slice(1, -1), y if y is not None else slice(1, -1), slice(1, -1))
        n = lab[idx].size
        vals = np.concatenate([rng.permutation(256), rng.integers(0, 256, n - 256)])
        lab[idx] = vals.reshape(lab[idx].shape)
    lab[2, 5:8, 5:9] = 255                  # a patch of label 255 with same-label neighbours
    img = rng.integers(0, 256, shape).astype(np.uint8)
    flags = orc.plane_flags(lab, allaxis)
    got = orc.seeds(lab, img, flags)
    ref = _bruteforce_seeds(lab, img, flags)
    for g, r in zip(got, ref):
        np.testing.assert_array_equal(g, r)
    assert set(np.unique(got[1])) == set(range(256))
    sel = got[1] == 255
    assert sel.sum() >= 12 and np.any(got[2][sel] > 1.0)


def test_variance_closed_forms():
    img = np.full((5, 5, 5), 100, np.uint8)
    lab = np.zeros((5, 5, 5), np.uint8)
    lab[2] = 1
    flags = orc.plane_flags(lab, False)
    pos, l, var = orc.seeds(lab, img, flags)
    assert np.all(var == 1.0)                      # uniform patch -> clamp to 1
    img[2, 1, 1] = 110                             # one same-label neighbour differs by 10
    pos, l, var = orc.seeds(lab, img, flags)
    centre = np.flatnonzero(pos == 2 * 25 + 2 * 5 + 2)[0]
    assert var[centre] == np.float32(100.0 / 9.0)  # 3x3 XY patch, one term of 100


# --- transition law -------------------------------------------------------------------
def _single_seed_walk(img, var, nbrw, sorw, key=1, ref=False, pos=None):
    Z, Y, X = img.shape
    p = np.array([pos if pos is not None else (Z // 2) * Y * X + (Y // 2) * X + X // 2], np.int64)
    lab = np.zeros(1, np.uint8)
    v = np.array([var], np.float32)
    if ref:
        return orc.ref_walk(img, p, lab, v, 1, nbrw, sorw)
    return orc.philox_walk(img, p, lab, v, 1, nbrw, sorw, key)


def test_one_step_frequencies_match_weights():
    """sorw=1: each walk takes one step; counts ~ Multinomial(N, W_d / sum W)."""
    img = np.full((5, 5, 5), 100, np.uint8)
    nb = {(3, 2, 2): 104, (1, 2, 2): 97, (2, 3, 2): 100, (2, 1, 2): 90, (2, 2, 3): 108, (2, 2, 1): 101}
    for k, v in nb.items():
        img[k] = v
    var = 20.0
    w = np.array([np.exp(-(100.0 - v) ** 2 / (2 * var)) for v in nb.values()])
    p = w / w.sum()
    N = 400_000
    hits, acc = _single_seed_walk(img, var, N, 1, key=3)
    assert acc == N
    counts = np.array([hits[0][k] for k in nb], np.float64)
    chi2 = float(((counts - N * p) ** 2 / (N * p)).sum())
    assert chi2 < 20.5   # 5 dof, p = 0.001


def test_planted_direction_swap_fails_chi2():
    """The chi^2 test above rejects a walker whose +y/-y are swapped."""
    img = np.full((5, 5, 5), 100, np.uint8)
    img[2, 3, 2], img[2, 1, 2] = 100, 90
    w = np.array([1, 1, 1, np.exp(-100 / 40.0), 1, 1])
    p = w / w.sum()
    N = 400_000
    hits, _ = _single_seed_walk(img, 20.0, N, 1, key=3)
    order = [(3, 2, 2), (1, 2, 2), (2, 1, 2), (2, 3, 2), (2, 2, 3), (2, 2, 1)]   # swapped y
    counts = np.array([hits[0][k] for k in order], np.float64)
    chi2 = float(((counts - N * p) ** 2 / (N * p)).sum())
    assert chi2 > 100


@pytest.mark.parametrize("ref", [False, True])
def test_uniform_image_isotropic(ref):
    img = np.full((5, 5, 5), 77, np.uint8)
    N = 300_000
    if ref:
        # the reference stream is sequential per seed: one seed, N walks of one step
        hits, acc = _single_seed_walk(img, 1.0, N, 1, ref=True)
    else:
        hits, acc = _single_seed_walk(img, 1.0, N, 1)
    assert acc == N
    nbrs = [(3, 2, 2), (1, 2, 2), (2, 3, 2), (2, 1, 2), (2, 2, 3), (2, 2, 1)]
    frac = np.array([hits[0][k] for k in nbrs]) / N
    # within 7 sigma of 1/6, for the Philox stream and for the reference stream
    # as its compilers contract it
    assert np.all(np.abs(frac - 1 / 6) < 0.0048), frac


@pytest.mark.parametrize("ref", [False, True])
def test_all_weights_underflow_steps_plus_z(ref):
    """Step 1: every neighbour weight underflows -> rand == 0 -> +z.  Step 2: the
    only non-zero weight points back at the seed (same intensity as B) -> -z."""
    img = np.full((7, 7, 7), 255, np.uint8)
    img[3, 3, 3] = 0
    hits, acc = _single_seed_walk(img, 1.0, 50, 2, ref=ref, pos=3 * 49 + 3 * 7 + 3)
    assert acc == 100
    assert hits[0][4, 3, 3] == 50 and hits[0][3, 3, 3] == 50
    assert hits.sum() == acc


@pytest.mark.parametrize("ref", [False, True])
def test_moves_off_interior_rejected_but_counted(ref):
    img = np.random.default_rng(0).integers(0, 255, (3, 3, 3)).astype(np.uint8)
    hits, acc = _single_seed_walk(img, 5.0, 7, 13, ref=ref)
    assert acc == 0 and hits.sum() == 0


def test_hit_conservation_no_rejection():
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (40, 40, 40)).astype(np.uint8)
    pos = np.array([20 * 1600 + 20 * 40 + 20, 19 * 1600 + 21 * 40 + 20], np.int64)
    lab = np.array([0, 1], np.uint8)
    var = np.array([30.0, 400.0], np.float32)
    hits, acc = orc.philox_walk(img, pos, lab, var, 2, 9, 17, 11)   # 17 steps never reach a wall
    assert acc == 2 * 9 * 17
    assert int(hits.sum(dtype=np.int64)) == acc


def test_philox_walks_are_independent_of_seed_order_and_subsets():
    rng = np.random.default_rng(6)
    img = rng.integers(0, 256, (16, 16, 16)).astype(np.uint8)
    pos = rng.choice(np.arange(16 ** 3), 40, replace=False).astype(np.int64)
    z, y, x = np.unravel_index(pos, img.shape)
    keep = (z > 0) & (z < 15) & (y > 0) & (y < 15) & (x > 0) & (x < 15)
    pos = pos[keep]
    lab = rng.integers(0, 3, pos.size).astype(np.uint8)
    var = rng.uniform(1, 500, pos.size).astype(np.float32)
    full, a_full = orc.philox_walk(img, pos, lab, var, 3, 5, 101, 9)
    perm = rng.permutation(pos.size)
    shuf, _ = orc.philox_walk(img, pos[perm], lab[perm], var[perm], 3, 5, 101, 9)
    np.testing.assert_array_equal(full, shuf)
    h = pos.size // 2
    a, aa = orc.philox_walk(img, pos[:h], lab[:h], var[:h], 3, 5, 101, 9)
    b, ab = orc.philox_walk(img, pos[h:], lab[h:], var[h:], 3, 5, 101, 9)
    np.testing.assert_array_equal(full, a + b)
    assert a_full == aa + ab


def test_philox_stream_layout():
    """Walk w, step s of seed p uses word s%4 of Philox((p, 0, w, s//4), key)."""
    img = np.full((9, 9, 9), 50, np.uint8)    # uniform: d = number of prefixes <= 6u*1
    p0 = 4 * 81 + 4 * 9 + 4
    key = 0x0123456789ABCDEF
    hits, _ = orc.philox_walk(img, np.array([p0], np.int64), np.zeros(1, np.uint8),
                              np.ones(1, np.float32), 1, 2, 6, key)
    expect = np.zeros((9, 9, 9), np.uint32)
    step = {0: (1, 0, 0), 1: (-1, 0, 0), 2: (0, 1, 0), 3: (0, -1, 0), 4: (0, 0, 1), 5: (0, 0, -1)}
    for w in range(2):
        z = y = x = 4
        for s in range(6):
            r = orc.philox4x32_10(np.array([[p0, 0, w, s // 4]], np.uint32),
                                  np.array([key & 0xffffffff, key >> 32], np.uint32))[0]
            u = np.float32(int(r[s % 4]) >> 8) * np.float32(2.0 ** -24)
            W = np.cumsum(np.ones(6, np.float32))
            rand = W[5] * u
            d = 0 if rand == 0 else int(np.sum(rand >= W[:5]))
            dz, dy, dx = step[d]
            if 0 < z + dz < 8 and 0 < y + dy < 8 and 0 < x + dx < 8:
                z, y, x = z + dz, y + dy, x + dx
                expect[z, y, x] += 1
    np.testing.assert_array_equal(hits[0], expect)


# --- the reference stream -------------------------------------------------------------
F32 = np.float32
MRG = dict(norm=F32(2.328306549295728e-10), m1=F32(4294967087.0), m2=F32(4294944443.0), a12=F32(1403580.0),
           a13n=F32(810728.0), a21=F32(527612.0), a23n=F32(1370589.0))


def _mrg_model(index, n, contract=True):
    """pyopencl_small.py:213-223, 256-282 in exact rationals rounded once per
    operation; contract=True fuses a*b - c*d into fma(a, b, -(c*d)) and p -= k*m
    into fma(-k, m, p), as FP_CONTRACT (on by default) builds it."""
    from tests.exact_f32 import add, div, fma, mul, sub
    c = MRG
    s1 = [F32(index)] * 3
    s2 = [F32(index)] * 3
    out = []
    for _ in range(n):
        comps = []
        for st, a, b, m in ((s1, c["a12"], c["a13n"], c["m1"]), (s2, c["a21"], c["a23n"], c["m2"])):
            lead = st[1] if a is c["a12"] else st[2]    # a12 * s11, a21 * s22
            p = fma(a, lead, -mul(b, st[0])) if contract else sub(mul(a, lead), mul(b, st[0]))
            k = int(div(p, m))                          # float quotient truncated to long
            p = fma(-F32(k), m, p) if contract else sub(p, mul(F32(k), m))
            if p < 0:
                p = add(p, m)
            st[0], st[1], st[2] = st[1], st[2], p
            comps.append(p)
        p1, p2 = comps
        out.append(mul(add(sub(p1, p2), c["m1"]), c["norm"]) if p1 <= p2 else mul(sub(p1, p2), c["norm"]))
    return np.array(out, np.float32)


def test_reference_stream_known_answer():
    """First uniform of the stream seeded with voxel 1, derived by hand:
    component 1: fma(1403580, 1, -810728) = 592852 (k1 = 0), component 2:
    fma(527612, 1, -1370589) = -842977 (k1 = 0), + m2 (4294944443 -> 4294944512
    as float32) = 4294101535 -> 4294101504 (multiple of 256); p1 <= p2, so
    (592852 - 4294101504) = -4293508652 -> -4293508608, + m1 (4294967040) =
    1458432, times norm (2.3283064365386963e-10) -> 3.3956766e-04.  The next
    values involve rounded products and come from the exact model."""
    assert float(MRG["m1"]) == 4294967040.0 and float(MRG["m2"]) == 4294944512.0
    u = orc.mrg_stream(1, 64)
    assert u[0] == np.float32(1458432 * np.float64(MRG["norm"]))
    assert u[0] == np.float32(3.3956766e-04)
    for index in (1, 7, 1234567, 98765432, 2 ** 26 + 3):
        np.testing.assert_array_equal(orc.mrg_stream(index, 64), _mrg_model(index, 64))


def test_reference_stream_statistics_and_uncontracted_planting():
    """Across seeds the stream's directions on a uniform image are 1/6 each and
    4000-step walks do not drift; the same recurrence without contraction
    (the planted mistake, in the exact model) drifts by tens of voxels."""
    rng = np.random.default_rng(0)
    idx = rng.integers(1, 512 ** 3, size=1500)
    d = np.stack([np.minimum((orc.mrg_stream(int(i), 4000) * 6).astype(np.int64), 5) for i in idx])
    freq = np.bincount(d.ravel(), minlength=6) / d.size
    assert np.all(np.abs(freq - 1 / 6) < 0.002), freq
    drift = [((d == 2 * a).sum(1) - (d == 2 * a + 1).sum(1)).mean() for a in range(3)]
    assert np.all(np.abs(drift) < 5.0), drift      # sigma of the mean ~ 36.5 / sqrt(1500) ~ 0.9
    bad = np.stack([np.minimum((_mrg_model(int(i), 1200, contract=False) * 6).astype(np.int64), 5)
                    for i in idx[:40]])
    fb = np.bincount(bad.ravel(), minlength=6) / bad.size
    assert np.max(np.abs(fb - 1 / 6)) > 0.004, fb


def test_reference_walk_consumes_the_stream_in_order():
    """One seed on a uniform image, N walks of one step: walk n moves by the
    direction of its uniform u_n (rand = 6 u_n against W_d = d + 1), so the
    hit counts of the six neighbours follow from the pinned stream."""
    img = np.full((5, 5, 5), 77, np.uint8)
    N = 5000
    hits, acc = _single_seed_walk(img, 1.0, N, 1, ref=True)
    p0 = 2 * 25 + 2 * 5 + 2
    u = orc.mrg_stream(p0, N)
    rand = (np.float32(6.0) * u).astype(np.float32)
    d = np.where(rand == 0, 0, (rand[:, None] >= np.arange(1, 6, dtype=np.float32)[None, :]).sum(1))
    want = np.bincount(d, minlength=6)
    nbrs = [(3, 2, 2), (1, 2, 2), (2, 3, 2), (2, 1, 2), (2, 2, 3), (2, 2, 1)]
    np.testing.assert_array_equal([hits[0][k] for k in nbrs], want)


def test_reference_stream_and_philox_agree_statistically():
    import oracle.pipeline as opipe
    from biomedisa_amd.synth import ellipsoids_volume
    from biomedisa_amd.helpers import per_label_dice
    vol = ellipsoids_volume(shape=(48, 48, 48), seed=11, semi=(7, 14), every=4)
    a = opipe.smart_interpolation(vol.img, vol.labels, nbrw=6, sorw=300, rng="philox")
    b = opipe.smart_interpolation(vol.img, vol.labels, nbrw=6, sorw=300, rng="reference")
    d = per_label_dice(a["regular"], b["regular"])
    assert min(d.values()) >= 0.95, d
    assert np.mean(list(d.values())) >= 0.98, d


# --- argmax -----------------------------------------------------------------------
def test_argmax_ties_and_zero_bruteforce():
    rng = np.random.default_rng(3)
    hits = rng.integers(0, 4, (5, 7, 6, 5)).astype(np.uint32)
    hits[:, 0, 0, 0] = 0
    values = np.array([0, 3, 9, 17, 200], np.uint8)
    got = orc.argmax(hits, values)
    for idx in np.ndindex(hits.shape[1:]):
        col = [int(hits[(l,) + idx]) for l in range(5)]
        best = col.index(max(col))          # first maximum
        assert got[idx] == values[best]
    assert got[0, 0, 0] == 0


# --- curvature smoothing -----------------------------------------------------------------
def test_smooth_linear_field_is_fixed():
    z, y, x = np.meshgrid(np.arange(12), np.arange(13), np.arange(14), indexing="ij")
    phi = (3 * z - 2 * y + 5 * x + 7).astype(np.float32)
    out = orc.smooth(phi, 5)
    np.testing.assert_array_equal(out, phi)


def test_smooth_paraboloid_curvature_is_two():
    """phi = x^2 + y^2: kappa = (2*4y^2 + 2*4x^2)/(4x^2+4y^2) = 2 away from the axis."""
    z, y, x = np.meshgrid(np.arange(8), np.arange(-6, 7), np.arange(-6, 7), indexing="ij")
    phi = (x * x + y * y).astype(np.float32)
    out = orc.smooth(phi, 1)
    inner = np.zeros(phi.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    axis = (x == 0) & (y == 0)
    d = out - phi
    np.testing.assert_allclose(d[inner & ~axis], 0.02, rtol=0, atol=2e-5)
    assert np.all(d[inner & axis] == 0)              # 0 / (0 + eps) = 0
    np.testing.assert_array_equal(out[~inner], phi[~inner])


def test_smooth_sphere_reference_stencil_closed_form():
    """phi = |p - c|.  The reference stencil (opencl_kernels.py:193-215) takes
    dx, dy, dz with a flipped sign and dxz, dyz with a flipped sign but dxy with
    the textbook sign, so its kappa is
        (sum phi_ii (|grad|^2 - phi_i^2) - 2 phi_x phi_y phi_xy
         + 2 phi_x phi_z phi_xz + 2 phi_y phi_z phi_yz) / |grad|^2
      = 2/r - 4 z^2 (x^2 + y^2) / r^5          for phi = r,
    not the mean curvature 2/r.  The oracle must keep that."""
    n = 41
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) - 20] * 3, indexing="ij")
    r = np.sqrt(x * x + y * y + z * z)
    phi = r.astype(np.float32)
    out = orc.smooth(phi, 1)
    sel = (r > 8) & (r < 15)
    kappa = (out - phi)[sel] / 0.01
    expect = 2 / r - 4 * z * z * (x * x + y * y) / r ** 5
    np.testing.assert_allclose(kappa, expect[sel], rtol=0.05, atol=3e-3)
    textbook = 2 / r[sel]
    assert np.abs(kappa - textbook).max() > 0.05     # the quirk is visible


# --- uncertainty ----------------------------------------------------------------------------
def test_uncertainty_closed_forms():
    hits = np.zeros((4, 3, 3, 5), np.float32)
    hits[:, 1, 1, 1] = [10, 5, 2, 0]     # 1 - (1-2/10)(1-5/10) = 0.6
    hits[:, 1, 1, 2] = [0, 0, 7, 0]      # 1 - (1-0)(1-0) = 0
    hits[:, 1, 1, 3] = [4, 4, 4, 4]      # 1 - 0*0 = 1
    u = orc.uncertainty(hits)
    np.testing.assert_allclose(u[1, 1, 1], 0.6, rtol=1e-6)
    assert u[1, 1, 2] == 0.0 and u[1, 1, 3] == 1.0
    assert (u[1, 1, 1] * np.float32(255)).astype(np.uint8) == 153
