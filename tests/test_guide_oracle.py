"""Pins of oracle/guide.py (the surrogate-guided kernels' CPU restatement)
against library routines, float64 arithmetic and brute force."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import guide as G
from oracle import orc
from oracle.net import from_bf16_bits


@pytest.mark.parametrize("r", [0, 1, 2, 5, 20])
def test_dilate_is_cube_dilation(r):
    rng = np.random.default_rng(r)
    m = rng.random((13, 17, 11)) < 0.002
    ref = ndimage.binary_dilation(m, structure=np.ones((2 * r + 1,) * 3, bool)) if r else m
    np.testing.assert_array_equal(G.dilate(m, r), ref.astype(np.uint8))


def test_dilate_planted_mistake_fails():
    """A 6-connected (cross) structuring element is not the Chebyshev cube."""
    m = np.zeros((9, 9, 9), bool)
    m[4, 4, 4] = True
    cross = ndimage.binary_dilation(m, iterations=2)
    assert not np.array_equal(G.dilate(m, 2), cross.astype(np.uint8))
    assert G.dilate(m, 2).sum() == 125


def test_boundary_brute_force():
    rng = np.random.default_rng(1)
    p = rng.integers(0, 3, (6, 7, 8)).astype(np.uint8)
    p[2:5, 2:5, 2:6] = 1
    ref = np.zeros(p.shape, np.uint8)
    Z, Y, X = p.shape
    for z in range(Z):
        for y in range(Y):
            for x in range(X):
                for dz, dy, dx in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]:
                    zz, yy, xx = z + dz, y + dy, x + dx
                    if 0 <= zz < Z and 0 <= yy < Y and 0 <= xx < X and p[zz, yy, xx] != p[z, y, x]:
                        ref[z, y, x] = 1
    np.testing.assert_array_equal(G.boundary(p), ref)


def test_guided_masks_definition():
    """U and need rebuilt with scipy: the 6-neighbour label change as grey dilation/erosion
    with a cross (edge-replicated), then cube dilations."""
    rng = np.random.default_rng(6)
    p = np.zeros((20, 21, 22), np.uint8)
    p[5:15, 5:15, 5:15] = 1
    p[2:6, 12:19, 3:9] = 2
    unc = (rng.random(p.shape) < 0.002).astype(np.uint8)
    cross = ndimage.generate_binary_structure(3, 1)
    b = (ndimage.grey_dilation(p, footprint=cross, mode="nearest") != p) | \
        (ndimage.grey_erosion(p, footprint=cross, mode="nearest") != p)
    np.testing.assert_array_equal(G.boundary(p), b.astype(np.uint8))
    for brad, srad in [(1, 2), (3, 0), (0, 4)]:
        U, need = G.guided_masks(p, unc, brad, srad)
        Uref = (ndimage.binary_dilation(b, np.ones((2 * brad + 1,) * 3, bool)) if brad else b) | (unc != 0)
        nref = ndimage.binary_dilation(Uref, np.ones((2 * srad + 1,) * 3, bool)) if srad else Uref
        np.testing.assert_array_equal(U, Uref.astype(np.uint8))
        np.testing.assert_array_equal(need, nref.astype(np.uint8))


def test_hits_target_hard_and_soft():
    rng = np.random.default_rng(2)
    nl, dims, grid = 5, (4, 5, 6), (8, 8, 8)
    h = rng.integers(0, 4, (nl,) + dims).astype(np.uint32)
    h[:, 0, 0, 0] = 0
    h[:, 1, 1, 1] = [2, 3, 3, 0, 1]
    t, m = G.hits_target(h, grid, 14, hard=True)
    assert t.shape == (8, 8, 8, 14) and m.sum() == np.prod(dims)
    idx = np.argmax(h, axis=0)  # NumPy's first maximum; all-zero -> 0
    np.testing.assert_array_equal(t[:4, :5, :6].argmax(-1), idx)
    assert t[1, 1, 1, 1] == 1 and t[0, 0, 0, 0] == 1
    np.testing.assert_array_equal(t.sum(-1), m)
    ts, ms = G.hits_target(h, grid, 14, hard=False)
    tot = h.sum(0).astype(np.float64)
    ok = tot > 0
    np.testing.assert_array_equal(ms[:4, :5, :6], ok)
    ref = np.where(ok, h / np.where(ok, tot, 1), 0).transpose(1, 2, 3, 0)
    np.testing.assert_allclose(ts[:4, :5, :6, :nl], ref, rtol=2 ** -23, atol=0)
    assert ts[4:].sum() == 0 and ms[4:].sum() == 0


def test_net_input_channels():
    rng = np.random.default_rng(3)
    Z, Y, X = 5, 6, 7
    img = rng.integers(0, 256, (Z, Y, X)).astype(np.uint8)
    lab = rng.integers(0, 34, (Z, Y, X)).astype(np.uint8)
    lab[0, 0, :3] = G.LABEL_NONE
    flags = np.zeros(Z + Y + X, np.uint8)
    flags[0] = 1          # z = 0 plane
    flags[Z + 3] = 1      # y = 3 plane
    out = G.net_input(img, lab, flags, (8, 8, 8))
    assert out.shape == (8, 8, 8, G.N_CH)
    x = from_bf16_bits(out).astype(np.float64)
    np.testing.assert_allclose(x[:Z, :Y, :X, 0], img / 127.5 - 1, atol=2 ** -8)
    assert not x[Z:].any() and not x[:, Y:].any() and not x[:, :, X:].any()
    for z in range(Z):
        for y in range(Y):
            for xx in range(X):
                seed = (z == 0 or y == 3) and lab[z, y, xx] != G.LABEL_NONE
                assert x[z, y, xx, 1] == (1.0 if seed else 0.0)
                hot = np.zeros(G.L_ONEHOT)
                if seed and lab[z, y, xx] < G.L_ONEHOT:
                    hot[lab[z, y, xx]] = 1
                np.testing.assert_array_equal(x[z, y, xx, 2:], hot)


def test_head_predict_against_float64_softmax():
    rng = np.random.default_rng(4)
    dims, grid = (6, 7, 9), (8, 8, 16)
    xb = (rng.standard_normal(grid + (16,)).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)
    W = (rng.standard_normal((9, 16)) * 0.4).astype(np.float32)
    b = rng.standard_normal(9).astype(np.float32)
    for thr in (0.5, 0.9):
        pred, unc = G.head_predict(xb, dims, W, b, thr)
        x = from_bf16_bits(xb[:6, :7, :9]).astype(np.float64)
        lg = x @ W.T.astype(np.float64) + b
        srt = np.sort(lg, -1)
        clear = srt[..., -1] - srt[..., -2] > 1e-4
        np.testing.assert_array_equal(pred[clear], lg.argmax(-1)[clear])
        p = np.exp(lg - lg.max(-1, keepdims=True))
        pmax = 1 / p.sum(-1)
        far = np.abs(pmax - thr) > 1e-5
        np.testing.assert_array_equal(unc[far], (pmax < thr)[far])
        assert 0 < unc.mean() < 1


def test_merge_guided_and_select():
    rng = np.random.default_rng(5)
    h = rng.integers(0, 5, (4, 3, 4, 5)).astype(np.uint32)
    values = np.array([0, 7, 9, 200], np.uint8)
    pred = rng.integers(0, 4, (3, 4, 5)).astype(np.uint8)
    U = (rng.random((3, 4, 5)) < 0.5).astype(np.uint8)
    got = G.merge_guided(h, values, pred, U)
    ref = np.where(U == 1, values[np.argmax(h, axis=0)], values[pred])
    np.testing.assert_array_equal(got, ref)
    a = rng.integers(0, 255, (3, 4, 5)).astype(np.uint8)
    np.testing.assert_array_equal(G.select_labels(a, pred, U, values), np.where(U == 1, a, values[pred]))


def test_select_seeds():
    from biomedisa_amd._hip import SEED_DTYPE
    s = np.zeros(10, SEED_DTYPE)
    s["pos"] = np.arange(10) * 3
    mask = np.zeros(40, np.uint8)
    mask[[0, 9, 27, 28]] = 1
    np.testing.assert_array_equal(G.select_seeds(s, mask)["pos"], [0, 9, 27])


def test_guided_disagree_brute_force():
    """Voxel by voxel: outside U, at least minhits hits, and the first label with
    the largest count differs from the proposal (ties go to the lower label)."""
    rng = np.random.default_rng(31)
    nl, dims = 4, (5, 6, 7)
    hits = rng.integers(0, 4, (nl,) + dims).astype(np.uint32)
    hits[:, 0, 0, :3] = 0
    hits[:, 1, 1, 1] = [2, 2, 0, 0]
    pred = rng.integers(0, nl, dims).astype(np.uint8)
    pred[1, 1, 1] = 1
    U = (rng.random(dims) < 0.3).astype(np.uint8)
    U[1, 1, 1] = 0
    for minh in (0, 1, 5):
        ref = np.zeros(dims, np.uint8)
        for z, y, x in np.ndindex(*dims):
            h = [int(hits[l, z, y, x]) for l in range(nl)]
            best = max(range(nl), key=lambda l: (h[l], -l))
            ref[z, y, x] = int(not U[z, y, x] and sum(h) >= max(minh, 1) and best != pred[z, y, x])
        np.testing.assert_array_equal(G.guided_disagree(hits, pred, U, minh), ref)
    assert G.guided_disagree(hits, pred, U, 1)[1, 1, 1] == 1   # tie 2/2 -> label 0 != 1


def test_repair_masks_and_hits():
    rng = np.random.default_rng(32)
    dims = (12, 13, 14)
    D = rng.random(dims) < 0.004
    U = rng.random(dims) < 0.2
    U2, need = G.repair_masks(D, U, 2, 3)
    ref = ndimage.binary_dilation(D, structure=np.ones((5, 5, 5), bool)) & ~U
    np.testing.assert_array_equal(U2, ref.astype(np.uint8))
    np.testing.assert_array_equal(need, ndimage.binary_dilation(ref, structure=np.ones((7,) * 3, bool))
                                  .astype(np.uint8))
    hits = rng.integers(0, 9, (3,) + dims).astype(np.uint32)
    hits2 = rng.integers(0, 9, (3,) + dims).astype(np.uint32)
    h, U3 = G.repair_hits(hits, hits2, U2, U)
    m = U2.astype(bool)
    np.testing.assert_array_equal(h[:, m], hits2[:, m])
    np.testing.assert_array_equal(h[:, ~m], hits[:, ~m])
    np.testing.assert_array_equal(U3, (U | m).astype(np.uint8))
