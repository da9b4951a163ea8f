"""NumPy restatement of the surrogate-guided kernels (test infrastructure only).

Only tests/, smoke() and bench's CPU baseline use this module; the product
path never does.  The guided mode has no counterpart in the reference: these
functions define what biomedisa_amd/csrc/bmx_guide.hip computes, in plain
array code, so that the kernels are checked bit for bit.  They are pinned in
tests/test_guide_oracle.py against library routines (scipy's binary dilation
with a cube structuring element), the integer argmax of oracle/bmx_oracle.c,
float64 softmax away from ties, and brute-force loops on small volumes.

FP32 arithmetic follows the kernels' order with every operation rounded to
float32 (the kernels are compiled without contraction): logits
b[l] + W[l,0]*x[0] + ... + W[l,15]*x[15] left to right, the softmax
denominator summed in label order with the engine's own exp (orc.expf).
"""
import numpy as np

from . import orc
from .net import to_bf16_bits

LABEL_NONE = 0xFF
N_CH = 32          # network input channels
L_ONEHOT = 30      # label indices with a one-hot input channel
HEAD_CH = 16       # channels of the decoder output read by the head


def net_input(img, lab, flags, grid):
    """bf16 bits [Zp,Yp,Xp,16]: image/127.5-1, seed indicator, one-hot label index.

    A seed voxel lies on a labelled plane of any axis (flags = [fz | fy | fx])
    and carries a label index other than 0xFF (ignored)."""
    Z, Y, X = lab.shape
    Zp, Yp, Xp = grid
    out = np.zeros((Zp, Yp, Xp, N_CH), np.uint16)
    f = np.asarray(flags, np.uint8)
    ch0 = np.asarray(img, np.float32) * np.float32(1.0 / 127.5) - np.float32(1.0)
    out[:Z, :Y, :X, 0] = to_bf16_bits(ch0.astype(np.float32))
    plane = (f[:Z][:, None, None] | f[Z:Z + Y][None, :, None] | f[Z + Y:Z + Y + X][None, None, :]) != 0
    seed = plane & (lab != LABEL_NONE)
    one = np.uint16(0x3F80)
    out[:Z, :Y, :X, 1][seed] = one
    for l in range(L_ONEHOT):
        out[:Z, :Y, :X, 2 + l][seed & (lab == l)] = one
    return out


def head_logits(xbits, W, b):
    """FP32 logits [..., L] of the 1x1x1 head, accumulated in channel order."""
    x = (np.asarray(xbits, np.uint32) << 16).view(np.float32)
    W = np.asarray(W, np.float32)
    b = np.asarray(b, np.float32)
    L = W.shape[0]
    lg = np.empty(x.shape[:-1] + (L,), np.float32)
    for l in range(L):
        s = np.full(x.shape[:-1], b[l], np.float32)
        for c in range(HEAD_CH):
            s = s + W[l, c] * x[..., c]
        lg[..., l] = s
    return lg


def head_predict(xbits, dims, W, b, thr):
    """(pred label index u8, uncertain flag u8) for the (Z,Y,X) corner of the padded grid."""
    Z, Y, X = dims
    lg = head_logits(xbits[:Z, :Y, :X], W, b)
    pred = np.argmax(lg, axis=-1)
    best = np.take_along_axis(lg, pred[..., None], -1)[..., 0]
    den = np.zeros((Z, Y, X), np.float32)
    for l in range(lg.shape[-1]):
        d = (lg[..., l] - best).astype(np.float32)
        den = den + orc.expf(d.ravel()).reshape(d.shape)
    unc = (np.float32(1.0) / den) < np.float32(thr)
    return pred.astype(np.uint8), unc.astype(np.uint8)


def boundary(pred):
    """1 where a 6-neighbour inside the volume has another label."""
    p = np.asarray(pred)
    b = np.zeros(p.shape, bool)
    for ax in range(3):
        d = np.diff(p, axis=ax) != 0
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax] = slice(0, -1)
        hi[ax] = slice(1, None)
        b[tuple(lo)] |= d
        b[tuple(hi)] |= d
    return b.astype(np.uint8)


def dilate(mask, r):
    """Chebyshev dilation by r (a (2r+1)^3 cube, clipped at the volume border),
    as three separable running maxima."""
    m = np.asarray(mask) != 0
    if r <= 0:
        return m.astype(np.uint8)
    for ax in range(3):
        n = m.shape[ax]
        out = np.zeros_like(m)
        rr = min(r, n - 1)
        for k in range(-rr, rr + 1):
            src = [slice(None)] * 3
            dst = [slice(None)] * 3
            if k >= 0:
                src[ax], dst[ax] = slice(k, n), slice(0, n - k)
            else:
                src[ax], dst[ax] = slice(0, n + k), slice(-k, n)
            out[tuple(dst)] |= m[tuple(src)]
        m = out
    return m.astype(np.uint8)


def guided_masks(pred, unc, brad, srad):
    """U = dilate(boundary(pred), brad) | unc;  need = dilate(U, srad)."""
    U = dilate(boundary(pred), brad) | (np.asarray(unc) != 0)
    return U.astype(np.uint8), dilate(U, srad)


def ring_mask(need, r):
    """dilate(need, r) & ~need: the far-field ring around the near set."""
    n = np.asarray(need) != 0
    return (dilate(n, r).astype(bool) & ~n).astype(np.uint8)


def guided_disagree(hits, pred, U, minhits):
    """Outside U: total hits >= minhits (and > 0) and first argmax != pred."""
    h = np.asarray(hits, np.uint32)
    tot = h.sum(0, dtype=np.uint64)
    am = np.argmax(h, axis=0)
    d = (np.asarray(U) == 0) & (tot >= max(int(minhits), 1)) & (am != np.asarray(pred))
    return d.astype(np.uint8)


def repair_masks(D, U, r, srad):
    """U2 = dilate(D, r) & ~U;  need = dilate(U2, srad)."""
    U2 = dilate(D, r).astype(bool) & (np.asarray(U) == 0)
    return U2.astype(np.uint8), dilate(U2, srad)


def repair_hits(hits, hits2, U2, U):
    """hits with hits2's counts on U2, and U | U2."""
    m = np.asarray(U2) != 0
    h = np.array(hits, np.uint32)
    h[:, m] = np.asarray(hits2, np.uint32)[:, m]
    return h, ((np.asarray(U) != 0) | m).astype(np.uint8)


def merge_guided(hits, values, pred, U):
    """values[first argmax of hits] on U, values[pred] elsewhere (hits [L,...])."""
    values = np.asarray(values, np.uint8)
    walk = orc.argmax(np.ascontiguousarray(hits, np.uint32), values)
    return np.where(np.asarray(U) != 0, walk, values[pred]).astype(np.uint8)


def select_labels(a, pred, U, values):
    values = np.asarray(values, np.uint8)
    return np.where(np.asarray(U) != 0, a, values[pred]).astype(np.uint8)


def hits_target(hits, grid, L, hard):
    """Training targets [Zp,Yp,Xp,L] float32 and mask [Zp,Yp,Xp] u8 from hits [nl,Z,Y,X]."""
    h = np.asarray(hits, np.uint32)
    nl, Z, Y, X = h.shape
    Zp, Yp, Xp = grid
    t = np.zeros((Zp, Yp, Xp, L), np.float32)
    m = np.zeros((Zp, Yp, Xp), np.uint8)
    if hard:
        idx = orc.argmax(np.ascontiguousarray(h), np.arange(nl, dtype=np.uint8))
        t[:Z, :Y, :X] = (np.arange(L)[None, None, None, :] == idx[..., None]).astype(np.float32)
        m[:Z, :Y, :X] = 1
        return t, m
    tot = np.zeros((Z, Y, X), np.float32)
    for l in range(nl):
        tot = tot + h[l].astype(np.float32)
    ok = tot > 0
    safe = np.where(ok, tot, np.float32(1))
    for l in range(nl):
        t[:Z, :Y, :X, l] = np.where(ok, h[l].astype(np.float32) / safe, np.float32(0))
    m[:Z, :Y, :X] = ok
    return t, m


def hits_index(hits, grid):
    """Training labels [Zp,Yp,Xp] u8 (first argmax index, 0 for all-zero) and mask."""
    h = np.asarray(hits, np.uint32)
    nl, Z, Y, X = h.shape
    idx = np.zeros(grid, np.uint8)
    m = np.zeros(grid, np.uint8)
    idx[:Z, :Y, :X] = orc.argmax(np.ascontiguousarray(h), np.arange(nl, dtype=np.uint8))
    m[:Z, :Y, :X] = 1
    return idx, m


def select_seeds(seeds, mask):
    """The seeds (structured array with field 'pos') whose voxel is set in mask."""
    return seeds[np.asarray(mask).ravel()[seeds["pos"].astype(np.int64)] != 0]
