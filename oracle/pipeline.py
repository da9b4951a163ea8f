"""CPU restatement of the whole smart-interpolation call (test infrastructure).

Follows biomedisa/interpolation.py:135-204 (dtype handling, pad, bounding-box
crop), biomedisa_helper.py:639-724 (label set), the backend flow of
pyopencl_small.py:80-135 and rw_small.py:114-201 (walk, smoothing,
uncertainty, argmax, get_labels).  Heavy loops run in libbmx_oracle.so.
Written independently of biomedisa_amd so that the two can check each other.
"""
import numpy as np

from . import orc

PHILOX_KEY = 0x5EED0B10D1CE0001


def _label_set(labels):
    vals = np.unique(labels)
    vals = vals[vals >= 0]
    if np.any(vals > 255):
        labels = labels.copy()
        labels[labels > 255] = 0
        vals = vals[vals <= 255]
    if not np.any(vals == 0):
        vals = np.concatenate([[0], vals])
    return labels, vals


def _bbox(lab):
    """predict_blocksize (biomedisa_helper.py:1105-1124), slice-by-slice as written there."""
    zsh, ysh, xsh = lab.shape
    amin_z, amax_z, amin_y, amax_y, amin_x, amax_x = zsh, 0, ysh, 0, xsh, 0
    for k in range(zsh):
        y, x = np.nonzero(lab[k])
        if x.any():
            amin_x = min(amin_x, int(x.min())); amax_x = max(amax_x, int(x.max()))
            amin_y = min(amin_y, int(y.min())); amax_y = max(amax_y, int(y.max()))
            amin_z = min(amin_z, k); amax_z = max(amax_z, k)
    lo = lambda v: v - 100 if v - 100 > 0 else 0
    hi = lambda v, n: v + 100 if v + 100 < n else n
    return lo(amin_z), hi(amax_z, zsh), lo(amin_y), hi(amax_y, ysh), lo(amin_x), hi(amax_x, xsh)


def prepare(data, labelData):
    """-> dict(img, labidx, values, crop, padded_shape, nol) exactly as the walker sees them."""
    labels = np.array(labelData)
    if labels.dtype.kind == "f" or labels.dtype in (np.uint32, np.int64, np.uint64):
        labels = labels.astype(np.int32)
    labels, vals = _label_set(labels)
    if data.dtype == np.int8:
        img = (data.astype(np.int16) + 128).astype(np.uint8)
    elif data.dtype == np.uint8:
        img = data.copy()
    else:
        img = data.astype(np.float32)
        img -= np.amin(img)
        img /= np.amax(img)
        img *= 255.0
    img = np.pad(img, 1)
    labels = np.pad(labels, 1)
    z0, z1, y0, y1, x0, x1 = _bbox(labels)
    img_c = np.ascontiguousarray(img[z0:z1, y0:y1, x0:x1])
    lab_c = np.ascontiguousarray(labels[z0:z1, y0:y1, x0:x1])
    labidx = np.full(lab_c.shape, 0xFF, np.uint8)
    for i, v in enumerate(vals):
        labidx[lab_c == v] = i
    return dict(img=img_c, labidx=labidx, lab=lab_c, values=vals.astype(np.uint8),
                crop=(z0, z1, y0, y1, x0, x1), padded_shape=img.shape, nol=len(vals))


def _uncrop(vol_c, p, dtype=np.uint8):
    z0, z1, y0, y1, x0, x1 = p["crop"]
    full = np.zeros(vol_c.shape[:-3] + tuple(p["padded_shape"]), dtype=dtype)
    full[..., z0:z1, y0:y1, x0:x1] = vol_c
    return full[..., 1:-1, 1:-1, 1:-1]


def smart_interpolation(data, labelData, nbrw=10, sorw=4000, allaxis=False, smooth=0,
                        uncertainty=False, return_hits=False, rng="philox", key=PHILOX_KEY):
    """CPU oracle of the drop-in -> result dict (+ '_hits_u32', '_accepted', '_nseeds')."""
    p = prepare(data, labelData)
    flags = orc.plane_flags(p["labidx"], allaxis)
    if allaxis and any(np.any(f[1:] & f[:-1]) for f in flags):
        return None
    pos, lab, var = orc.seeds(p["labidx"], p["img"], flags)
    nol = p["nol"]
    if rng == "philox":
        hits, acc = orc.philox_walk(p["img"], pos, lab, var, nol, nbrw, sorw, key)
    else:
        hits, acc = orc.ref_walk(p["img"], pos, lab, var, nol, nbrw, sorw)
    res = {"_hits_u32": hits, "_accepted": acc, "_nseeds": pos.size, "_prep": p}
    res["regular"] = _uncrop(orc.argmax(hits, p["values"]), p)
    if smooth:
        phi = np.stack([orc.smooth(h.astype(np.float32), smooth) for h in hits])
        res["smooth"] = _uncrop(orc.argmax(phi, p["values"]), p)
    if uncertainty:
        u = orc.uncertainty(hits.astype(np.float32))
        res["uncertainty"] = _uncrop((u * np.float32(255.0)).astype(np.uint8), p)
    if return_hits:
        res["hits"] = _uncrop(hits.astype(np.float32), p, np.float32)
    return res
