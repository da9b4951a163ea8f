"""CPU restatement of the whole smart-interpolation call (test infrastructure).

Follows biomedisa/interpolation.py:135-204 (dtype handling, pad, bounding-box
crop), biomedisa_helper.py:639-724 (label set), the backend flow of
pyopencl_small.py:80-135 and rw_small.py:114-201 (walk, smoothing,
uncertainty, argmax, get_labels), and for path='large' that of
pyopencl_large.py:243-496 with rw_large.py:40-123 on one GPU (planes,
adaptive seeds, running maxima).  Heavy loops run in libbmx_oracle.so.
Written independently of biomedisa_amd so that the two can check each other.
"""
import numpy as np

from . import orc
from . import post

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


def prepare(data, labelData, denoise=False):
    """-> dict(img, labidx, values, crop, padded_shape, nol) exactly as the walker
    sees them; denoise applies smooth_img_3x3 after the dtype handling
    (interpolation.py:149-151)."""
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
    if denoise:
        img = orc.smooth_img(img)
    img = np.pad(img, 1)
    labels = np.pad(labels, 1)
    z0, z1, y0, y1, x0, x1 = _bbox(labels)
    img_c = np.ascontiguousarray(img[z0:z1, y0:y1, x0:x1])
    lab_c = np.ascontiguousarray(labels[z0:z1, y0:y1, x0:x1])
    # uint16 label indices (0xFFFF: negative, ignored): the reference switches to
    # short labels with -1 as ignore value for 256 labels (pyopencl_small.py:90-95)
    labidx = np.full(lab_c.shape, 0xFFFF, np.uint16)
    for i, v in enumerate(vals):
        labidx[lab_c == v] = i
    return dict(img=img_c, labidx=labidx, lab=lab_c, values=vals.astype(np.uint8),
                crop=(z0, z1, y0, y1, x0, x1), padded_shape=img.shape, nol=len(vals))


def _thin_planes_positive(arr):
    """read_labeled_slices_allx_large (biomedisa_helper.py:1082-1103): indices of
    the planes of arr (axis 0) holding a positive label that survives the removal
    of 1-voxel-thin lines along axis 1, then axis 2."""
    ones = (arr > 0).astype(np.int8)
    grad = np.zeros_like(ones)
    t = np.abs(ones[:, :-1] - ones[:, 1:])
    grad[:, :-1] += t
    grad[:, 1:] += t
    ones[grad == 2] = 0
    grad.fill(0)
    t = np.abs(ones[:, :, :-1] - ones[:, :, 1:])
    grad[:, :, :-1] += t
    grad[:, :, 1:] += t
    ones[grad == 2] = 0
    return [k for k in range(arr.shape[0]) if np.any(ones[k])]


def plane_flags_large(lab, allaxis):
    """Labelled planes of the large-volume path on the label VALUE volume `lab`
    (negative = ignored): XY planes by read_labeled_slices_large
    (biomedisa_helper.py:1068-1080: a labelled plane hides the next 4), allaxis
    planes by read_labeled_slices_allx_large on the volume and its two swapped
    views (rw_large.py:67-99)."""
    Z, Y, X = lab.shape
    fz, fy, fx = np.zeros(Z, np.uint8), np.zeros(Y, np.uint8), np.zeros(X, np.uint8)
    if not allaxis:
        i = 0
        while i < Z:
            if np.any(lab[i]):
                fz[i] = 1
                i += 5
            else:
                i += 1
        return fz, fy, fx
    fz[_thin_planes_positive(lab)] = 1
    t = np.ascontiguousarray(np.swapaxes(lab, 0, 1))
    fy[_thin_planes_positive(t)] = 1
    t = np.ascontiguousarray(np.swapaxes(t, 0, 2))
    fx[_thin_planes_positive(t)] = 1
    return fz, fy, fx


def select_path(p):
    """interpolation.py:192-196: the small path when the cropped box's float32
    volume times the label count is below 1e10 bytes and the volume below 2e9."""
    nbytes = int(np.prod(p["img"].shape)) * 4
    return "small" if nbytes * p["nol"] < 1e10 and nbytes < 2e9 else "large"


def _uncrop(vol_c, p, dtype=np.uint8):
    z0, z1, y0, y1, x0, x1 = p["crop"]
    full = np.zeros(vol_c.shape[:-3] + tuple(p["padded_shape"]), dtype=dtype)
    full[..., z0:z1, y0:y1, x0:x1] = vol_c
    return full[..., 1:-1, 1:-1, 1:-1]


def smart_interpolation(data, labelData, nbrw=10, sorw=4000, allaxis=False, smooth=0,
                        uncertainty=False, return_hits=False, rng="philox", key=PHILOX_KEY, path="small",
                        clean=None, fill=None, acwe=False, acwe_alpha=1.0, acwe_smooth=1, acwe_steps=3,
                        denoise=False):
    """CPU oracle of the drop-in -> result dict (+ '_hits_u32', '_accepted', '_nseeds').
    path: 'small', 'large' or 'auto' (the reference's size rule)."""
    p = prepare(data, labelData, denoise)
    flags = orc.plane_flags(p["labidx"], allaxis)
    if allaxis and any(np.any(f[1:] & f[:-1]) for f in flags):
        return None
    if path == "auto":
        path = select_path(p)
    if path == "large":
        flags = plane_flags_large(p["lab"], allaxis)
    pos, lab, var = orc.seeds(p["labidx"], p["img"], flags)
    if path == "large":
        keep = orc.adaptive(p["labidx"], flags, pos)
        pos, lab, var = pos[keep], lab[keep], var[keep]
    nol = p["nol"]
    if rng == "philox":
        hits, acc = orc.philox_walk(p["img"], pos, lab, var, nol, nbrw, sorw, key)
    else:
        hits, acc = orc.ref_walk(p["img"], pos, lab, var, nol, nbrw, sorw)
    res = {"_hits_u32": hits, "_accepted": acc, "_nseeds": pos.size, "_prep": p, "_path": path}
    res["regular"] = _uncrop(orc.argmax(hits, p["values"]), p)
    if smooth:
        phi = np.stack([orc.smooth(h.astype(np.float32), smooth) for h in hits])
        if path == "large":
            phi[0][phi[0] < 0] = 0   # pyopencl_large.py:466
        res["smooth"] = _uncrop(orc.argmax(phi, p["values"]), p)
    if uncertainty:
        u = orc.uncertainty(hits.astype(np.float32))
        res["uncertainty"] = _uncrop((u * np.float32(255.0)).astype(np.uint8), p)
    if return_hits:
        res["hits"] = _uncrop(hits.astype(np.float32), p, np.float32)
    # rw_small.py:222-241
    if clean:
        res["cleaned"] = post.clean(res["regular"], clean)
        if smooth:
            res["smooth_cleaned"] = post.clean(res["smooth"], clean)
    if fill:
        res["filled"] = post.fill(res["regular"], fill)
    if clean and fill:
        res["cleaned_filled"] = post.cleaned_filled(res["regular"], res["cleaned"], res["filled"])
    # rw_small.py:244-253
    if acwe:
        z0, z1, y0, y1, x0, x1 = p["crop"]
        data = np.zeros(p["padded_shape"], dtype=p["img"].dtype)
        data[z0:z1, y0:y1, x0:x1] = p["img"]
        res["acwe"] = post.acwe(data[1:-1, 1:-1, 1:-1], res["regular"], acwe_alpha, acwe_smooth, acwe_steps)
        res["refined"] = post.refinement(data[1:-1, 1:-1, 1:-1], res["regular"])
    return res
