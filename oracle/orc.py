"""ctypes bindings of ``libbmx_oracle.so`` (the C restatement in bmx_oracle.c).

Test infrastructure only: see the package docstring.  Every function takes and
returns NumPy arrays; volumes are C-contiguous (Z, Y, X).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libbmx_oracle.so")
_lib = None

IGN = 0xFF
DTYPE_U8, DTYPE_F32 = 0, 1

_p = C.c_void_p
_i64 = C.c_int64


def build():
    """Compile libbmx_oracle.so in place (gcc, OpenMP, no FP contraction)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_expf.restype = C.c_float
        L.orc_expf.argtypes = [C.c_float]
        L.orc_expf_array.argtypes = [_p, _p, _i64]
        L.orc_philox_array.argtypes = [_p, _p, _p, _i64]
        L.orc_plane_flags.argtypes = [_p, C.c_int, C.c_int, C.c_int, C.c_int, _p, _p, _p]
        L.orc_seeds.restype = _i64
        L.orc_seeds.argtypes = [_p, _p, C.c_int, C.c_int, C.c_int, C.c_int, _p, _p, _p, _p, _p, _p]
        L.orc_philox_walk.restype = _i64
        L.orc_philox_walk.argtypes = [_p, C.c_int, C.c_int, C.c_int, C.c_int, _p, _p, _p, _i64,
                                      C.c_int, C.c_int, C.c_uint64, _p]
        L.orc_ref_walk.restype = _i64
        L.orc_mrg_stream.argtypes = [_i64, C.c_int, _p]
        L.orc_ref_walk.argtypes = [_p, C.c_int, C.c_int, C.c_int, C.c_int, _p, _p, _p, _i64,
                                   C.c_int, C.c_int, _p]
        L.orc_argmax_u32.argtypes = [_p, C.c_int, _i64, _p, _p]
        L.orc_argmax_f32.argtypes = [_p, C.c_int, _i64, _p, _p]
        L.orc_smooth.argtypes = [_p, _p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_uncertainty.argtypes = [_p, C.c_int, C.c_int, C.c_int, C.c_int, _p]
        L.orc_adaptive.argtypes = [_p, C.c_int, C.c_int, C.c_int, _p, _p, _p, _p, _i64, _p]
        L.orc_label6.restype = _i64
        L.orc_label6.argtypes = [_p, C.c_int, C.c_int, C.c_int, _p, _p]
        L.orc_evolution.argtypes = [_p, _p, _p, C.c_int, C.c_int, C.c_int, C.c_double, _p]
        L.orc_erosion.argtypes = [_p, _p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_dilation.argtypes = [_p, _p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_geodis.argtypes = [_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int]
        L.orc_smooth_img.argtypes = [_p, C.c_int, C.c_int, C.c_int, C.c_int, _p]
        L.orc_num_threads.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _img_dtype(img):
    if img.dtype == np.uint8:
        return DTYPE_U8
    if img.dtype == np.float32:
        return DTYPE_F32
    raise TypeError("image must be uint8 or float32, got %s" % img.dtype)


def num_threads():
    return lib().orc_num_threads()


def expf(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    lib().orc_expf_array(_ptr(x), _ptr(y), x.size)
    return y


def philox4x32_10(ctr, key):
    """ctr: (..., 4) uint32, key: (..., 2) uint32 -> (..., 4) uint32."""
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32)
    key = np.ascontiguousarray(np.broadcast_to(key, ctr.shape[:-1] + (2,)), dtype=np.uint32)
    out = np.empty_like(ctr)
    n = ctr.size // 4
    lib().orc_philox_array(_ptr(ctr), _ptr(key), _ptr(out), n)
    return out


def _lab16(labidx):
    """The oracle's uint16 label-index volume: uint16 as given (0xFFFF =
    ignored), uint8 with 0xFF as its ignored index."""
    labidx = np.asarray(labidx)
    if labidx.dtype == np.uint16:
        return np.ascontiguousarray(labidx)
    lab8 = np.ascontiguousarray(labidx, dtype=np.uint8)
    out = lab8.astype(np.uint16)
    out[lab8 == 0xFF] = 0xFFFF
    return out


def plane_flags(labidx, allaxis):
    """Labelled planes of a label-index volume (0 = background, ignored index:
    0xFF for uint8, 0xFFFF for uint16 volumes)."""
    labidx = _lab16(labidx)
    Z, Y, X = labidx.shape
    fz, fy, fx = np.zeros(Z, np.uint8), np.zeros(Y, np.uint8), np.zeros(X, np.uint8)
    lib().orc_plane_flags(_ptr(labidx), Z, Y, X, int(bool(allaxis)), _ptr(fz), _ptr(fy), _ptr(fx))
    return fz, fy, fx


def seeds(labidx, img, flags):
    """Seed list (pos int64, label index uint8, var float32), ascending pos."""
    labidx = _lab16(labidx)
    img = np.ascontiguousarray(img)
    Z, Y, X = labidx.shape
    fz, fy, fx = [np.ascontiguousarray(f, dtype=np.uint8) for f in flags]
    L = lib()
    n = L.orc_seeds(_ptr(labidx), _ptr(img), _img_dtype(img), Z, Y, X,
                    _ptr(fz), _ptr(fy), _ptr(fx), None, None, None)
    pos = np.empty(n, np.int64)
    lab = np.empty(n, np.uint8)
    var = np.empty(n, np.float32)
    L.orc_seeds(_ptr(labidx), _ptr(img), _img_dtype(img), Z, Y, X,
                _ptr(fz), _ptr(fy), _ptr(fx), _ptr(pos), _ptr(lab), _ptr(var))
    return pos, lab, var


def label6(mask):
    """6-connected components -> (int32 labels numbered in raster order of each
    component's first voxel, count)."""
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    Z, Y, X = mask.shape
    labels = np.empty(mask.shape, np.int32)
    queue = np.empty(mask.size, np.int64)
    n = lib().orc_label6(_ptr(mask), Z, Y, X, _ptr(labels), _ptr(queue))
    return labels, int(n)


def evolution(mean, A, data, alpha):
    mean = np.ascontiguousarray(mean, np.float32)
    A = np.ascontiguousarray(A, np.uint8)
    data = np.ascontiguousarray(data, np.uint8)
    B = np.empty_like(A)
    lib().orc_evolution(_ptr(mean), _ptr(A), _ptr(data), *A.shape, float(alpha), _ptr(B))
    return B


def erosion(start, final, label):
    """Reads start, writes into final (in place), as curvop_numba.erosion."""
    lib().orc_erosion(_ptr(start), _ptr(final), *start.shape, int(label))


def dilation(start, final, label):
    lib().orc_dilation(_ptr(start), _ptr(final), *start.shape, int(label))


def geodis(c, sqrt2, sqrt3, iterations):
    """In place on a float32 volume (active_contour.py:42-89)."""
    assert c.dtype == np.float32 and c.flags.c_contiguous
    lib().orc_geodis(_ptr(c), *c.shape, float(sqrt2), float(sqrt3), int(iterations))
    return c


def smooth_img(img):
    """smooth_img_3x3 (uint8 or float32)."""
    img = np.ascontiguousarray(img)
    out = np.empty_like(img)
    lib().orc_smooth_img(_ptr(img), 0 if img.dtype == np.uint8 else 1, *img.shape, _ptr(out))
    return out


def adaptive(labidx, flags, pos):
    """The large-volume kernel's `found` test per seed (bool array)."""
    labidx = _lab16(labidx)
    Z, Y, X = labidx.shape
    fz, fy, fx = [np.ascontiguousarray(f, dtype=np.uint8) for f in flags]
    pos = np.ascontiguousarray(pos, np.int64)
    out = np.zeros(pos.size, np.uint8)
    lib().orc_adaptive(_ptr(labidx), Z, Y, X, _ptr(fz), _ptr(fy), _ptr(fx), _ptr(pos), pos.size, _ptr(out))
    return out.astype(bool)


def philox_walk(img, pos, lab, var, nlabels, nbrw, sorw, key):
    """Philox-stream walker -> (hits uint32 (L,Z,Y,X), accepted moves)."""
    img = np.ascontiguousarray(img)
    Z, Y, X = img.shape
    hits = np.zeros((nlabels, Z, Y, X), np.uint32)
    pos = np.ascontiguousarray(pos, np.int64)
    lab = np.ascontiguousarray(lab, np.uint8)
    var = np.ascontiguousarray(var, np.float32)
    acc = lib().orc_philox_walk(_ptr(img), _img_dtype(img), Z, Y, X, _ptr(pos), _ptr(lab),
                                _ptr(var), pos.size, int(nbrw), int(sorw),
                                C.c_uint64(int(key) & 0xFFFFFFFFFFFFFFFF), _ptr(hits))
    return hits, acc


def mrg_stream(index, n):
    """The first n uniforms of the reference stream seeded with voxel `index`."""
    u = np.empty(int(n), np.float32)
    lib().orc_mrg_stream(C.c_int64(int(index)), int(n), _ptr(u))
    return u


def ref_walk(img, pos, lab, var, nlabels, nbrw, sorw):
    """Reference-stream walker (float MRG recurrence) -> (hits, accepted)."""
    img = np.ascontiguousarray(img)
    Z, Y, X = img.shape
    hits = np.zeros((nlabels, Z, Y, X), np.uint32)
    pos = np.ascontiguousarray(pos, np.int64)
    lab = np.ascontiguousarray(lab, np.uint8)
    var = np.ascontiguousarray(var, np.float32)
    acc = lib().orc_ref_walk(_ptr(img), _img_dtype(img), Z, Y, X, _ptr(pos), _ptr(lab),
                             _ptr(var), pos.size, int(nbrw), int(sorw), _ptr(hits))
    return hits, acc


def argmax(hits, values):
    """First-maximum argmax over axis 0, mapped through ``values``."""
    hits = np.ascontiguousarray(hits)
    values = np.ascontiguousarray(values, dtype=np.uint8)
    L = hits.shape[0]
    V = hits[0].size
    out = np.empty(hits.shape[1:], np.uint8)
    if hits.dtype == np.uint32:
        lib().orc_argmax_u32(_ptr(hits), L, V, _ptr(values), _ptr(out))
    elif hits.dtype == np.float32:
        lib().orc_argmax_f32(_ptr(hits), L, V, _ptr(values), _ptr(out))
    else:
        raise TypeError(hits.dtype)
    return out


def smooth(phi, iters):
    """Curvature smoothing of one float32 volume (in a copy)."""
    phi = np.array(phi, dtype=np.float32, order="C", copy=True)
    curv = np.zeros_like(phi)
    Z, Y, X = phi.shape
    lib().orc_smooth(_ptr(phi), _ptr(curv), Z, Y, X, int(iters))
    return phi


def uncertainty(hits):
    """Top-3 uncertainty of float32 hits (L,Z,Y,X) -> float32 (Z,Y,X)."""
    hits = np.ascontiguousarray(hits, dtype=np.float32)
    L, Z, Y, X = hits.shape
    out = np.empty((Z, Y, X), np.float32)
    lib().orc_uncertainty(_ptr(hits), L, Z, Y, X, _ptr(out))
    return out
