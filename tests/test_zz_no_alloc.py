"""bmx_walk_noalloc (the direct Philox walk with its scratch in the library's
static device buffer) against bmx_walk_ex with the direct algorithm, through
the C ABI: identical hits and accepted moves on two calls in a row, the second
on a smaller volume, for uint8 and float32 images; oversized volumes are
rejected before any device work."""
import numpy as np
import pytest

from biomedisa_amd import _hip
from biomedisa_amd.synth import ellipsoids_volume

pytestmark = pytest.mark.gpu


def _case(eng, shape, seed, f32):
    vol = ellipsoids_volume(shape=shape, seed=seed, semi=(4, 9), every=5)
    img = np.ascontiguousarray(np.pad(vol.img, 1))
    if f32:
        img = img.astype(np.float32) * np.float32(0.75) + np.float32(0.25)
    lab = np.ascontiguousarray(np.pad(vol.labels, 1))
    dims = img.shape
    ib, lb = eng.upload(img), eng.upload(lab)
    dt = _hip.IMG_F32 if f32 else _hip.IMG_U8
    flags = eng.plane_flags(lb, dims, False)
    seeds, n = eng.seeds(lb, ib, dt, dims, flags)
    return ib, dt, dims, seeds, n


@pytest.mark.parametrize("f32", [False, True])
def test_noalloc_walk_matches_direct_walk(eng, f32):
    key = 0x1234ABCD5678EF01
    for shape, seed in (((38, 42, 46), 3), ((22, 26, 30), 4)):    # the second call on a smaller volume
        print("noalloc case", shape, "f32" if f32 else "u8", flush=True)
        ib, dt, dims, seeds, n = _case(eng, shape, seed, f32)
        V = int(np.prod(dims))
        nlab, nbrw, sorw = 4, 10, 150
        want = eng.alloc(nlab * V * 4)
        want.zero()
        acc_want = eng.walk(ib, dt, dims, seeds, n, nlab, nbrw, sorw, _hip.RNG_PHILOX, key, want,
                            algorithm=_hip.WALK_DIRECT)
        got = eng.alloc(nlab * V * 4)
        got.zero()
        acc_got = eng.walk_noalloc(ib, dt, dims, seeds, n, nlab, nbrw, sorw, key, got)
        assert acc_got == acc_want > 0
        np.testing.assert_array_equal(got.download((nlab, V), np.uint32), want.download((nlab, V), np.uint32))


def test_noalloc_walk_rejects_oversized_volumes(eng):
    small = eng.alloc(1 << 12)
    hits = eng.alloc(1 << 12)
    # 260^3 uint8 voxels need 8 bytes each, above the 64 MiB buffer: refused by the host checks
    with pytest.raises(_hip.BmxError) as e:
        eng.walk_noalloc(small, _hip.IMG_U8, (260, 260, 260), None, 0, 2, 10, 100, 1, hits)
    assert e.value.code == _hip.BMX_E_INVALID and "maximum" in str(e.value)
