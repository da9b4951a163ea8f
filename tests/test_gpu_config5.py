"""The large-volume configuration at reduced size: a 192^3 uint16 volume with
12-bit values (normalised to float32 in [0, 255] as interpolation.py:140-144
does), 16 structures, noise sigma 40, bias +-200, labels on every 16th plane
of all three axes.  Plain, surrogate-guided and the reference's large-volume
driver (path='large') on the same GPU; the guided kernels bit for bit against
oracle/guide.py on the float image; per-label Dice of guided and of the large
driver against plain."""
import numpy as np
import pytest

from biomedisa_amd import _hip
from biomedisa_amd import helpers as H
from oracle import guide as G

pytestmark = pytest.mark.gpu

SIZE = 192


@pytest.fixture(scope="module")
def case(eng):
    from biomedisa_amd import surrogate as S
    from biomedisa_amd.interpolation import _prepare_device, run_pipeline, run_pipeline_guided
    from biomedisa_amd.synth import large_volume
    vol = large_volume(seed=9, size=SIZE, every=16)
    assert vol.img.dtype == np.uint16 and vol.img.max() <= 4095
    prep = _prepare_device(eng, vol.img, vol.labels, 10, 4000, False, None, True, "none", "all", 0, False, False)
    dv, nol, values = prep.vol, prep.bm.nol, prep.values
    assert dv.img_dtype == _hip.IMG_F32 and 12 <= nol <= 17
    print("config-5 case: %s float32, %d labels" % (dv.dims, nol), flush=True)
    p = run_pipeline(dv, nol, values, nbrw=10, sorw=4000, allaxis=True)
    g = run_pipeline_guided(dv, nol, values, S.net(0), nbrw=10, sorw=4000, allaxis=True, keep=True,
                            plain_above=None)
    lp = run_pipeline(dv, nol, values, nbrw=10, sorw=4000, allaxis=True, path="large")
    dl = lambda b: b.download(dv.dims, np.uint8)  # noqa: E731
    return dict(dv=dv, nol=nol, values=values, p=p, g=g, lp=lp, dl=dl)


def test_guided_and_large_driver_agree_with_plain(case):
    dl, p, g, lp = case["dl"], case["p"], case["g"], case["lp"]
    preg = dl(p["regular"])
    d = H.per_label_dice(dl(g["regular"]), preg)
    dlarge = H.per_label_dice(dl(lp["regular"]), preg)
    print("guided: %.3f of seeds walk all walks, %.3f ring; min Dice %.5f; large driver: %.3f walk, min Dice %.5f"
          % (g["walked_seeds"] / g["nseeds"], g["ring_seeds"] / g["nseeds"], min(d.values()),
             lp["walked_seeds"] / lp["nseeds"], min(dlarge.values())), flush=True)
    assert len(d) == case["nol"] - 1
    assert min(d.values()) >= 0.99, d
    assert min(dlarge.values()) >= 0.99, dlarge
    assert g["walked_seeds"] < g["nseeds"]


def test_guided_kernels_bitexact_on_float_image(case, eng):
    dv, dl, g, values = case["dv"], case["dl"], case["g"], case["values"]
    Z, Y, X = dv.dims
    # network input of the float image (one slab's worth: the first 24 planes)
    img = dv.img.download(dv.dims, np.float32)
    lab = dv.lab.download(dv.dims, np.uint8)
    flags = g["flags"]
    nz = 24
    grid = (nz, (Y + 7) // 8 * 8, (X + 7) // 8 * 8)
    fs = np.concatenate([flags[:nz], flags[Z:Z + Y], flags[Z + Y:]])
    out = eng.alloc(int(np.prod(grid)) * G.N_CH * 2)
    _hip.call("bmx_net_input", eng.ctx, dv.img.ptr, _hip.IMG_F32, dv.lab.ptr, eng.upload(fs).ptr, nz, Y, X,
              *grid, out.ptr)
    np.testing.assert_array_equal(out.download(grid + (G.N_CH,), np.uint16),
                                  G.net_input(img[:nz], lab[:nz], fs, grid))
    # uncertain set, merge and the smoothing's label selection on this volume
    pred = dl(g["pred"])
    hits = g["hits"].download((case["nol"],) + dv.dims, np.uint32)
    U = dl(g["U"]).astype(bool)
    np.testing.assert_array_equal(dl(g["regular"]), G.merge_guided(hits, values, pred, U))
    # the proposal's masks from its own prediction: U = unc | boundary band, need = its dilation
    unc = eng.alloc(dv.V)
    pb = eng.upload(pred)
    U1, need, tmp = eng.alloc(dv.V), eng.alloc(dv.V), eng.alloc(dv.V)
    uflag = (U & ~G.dilate(G.boundary(pred), 4).astype(bool)).astype(np.uint8)   # flagged-uncertain part of U
    eng.upload_into(unc, uflag)
    _hip.call("bmx_guided_masks", eng.ctx, pb.ptr, unc.ptr, Z, Y, X, 4, 6, U1.ptr, need.ptr, tmp.ptr)
    u_ref, need_ref = G.guided_masks(pred, uflag, 4, 6)
    np.testing.assert_array_equal(dl(U1), u_ref)
    np.testing.assert_array_equal(dl(need), need_ref)
