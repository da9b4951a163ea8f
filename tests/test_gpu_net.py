"""Learned-proposal kernels (bmx_net.hip) against the float64 oracle (oracle/net.py).

Inputs are bf16 values; the GPU accumulates in FP32 and rounds outputs to
bf16, so outputs agree to bf16 resolution (relative 2^-8) and the FP32 weight
gradients to 1e-3 of their scale.
"""
import ctypes as C

import numpy as np
import pytest

from biomedisa_amd import _hip
from oracle import net as O

pytestmark = pytest.mark.gpu


def up_bf16(eng, a):
    return eng.upload(O.to_bf16_bits(a))


def down_bf16(buf, shape):
    return O.from_bf16_bits(buf.download(shape, np.uint16)).astype(np.float64)


def _close_bf16(got, ref):
    scale = np.abs(ref).max() + 1e-12
    err = np.abs(got - ref)
    assert np.all(err <= np.abs(ref) * 2.0 ** -7 + scale * 2.0 ** -12), (err.max(), scale)


SHAPES = [(2, 4, 6, 8), (1, 3, 5, 7)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("ci,co", [(16, 16), (32, 32), (16, 64), (48, 16)])
@pytest.mark.parametrize("relu", [0, 1])
def test_conv_fwd(eng, shape, ci, co, relu):
    rng = np.random.default_rng(ci * 100 + co + relu)
    x = O.bf16_round(rng.standard_normal(shape + (ci,)))
    w = O.bf16_round(rng.standard_normal((co, 27, ci)) * 0.1)
    b = rng.standard_normal(co).astype(np.float32)
    xb, wb, bb = up_bf16(eng, x), up_bf16(eng, w), eng.upload(b)
    yb = eng.alloc(int(np.prod(shape)) * co * 2)
    _hip.call("bmx_conv3d_fwd", eng.ctx, xb.ptr, *shape, ci, wb.ptr, bb.ptr, yb.ptr, co, relu)
    ref = O.conv3d(x, w, b.astype(np.float64), relu=bool(relu))
    _close_bf16(down_bf16(yb, shape + (co,)), ref)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("ci,co", [(16, 16), (32, 64), (64, 32), (96, 32)])
@pytest.mark.parametrize("masked", [False, True])
def test_conv_dgrad(eng, shape, ci, co, masked):
    rng = np.random.default_rng(ci + co * 7 + masked)
    dy = O.bf16_round(rng.standard_normal(shape + (co,)))
    w = O.bf16_round(rng.standard_normal((co, 27, ci)) * 0.1)
    xact = O.bf16_round(np.maximum(rng.standard_normal(shape + (ci,)), 0))
    dyb, wb, xb = up_bf16(eng, dy), up_bf16(eng, w), up_bf16(eng, xact)
    dxb = eng.alloc(int(np.prod(shape)) * ci * 2)
    scratch = eng.alloc(co * 27 * ci * 2)
    _hip.call("bmx_conv3d_dgrad", eng.ctx, dyb.ptr, *shape, co, wb.ptr, xb.ptr if masked else None,
              dxb.ptr, ci, scratch.ptr)
    ref = O.conv3d_dgrad(dy, w)
    if masked:
        ref = ref * (xact > 0)
    _close_bf16(down_bf16(dxb, shape + (ci,)), ref)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("ci,co", [(16, 16), (32, 64), (48, 16)])
def test_conv_wgrad(eng, shape, ci, co):
    rng = np.random.default_rng(ci * 3 + co)
    x = O.bf16_round(rng.standard_normal(shape + (ci,)))
    dy = O.bf16_round(rng.standard_normal(shape + (co,)))
    xb, dyb = up_bf16(eng, x), up_bf16(eng, dy)
    dwb = eng.alloc(co * 27 * ci * 4)
    dbb = eng.alloc(co * 4)
    dwb.zero()
    dbb.zero()
    _hip.call("bmx_conv3d_wgrad", eng.ctx, xb.ptr, *shape, ci, dyb.ptr, co, dwb.ptr, dbb.ptr)
    dw, db = O.conv3d_wgrad(x, dy)
    got = dwb.download((co, 27, ci), np.float32)
    np.testing.assert_allclose(got, dw, rtol=1e-4, atol=1e-4 * np.abs(dw).max())
    np.testing.assert_allclose(dbb.download((co,), np.float32), db, rtol=1e-4, atol=1e-3)


@pytest.mark.parametrize("shape", [(1, 9, 13, 70), (2, 40, 6, 33)])
@pytest.mark.parametrize("ci,co", [(16, 32), (64, 128), (192, 64)])
def test_conv_wgrad_tiles(eng, shape, ci, co):
    """Weight gradient over several (x, y) tiles and z chunks of a block (x
    past 32 and 64, y past 4, depth 40 split into chunks), every output-channel
    tiling (16 x 2 and 16 x 4 per block, two output-channel blocks) and twelve
    input-channel blocks."""
    rng = np.random.default_rng(ci * 5 + co + shape[3])
    x = O.bf16_round(rng.standard_normal(shape + (ci,)))
    dy = O.bf16_round(rng.standard_normal(shape + (co,)))
    xb, dyb = up_bf16(eng, x), up_bf16(eng, dy)
    dwb = eng.alloc(co * 27 * ci * 4)
    dbb = eng.alloc(co * 4)
    dwb.zero()
    dbb.zero()
    _hip.call("bmx_conv3d_wgrad", eng.ctx, xb.ptr, *shape, ci, dyb.ptr, co, dwb.ptr, dbb.ptr)
    dw, db = O.conv3d_wgrad(x, dy)
    got = dwb.download((co, 27, ci), np.float32)
    np.testing.assert_allclose(got, dw, rtol=1e-4, atol=1e-4 * np.abs(dw).max())
    np.testing.assert_allclose(dbb.download((co,), np.float32), db, rtol=1e-4, atol=1e-3)


def test_pool_up_concat_split(eng):
    rng = np.random.default_rng(9)
    shape = (2, 4, 6, 8)
    x = O.bf16_round(rng.standard_normal(shape + (16,)))
    xb = up_bf16(eng, x)
    pb = eng.alloc(x.size // 8 * 2)
    _hip.call("bmx_pool2", eng.ctx, xb.ptr, *shape, 16, pb.ptr)
    _close_bf16(down_bf16(pb, (2, 2, 3, 4, 16)), O.pool2(x))
    g = O.bf16_round(rng.standard_normal((2, 2, 3, 4, 16)))
    gb = up_bf16(eng, g)
    gx = eng.alloc(x.size * 2)
    _hip.call("bmx_pool2_bwd", eng.ctx, gb.ptr, *shape, 16, gx.ptr)
    _close_bf16(down_bf16(gx, shape + (16,)), O.pool2_bwd(g))
    ub = eng.alloc(x.size * 2)
    _hip.call("bmx_up2", eng.ctx, gb.ptr, 2, 2, 3, 4, 16, ub.ptr)
    _close_bf16(down_bf16(ub, shape + (16,)), O.up2(g))
    db = eng.alloc(g.size * 2)
    _hip.call("bmx_up2_bwd", eng.ctx, xb.ptr, 2, 2, 3, 4, 16, db.ptr)
    _close_bf16(down_bf16(db, g.shape), O.up2_bwd(x))
    nv = int(np.prod(shape))
    y = O.bf16_round(rng.standard_normal(shape + (32,)))
    yb = up_bf16(eng, y)
    cb = eng.alloc(nv * 48 * 2)
    _hip.call("bmx_concat", eng.ctx, xb.ptr, 16, yb.ptr, 32, nv, cb.ptr)
    cat = down_bf16(cb, shape + (48,))
    np.testing.assert_array_equal(cat, np.concatenate([x, y], -1))
    a2, b2 = eng.alloc(nv * 16 * 2), eng.alloc(nv * 32 * 2)
    _hip.call("bmx_split", eng.ctx, cb.ptr, nv, a2.ptr, 16, b2.ptr, 32, 0)
    np.testing.assert_array_equal(down_bf16(a2, shape + (16,)), x)
    np.testing.assert_array_equal(down_bf16(b2, shape + (32,)), y)
    # split accumulating into a, and the ReLU mask / bf16 add, 16 bytes at a time
    _hip.call("bmx_split", eng.ctx, cb.ptr, nv, a2.ptr, 16, b2.ptr, 32, 1)
    np.testing.assert_array_equal(down_bf16(a2, shape + (16,)), O.bf16_round(x + x))
    gm = up_bf16(eng, y)
    _hip.call("bmx_relu_mask", eng.ctx, gm.ptr, xb.ptr, nv * 16)
    np.testing.assert_array_equal(down_bf16(gm, shape + (16,)).reshape(-1),
                                  np.where(x.reshape(-1) > 0, y.reshape(-1)[:nv * 16], 0.0))
    s1 = up_bf16(eng, x)
    _hip.call("bmx_add_bf16", eng.ctx, s1.ptr, yb.ptr, nv * 16)
    np.testing.assert_array_equal(down_bf16(s1, shape + (16,)), O.bf16_round(x + y.reshape(-1)[:nv * 16].reshape(x.shape)))


def test_concat_split_odd_channels(eng):
    """Channel counts that are not multiples of 8 take the element-wise kernels."""
    rng = np.random.default_rng(19)
    nv = 37
    a = O.bf16_round(rng.standard_normal((nv, 12)))
    b = O.bf16_round(rng.standard_normal((nv, 5)))
    ab, bb = up_bf16(eng, a), up_bf16(eng, b)
    cb = eng.alloc(nv * 17 * 2)
    _hip.call("bmx_concat", eng.ctx, ab.ptr, 12, bb.ptr, 5, nv, cb.ptr)
    np.testing.assert_array_equal(down_bf16(cb, (nv, 17)), np.concatenate([a, b], -1))
    a2, b2 = up_bf16(eng, a), eng.alloc(nv * 5 * 2)
    _hip.call("bmx_split", eng.ctx, cb.ptr, nv, a2.ptr, 12, b2.ptr, 5, 1)
    np.testing.assert_array_equal(down_bf16(a2, (nv, 12)), O.bf16_round(a + a))
    np.testing.assert_array_equal(down_bf16(b2, (nv, 5)), b)


@pytest.mark.parametrize("L,index_targets", [(9, False), (30, False), (30, True)])
def test_head_softmax_ce(eng, L, index_targets):
    rng = np.random.default_rng(10 + L)
    nv = 300
    x = O.bf16_round(rng.standard_normal((nv, 16)))
    W = (rng.standard_normal((L, 16)) * 0.3).astype(np.float32)
    b = rng.standard_normal(L).astype(np.float32)
    if index_targets:
        ti = rng.integers(0, L, nv).astype(np.uint8)
        t = np.eye(L)[ti]
    else:
        t = rng.random((nv, L)); t /= t.sum(-1, keepdims=True)
    vm = (rng.random(nv) > 0.2).astype(np.uint8)
    xb, Wb, bb = up_bf16(eng, x), eng.upload(W), eng.upload(b)
    tb, mb = (None, eng.upload(vm)) if index_targets else (eng.upload(t.astype(np.float32)), eng.upload(vm))
    tib = eng.upload(ti) if index_targets else None
    pb, dxb = eng.alloc(nv * L * 4), eng.alloc(nv * 16 * 2)
    dWb, dbb, lb = eng.alloc(L * 16 * 4), eng.alloc(L * 4), eng.alloc(4)
    for bf in (dWb, dbb, lb):
        bf.zero()
    scale = 1.0 / nv
    _hip.call("bmx_head", eng.ctx, xb.ptr, nv, L, Wb.ptr, bb.ptr, pb.ptr, None if tb is None else tb.ptr,
              None if tib is None else tib.ptr, mb.ptr, C.c_float(scale), dxb.ptr, dWb.ptr, dbb.ptr, lb.ptr)
    p, loss, dx, dW, db = O.head(x, W.astype(np.float64), b.astype(np.float64),
                                 t.astype(np.float32).astype(np.float64), vm, scale)
    np.testing.assert_allclose(pb.download((nv, L), np.float32), p, rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(lb.download((1,), np.float32)[0], loss, rtol=1e-4)
    _close_bf16(down_bf16(dxb, (nv, 16)), dx)
    np.testing.assert_allclose(dWb.download((L, 16), np.float32), dW, rtol=1e-3, atol=1e-6)
    np.testing.assert_allclose(dbb.download((L,), np.float32), db, rtol=1e-3, atol=1e-6)


def test_adam_step(eng):
    rng = np.random.default_rng(11)
    n = 1000
    p = rng.standard_normal(n).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    pb, gb = eng.upload(p), eng.upload(g)
    mb, vb = eng.alloc(n * 4), eng.alloc(n * 4)
    mb.zero(); vb.zero()
    bf = eng.alloc(n * 2)
    _hip.call("bmx_adam", eng.ctx, pb.ptr, gb.ptr, mb.ptr, vb.ptr, bf.ptr, n, C.c_float(1e-2),
              C.c_float(0.9), C.c_float(0.999), C.c_float(1e-8), 1, C.c_float(0.0))
    m = 0.1 * g
    v = 0.001 * g * g
    ref = p - 1e-2 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    got = pb.download((n,), np.float32)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(bf.download((n,), np.uint16), O.to_bf16_bits(got))


def test_whole_network_gradients_match_oracle(eng):
    """Loss and every parameter gradient of the GPU encoder-decoder (forward,
    head, backward through pooling, upsampling, skips and ReLUs) against the
    float64 oracle network with bf16 rounding at the same points."""
    from biomedisa_amd import surrogate as S
    net = S.Net(eng, seed=3)
    N, D, H, W = 1, 8, 8, 16
    rng = np.random.default_rng(12)
    x = O.bf16_round(np.concatenate([rng.standard_normal((N, D, H, W, 2)),
                                     (rng.random((N, D, H, W, S.CIN - 2)) < 0.1)], -1))
    t = rng.random((N, D, H, W, S.L_MAX)); t /= t.sum(-1, keepdims=True); t = t.astype(np.float32)
    vm = (rng.random((N, D, H, W)) > 0.1).astype(np.uint8)
    xb, tb, mb, lb = up_bf16(eng, x), eng.upload(t), eng.upload(vm), eng.alloc(4)
    net.zero_grad()
    lb.zero()
    net.loss_and_grad(xb, tb, mb, N, D, H, W, lb)
    params = net.state_dict()
    loss, G = O.unet_loss_and_grads(params, x, t.astype(np.float64), vm)
    np.testing.assert_allclose(lb.download((1,), np.float32)[0], loss, rtol=2e-2)
    grads = {}
    for name, _, _ in S.LAYERS:
        w, b, _, _ = net.conv[name]
        grads[name + ".w"] = w.g.download(w.shape, np.float32)
        grads[name + ".b"] = b.g.download(b.shape, np.float32)
    grads["head.W"] = net.hW.g.download(net.hW.shape, np.float32)
    grads["head.b"] = net.hb.g.download(net.hb.shape, np.float32)
    for k, ref in G.items():
        got = grads[k].astype(np.float64)
        cos = (got * ref).sum() / (np.linalg.norm(got) * np.linalg.norm(ref) + 1e-30)
        rel = np.linalg.norm(got - ref) / (np.linalg.norm(ref) + 1e-30)
        assert cos > 0.995 and rel < 0.1, (k, cos, rel)


def test_concat_past_2pow32_elements(eng):
    """Element-wise kernels stride over ranges beyond 2^32 elements (one dispatch
    covers fewer work-items): upsample two tensors on the device and concatenate
    them into 4.4e9 elements, then check the tail of the result."""
    small = (1, 112, 128, 800)
    nsmall = int(np.prod(small))
    nbig = nsmall * 8
    assert nbig * 48 > 2 ** 32
    rng = np.random.default_rng(13)
    xa = rng.integers(0, 65536, (nsmall, 16), dtype=np.uint16)
    xb = rng.integers(0, 65536, (nsmall, 32), dtype=np.uint16)
    ba, bb = eng.upload(xa), eng.upload(xb)
    ua, ub = eng.alloc(nbig * 16 * 2), eng.alloc(nbig * 32 * 2)
    _hip.call("bmx_up2", eng.ctx, ba.ptr, *small, 16, ua.ptr)
    _hip.call("bmx_up2", eng.ctx, bb.ptr, *small, 32, ub.ptr)
    del ba, bb
    y = eng.alloc(nbig * 48 * 2)
    _hip.call("bmx_concat", eng.ctx, ua.ptr, 16, ub.ptr, 32, nbig, y.ptr)
    ntail = 1 << 16
    tail = np.empty((ntail, 48), np.uint16)
    _hip._check(_hip.load().bmx_download(eng.ctx, tail.ctypes.data, y.ptr + (nbig - ntail) * 96, tail.nbytes))
    v = np.arange(nbig - ntail, nbig)
    _, D, H, W = small
    xx, t = v % (2 * W), v // (2 * W)
    yy, zz = t % (2 * H), t // (2 * H)
    src = ((zz // 2) * H + yy // 2) * W + xx // 2
    np.testing.assert_array_equal(tail[:, :16], xa[src])
    np.testing.assert_array_equal(tail[:, 16:], xb[src])


@pytest.mark.parametrize("ca,cb,co", [(16, 32, 16), (32, 64, 32), (64, 128, 64)])
def test_conv_fwd_cat_equals_materialised(eng, ca, cb, co):
    """bmx_conv3d_fwd_cat(a, b) == bmx_conv3d_fwd(concat(a, up2(b))), bit for bit."""
    rng = np.random.default_rng(ca + cb)
    shape = (2, 4, 6, 8)
    half = (2, 2, 3, 4)
    a = O.bf16_round(rng.standard_normal(shape + (ca,)))
    b = O.bf16_round(rng.standard_normal(half + (cb,)))
    w = O.bf16_round(rng.standard_normal((co, 27, ca + cb)) * 0.1)
    bias = rng.standard_normal(co).astype(np.float32)
    ab, bb, wb, biasb = up_bf16(eng, a), up_bf16(eng, b), up_bf16(eng, w), eng.upload(bias)
    nv = int(np.prod(shape))
    ub, cb_buf = eng.alloc(nv * cb * 2), eng.alloc(nv * (ca + cb) * 2)
    _hip.call("bmx_up2", eng.ctx, bb.ptr, *half, cb, ub.ptr)
    _hip.call("bmx_concat", eng.ctx, ab.ptr, ca, ub.ptr, cb, nv, cb_buf.ptr)
    y1, y2 = eng.alloc(nv * co * 2), eng.alloc(nv * co * 2)
    _hip.call("bmx_conv3d_fwd", eng.ctx, cb_buf.ptr, *shape, ca + cb, wb.ptr, biasb.ptr, y1.ptr, co, 1)
    _hip.call("bmx_conv3d_fwd_cat", eng.ctx, ab.ptr, *shape, ca, bb.ptr, cb, wb.ptr, biasb.ptr, y2.ptr, co, 1)
    np.testing.assert_array_equal(y2.download(shape + (co,), np.uint16), y1.download(shape + (co,), np.uint16))
    ref = O.conv3d(np.concatenate([a, O.up2(b)], -1), w, bias.astype(np.float64), relu=True)
    _close_bf16(down_bf16(y2, shape + (co,)), ref)


@pytest.mark.parametrize("shape,ci,cat", [((2, 20, 9, 70), 32, False), ((1, 34, 10, 66), 16, False),
                                          ((1, 34, 10, 66), 48, True), ((2, 18, 6, 40), 48, True)])
def test_conv_fwd16_lds_bit_identical(eng, shape, ci, cat):
    """The 16-output-channel forward (staged in LDS, k_conv_fwd_zm, for plain
    inputs) runs the generic kernel's MFMA sequence per output channel, so it
    equals channels 0..15 of a 32-channel convolution (generic kernel) whose
    first 16 filters are its filters, bit for bit, across x/y tiles, z chunks
    and N; the concatenated-input cases check the same of the generic kernel."""
    rng = np.random.default_rng(sum(shape) + ci)
    w32 = O.bf16_round(rng.standard_normal((32, 27, ci)) * 0.1)
    b32 = rng.standard_normal(32).astype(np.float32)
    nv = int(np.prod(shape))
    wb32, bb32 = up_bf16(eng, w32), eng.upload(b32)
    wb16, bb16 = up_bf16(eng, w32[:16]), eng.upload(b32[:16].copy())
    y32, y16 = eng.alloc(nv * 32 * 2), eng.alloc(nv * 16 * 2)
    if cat:
        ca, cb = 16, ci - 16
        half = (shape[0], shape[1] // 2, shape[2] // 2, shape[3] // 2)
        a = O.bf16_round(rng.standard_normal(shape + (ca,)))
        b = O.bf16_round(rng.standard_normal(half + (cb,)))
        ab, bb = up_bf16(eng, a), up_bf16(eng, b)
        _hip.call("bmx_conv3d_fwd_cat", eng.ctx, ab.ptr, *shape, ca, bb.ptr, cb, wb32.ptr, bb32.ptr, y32.ptr, 32, 1)
        _hip.call("bmx_conv3d_fwd_cat", eng.ctx, ab.ptr, *shape, ca, bb.ptr, cb, wb16.ptr, bb16.ptr, y16.ptr, 16, 1)
        ref = O.conv3d(np.concatenate([a, O.up2(b)], -1), w32[:16], b32[:16].astype(np.float64), relu=True)
    else:
        x = O.bf16_round(rng.standard_normal(shape + (ci,)))
        xb = up_bf16(eng, x)
        _hip.call("bmx_conv3d_fwd", eng.ctx, xb.ptr, *shape, ci, wb32.ptr, bb32.ptr, y32.ptr, 32, 0)
        _hip.call("bmx_conv3d_fwd", eng.ctx, xb.ptr, *shape, ci, wb16.ptr, bb16.ptr, y16.ptr, 16, 0)
        ref = O.conv3d(x, w32[:16], b32[:16].astype(np.float64), relu=False)
    got16 = y16.download(shape + (16,), np.uint16)
    np.testing.assert_array_equal(got16, y32.download(shape + (32,), np.uint16)[..., :16])
    _close_bf16(O.from_bf16_bits(got16).astype(np.float64), ref)
