"""Pins of the float64 net oracle (oracle/net.py) by finite differences and
closed forms: the GPU kernels of the learned proposal are checked against it."""
import numpy as np
import pytest

from oracle import net as O


def _fd(f, x, eps=1e-6, idx=None):
    """Central differences of scalar f at a few coordinates of x."""
    rng = np.random.default_rng(0)
    flat = x.reshape(-1)
    picks = idx if idx is not None else rng.choice(flat.size, 12, replace=False)
    out = []
    for i in picks:
        old = flat[i]
        flat[i] = old + eps; fp = f()
        flat[i] = old - eps; fm = f()
        flat[i] = old
        out.append((fp - fm) / (2 * eps))
    return np.array(out), picks


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 3, 4, 5, 2))
    w = rng.standard_normal((3, 27, 2))
    b = rng.standard_normal(3)
    y = O.conv3d(x, w, b)
    n, z, yy, xx = 0, 1, 0, 4
    ref = b.copy()
    for t, (dz, dy, dx) in enumerate(O.TAPS):
        zz, y2, x2 = z + dz, yy + dy, xx + dx
        if 0 <= zz < 3 and 0 <= y2 < 4 and 0 <= x2 < 5:
            ref += w[:, t, :] @ x[n, zz, y2, x2]
    np.testing.assert_allclose(y[n, z, yy, xx], ref, rtol=1e-12)


def test_conv_gradients_finite_differences():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 3, 4, 3, 2))
    w = rng.standard_normal((3, 27, 2))
    g = rng.standard_normal((2, 3, 4, 3, 3))
    f = lambda: float((O.conv3d(x, w) * g).sum())
    num, picks = _fd(f, x)
    np.testing.assert_allclose(O.conv3d_dgrad(g, w).reshape(-1)[picks], num, rtol=1e-6, atol=1e-8)
    num, picks = _fd(f, w)
    dw, db = O.conv3d_wgrad(x, g)
    np.testing.assert_allclose(dw.reshape(-1)[picks], num, rtol=1e-6, atol=1e-8)
    bb = np.zeros(3)
    fb = lambda: float(((O.conv3d(x, w) + bb) * g).sum())
    num, picks = _fd(fb, bb, idx=[0, 1, 2])
    np.testing.assert_allclose(db, num, rtol=1e-6)


def test_pool_up_adjoints():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1, 4, 6, 2, 3))
    g = rng.standard_normal((1, 2, 3, 1, 3))
    # <pool(x), g> = <x, pool_bwd(g)>, <up(g), x> = <g, up_bwd(x)>
    assert np.isclose((O.pool2(x) * g).sum(), (x * O.pool2_bwd(g)).sum())
    assert np.isclose((O.up2(g) * x).sum(), (g * O.up2_bwd(x)).sum())


def test_head_gradients_finite_differences():
    rng = np.random.default_rng(4)
    x = rng.standard_normal((5, 16))
    W = rng.standard_normal((4, 16)) * 0.3
    b = rng.standard_normal(4)
    t = rng.random((5, 4)); t /= t.sum(-1, keepdims=True)
    p, loss, dx, dW, db = O.head(x, W, b, t)
    np.testing.assert_allclose(p.sum(-1), 1.0)
    f = lambda: O.head(x, W, b, t)[1]
    num, picks = _fd(f, x); np.testing.assert_allclose(dx.reshape(-1)[picks], num, rtol=1e-5, atol=1e-8)
    num, picks = _fd(f, W); np.testing.assert_allclose(dW.reshape(-1)[picks], num, rtol=1e-5, atol=1e-8)
    num, picks = _fd(f, b, idx=[0, 1, 2, 3]); np.testing.assert_allclose(db, num, rtol=1e-5, atol=1e-8)


def test_bf16_rounding_ties_to_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0], np.float32)
    r = O.bf16_round(x)
    np.testing.assert_array_equal(r, [1.0, 1.0, 1.0 + 2 ** -6, -2.5, 0.0])


def test_whole_network_gradient_finite_differences():
    """The oracle encoder-decoder's analytic gradients (float64, no rounding; they
    are gradients of the mean loss, the loss returned is the sum) against central
    differences of its loss, for a few weights of every layer."""
    rng = np.random.default_rng(5)
    params = {}
    for n, ci, co in O.UNET_LAYERS:
        params[n + ".w"] = rng.standard_normal((co, 27, ci)) * np.sqrt(2 / (27 * ci))
        params[n + ".b"] = rng.standard_normal(co) * 0.1
    params["head.W"] = rng.standard_normal((14, 16)) * 0.25
    params["head.b"] = np.zeros(14)
    x = rng.standard_normal((1, 8, 8, 8, O.UNET_LAYERS[0][1]))
    t = rng.random((1, 8, 8, 8, 14)); t /= t.sum(-1, keepdims=True)
    vm = np.ones((1, 8, 8, 8))
    loss, G = O.unet_loss_and_grads(params, x, t, vm, round_bf16=False)
    for k in ["enc0.w", "enc1.w", "enc2.w", "bott.w", "dec2.w", "dec1.w", "dec0.w", "dec2.b", "head.W"]:
        flat = params[k].reshape(-1)
        g = G[k].reshape(-1)
        for i in rng.choice(flat.size, 3, replace=False):
            old = flat[i]
            flat[i] = old + 1e-5; lp, _ = O.unet_loss_and_grads(params, x, t, vm, False)
            flat[i] = old - 1e-5; lm, _ = O.unet_loss_and_grads(params, x, t, vm, False)
            flat[i] = old
            fd = (lp - lm) / 2e-5 / x[..., 0].size
            assert abs(fd - g[i]) <= 1e-4 * max(1e-2, abs(fd)), (k, i, fd, g[i])
