"""Float64 NumPy restatement of the learned-proposal ops (test infrastructure).

Conv3d 3x3x3 with zero padding 1, channels-last [N,D,H,W,C], weights
[Co][27][Ci] with tap = (kz*3 + ky)*3 + kx for offsets (kz-1, ky-1, kx-1);
its input and weight gradients; 2x average pooling; 2x nearest upsampling;
1x1x1 softmax head with soft-target cross-entropy.  Pinned against finite
differences in tests/test_net_oracle.py.
"""
import numpy as np

TAPS = [(kz - 1, ky - 1, kx - 1) for kz in range(3) for ky in range(3) for kx in range(3)]


def to_bf16_bits(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


def bf16_round(x):
    return from_bf16_bits(to_bf16_bits(x)).astype(np.float64)


def _shifted(x, dz, dy, dx):
    """s[n,z,y,x,c] = x[n,z+dz,y+dy,x+dx,c] with zeros outside."""
    p = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    N, D, H, W, C = x.shape
    return p[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W, :]


def conv3d(x, w, b=None, relu=False):
    y = np.zeros(x.shape[:4] + (w.shape[0],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        y += np.einsum("ndhwc,oc->ndhwo", _shifted(x, dz, dy, dx), w[:, t, :])
    if b is not None:
        y += b
    return np.maximum(y, 0.0) if relu else y


def conv3d_dgrad(dy, w):
    dx = np.zeros(dy.shape[:4] + (w.shape[2],))
    for t, (oz, oy, ox) in enumerate(TAPS):
        dx += np.einsum("ndhwo,oc->ndhwc", _shifted(dy, -oz, -oy, -ox), w[:, t, :])
    return dx


def conv3d_wgrad(x, dy):
    Co, Ci = dy.shape[-1], x.shape[-1]
    dw = np.zeros((Co, 27, Ci))
    for t, (dz, dy_, dx) in enumerate(TAPS):
        dw[:, t, :] = np.einsum("ndhwo,ndhwc->oc", dy, _shifted(x, dz, dy_, dx))
    return dw, dy.sum(axis=(0, 1, 2, 3))


def pool2(x):
    N, D, H, W, C = x.shape
    return x.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def pool2_bwd(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 1), 2, 2), 2, 3) / 8.0


def up2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up2_bwd(dy):
    N, D, H, W, C = dy.shape
    return dy.reshape(N, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


def head(x, W, b, target=None, vmask=None, scale=1.0):
    """x [...,16], W [L,16], b [L] -> (prob, loss, dx, dW, db)."""
    lg = x @ W.T + b
    lg -= lg.max(axis=-1, keepdims=True)
    e = np.exp(lg)
    p = e / e.sum(axis=-1, keepdims=True)
    if target is None:
        return p, None, None, None, None
    m = np.ones(x.shape[:-1]) if vmask is None else vmask.astype(np.float64)
    loss = -(m[..., None] * target * np.log(np.maximum(p, 1e-30))).sum()
    g = scale * m[..., None] * (p - target)
    dx = g @ W
    dW = g.reshape(-1, g.shape[-1]).T @ x.reshape(-1, x.shape[-1])
    db = g.reshape(-1, g.shape[-1]).sum(0)
    return p, loss, dx, dW, db


# ---------------------------------------------------------------------------
# The whole encoder-decoder (same layer list as the GPU network), float64,
# with optional bf16 rounding of activations and activation gradients at the
# points where the GPU stores them in bf16.
# ---------------------------------------------------------------------------
UNET_LAYERS = [("enc0", 32, 16), ("enc1", 16, 32), ("enc2", 32, 64), ("bott", 64, 128),
               ("dec2", 192, 64), ("dec1", 96, 32), ("dec0", 48, 16)]


def unet_loss_and_grads(params, x, target, vmask, round_bf16=True):
    """params: {name.w [Co,27,Ci], name.b, head.W [L,16], head.b}; x [N,D,H,W,Ci of enc0];
    target [N,D,H,W,L]; vmask [N,D,H,W].  Returns (summed loss, grads dict)."""
    r = bf16_round if round_bf16 else (lambda a: a)
    P = {k: (bf16_round(v) if (round_bf16 and k.endswith(".w")) else v.astype(np.float64))
         for k, v in params.items()}

    def conv(name, inp):
        return r(conv3d(inp, P[name + ".w"], P[name + ".b"], relu=True))

    a0 = conv("enc0", x)
    p0 = r(pool2(a0)); a1 = conv("enc1", p0)
    p1 = r(pool2(a1)); a2 = conv("enc2", p1)
    p2 = r(pool2(a2)); a3 = conv("bott", p2)
    c2 = np.concatenate([a2, up2(a3)], -1); d2 = conv("dec2", c2)
    c1 = np.concatenate([a1, up2(d2)], -1); d1 = conv("dec1", c1)
    c0 = np.concatenate([a0, up2(d1)], -1); d0 = conv("dec0", c0)
    nvox = float(np.prod(x.shape[:4]))
    _, loss, g_d0, dW, db = head(d0, P["head.W"], P["head.b"], target, vmask, 1.0 / nvox)
    G = {"head.W": dW, "head.b": db}

    def back(name, inp, g_out, out, mask_in):
        g = r(g_out) * (out > 0)
        G[name + ".w"], G[name + ".b"] = conv3d_wgrad(inp, g)
        gi = conv3d_dgrad(g, P[name + ".w"])
        return r(gi * (inp > 0)) if mask_in else r(gi)

    g_c0 = back("dec0", c0, g_d0, d0, True)
    g_a0, g_d1 = g_c0[..., :16], r(up2_bwd(g_c0[..., 16:]))
    g_c1 = back("dec1", c1, g_d1, d1, True)
    g_a1, g_d2 = g_c1[..., :32], r(up2_bwd(g_c1[..., 32:]))
    g_c2 = back("dec2", c2, g_d2, d2, True)
    g_a2, g_a3 = g_c2[..., :64], r(up2_bwd(g_c2[..., 64:]))
    g_p2 = back("bott", p2, g_a3, a3, False)
    g_a2 = r(g_a2 + r(pool2_bwd(g_p2)))
    g_p1 = back("enc2", p1, g_a2, a2, False)
    g_a1 = r(g_a1 + r(pool2_bwd(g_p1)))
    g_p0 = back("enc1", p0, g_a1, a1, False)
    g_a0 = r(g_a0 + r(pool2_bwd(g_p0)))
    back("enc0", x, g_a0, a0, False)
    return loss, G
