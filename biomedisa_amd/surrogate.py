"""Learned proposal for smart interpolation: a 3-D conv encoder-decoder.

Architecture (4 levels, 16/32/64/128 channels, one 3x3x3 conv + ReLU per
level and path, 2x average pooling down, 2x nearest upsampling up, skip
concatenations, 1x1x1 softmax head):

    x(32) -enc0-> a0(16) -pool-> -enc1-> a1(32) -pool-> -enc2-> a2(64) -pool-> -bott-> a3(128)
    [a2, up a3](192) -dec2-> d2(64);  [a1, up d2](96) -dec1-> d1(32);
    [a0, up d1](48) -dec0-> d0(16) -head-> softmax over L = 30 label indices

Input channels (bmx_net_input, 32): image/127.5 - 1, seed indicator, one-hot
of the label index (< 30) on seed voxels.  The network is trained with cross-entropy
against the plain walker's per-voxel label (argmax of its hits) on seeded
synthetic volumes, so its top probability estimates how likely the plain
walker is to agree with its proposal.  BF16 weights and activations, FP32 accumulation, FP32
master weights and gradients; every forward/backward op is a libbmx_hip
kernel (bmx_net.hip).  No counterpart in the reference.
"""
import ctypes as C
import os
import time

import numpy as np

from . import _hip

L_MAX = 30        # label indices the input encoding and the head support
CIN = 32          # input channels: image, seed indicator, one-hot of 30 label indices
LAYERS = [("enc0", 32, 16), ("enc1", 16, 32), ("enc2", 32, 64), ("bott", 64, 128),
          ("dec2", 192, 64), ("dec1", 96, 32), ("dec0", 48, 16)]
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "surrogate_weights.npz")


def _f(x):
    return C.c_float(float(x))


def pad8(n):
    return (n + 7) // 8 * 8


class Param:
    """FP32 master tensor with gradient, Adam moments and (for conv weights) a bf16 copy."""

    def __init__(self, eng, init, bf16=False):
        self.shape = init.shape
        self.n = int(init.size)
        self.w = eng.upload(init.astype(np.float32))
        self.g = eng.alloc(self.n * 4)
        self.m = eng.alloc(self.n * 4)
        self.v = eng.alloc(self.n * 4)
        for b in (self.g, self.m, self.v):
            b.zero()
        self.bf = eng.alloc(self.n * 2) if bf16 else None
        self.eng = eng
        self.refresh()

    def refresh(self):
        if self.bf is not None:
            _hip.call("bmx_f32_to_bf16", self.eng.ctx, self.w.ptr, self.bf.ptr, self.n)

    def set(self, arr):
        _hip.load().bmx_upload(self.eng.ctx, self.w.ptr, np.ascontiguousarray(arr, np.float32).ctypes.data,
                               self.n * 4)
        self.refresh()

    def get(self):
        return self.w.download(self.shape, np.float32)


class Net:
    """The encoder-decoder on one device (activations cached per input shape)."""

    def __init__(self, eng, seed=0):
        self.eng = eng
        rng = np.random.default_rng(seed)
        self.conv = {}
        for name, ci, co in LAYERS:
            w = rng.standard_normal((co, 27, ci)) * np.sqrt(2.0 / (27 * ci))
            self.conv[name] = (Param(eng, w, bf16=True), Param(eng, np.zeros(co)), ci, co)
        self.hW = Param(eng, rng.standard_normal((L_MAX, 16)) * 0.25)
        self.hb = Param(eng, np.zeros(L_MAX))
        self._bufs = {}
        self.step_count = 0
        self.T = 1.0             # softmax temperature used by predict (fitted by calibrate)
        self._head_inf = None    # head weights / T on the device, rebuilt after changes
        self.scratch = eng.alloc(max(ci * co * 27 for _, ci, co in LAYERS) * 2)

    # -- parameters -----------------------------------------------------------
    def params(self):
        out = []
        for name, _, _ in LAYERS:
            w, b, _, _ = self.conv[name]
            out += [w, b]
        return out + [self.hW, self.hb]

    def state_dict(self):
        d = {}
        for name, _, _ in LAYERS:
            w, b, _, _ = self.conv[name]
            d[name + ".w"], d[name + ".b"] = w.get(), b.get()
        d["head.W"], d["head.b"] = self.hW.get(), self.hb.get()
        d["head.T"] = np.array(self.T, np.float32)
        return d

    def load_state_dict(self, d):
        for name, _, _ in LAYERS:
            w, b, _, _ = self.conv[name]
            w.set(d[name + ".w"])
            b.set(d[name + ".b"])
        self.hW.set(d["head.W"])
        self.hb.set(d["head.b"])
        self.T = float(d["head.T"]) if "head.T" in d else 1.0
        self._head_inf = None

    def save(self, path=WEIGHTS_PATH):
        np.savez_compressed(path, **self.state_dict())

    def load(self, path=WEIGHTS_PATH):
        with np.load(path, allow_pickle=False) as f:
            self.load_state_dict({k: f[k] for k in f.files})

    # -- buffers ----------------------------------------------------------------
    def _get(self, key, nbytes):
        b = self._bufs.get(key)
        if b is None or b.nbytes < nbytes:
            b = self.eng.alloc(nbytes)
            self._bufs[key] = b
        return b

    def _acts(self, N, D, H, W, grads):
        s0 = N * D * H * W
        s1, s2, s3 = s0 // 8, s0 // 64, s0 // 512
        sizes = {"a0": s0 * 16, "p0": s1 * 16, "a1": s1 * 32, "p1": s2 * 32, "a2": s2 * 64, "p2": s3 * 64,
                 "a3": s3 * 128, "d2": s2 * 64, "d1": s1 * 32, "d0": s0 * 16}
        if grads:
            sizes.update({"u2": s2 * 128, "c2": s2 * 192, "u1": s1 * 64, "c1": s1 * 96, "u0": s0 * 32,
                          "c0": s0 * 48})
            sizes.update({"g_d0": s0 * 16, "g_c0": s0 * 48, "g_a0": s0 * 16, "g_u0": s0 * 32,
                          "g_d1": s1 * 32, "g_c1": s1 * 96, "g_a1": s1 * 32, "g_u1": s1 * 64,
                          "g_d2": s2 * 64, "g_c2": s2 * 192, "g_a2": s2 * 64, "g_u2": s2 * 128,
                          "g_a3": s3 * 128, "g_p2": s3 * 64, "g_p1": s2 * 32, "g_p0": s1 * 16,
                          "g_tmp2": s2 * 64, "g_tmp1": s1 * 32, "g_tmp0": s0 * 16})
        return {k: self._get(k, n * 2) for k, n in sizes.items()}

    # -- forward / backward ---------------------------------------------------------
    def _conv(self, name, x, shape, y):
        w, b, ci, co = self.conv[name]
        _hip.call("bmx_conv3d_fwd", self.eng.ctx, x.ptr, *shape, ci, w.bf.ptr, b.w.ptr, y.ptr, co, 1)

    def _conv_cat(self, name, a, ca, b, cb, shape, y):
        w, bias, ci, co = self.conv[name]
        assert ci == ca + cb
        _hip.call("bmx_conv3d_fwd_cat", self.eng.ctx, a.ptr, *shape, ca, b.ptr, cb, w.bf.ptr, bias.w.ptr, y.ptr, co, 1)

    def forward(self, x, N, D, H, W, grads=False):
        """x: bf16 [N,D,H,W,CIN] (D,H,W multiples of 8) -> d0 buffer (bf16 [.., 16])."""
        assert D % 8 == 0 and H % 8 == 0 and W % 8 == 0
        a = self._acts(N, D, H, W, grads)
        ctx = self.eng.ctx
        S0, S1, S2, S3 = (N, D, H, W), (N, D // 2, H // 2, W // 2), (N, D // 4, H // 4, W // 4), \
            (N, D // 8, H // 8, W // 8)
        n1, n2 = int(np.prod(S1)), int(np.prod(S2))
        n0 = int(np.prod(S0))
        self._conv("enc0", x, S0, a["a0"])
        _hip.call("bmx_pool2", ctx, a["a0"].ptr, *S0, 16, a["p0"].ptr)
        self._conv("enc1", a["p0"], S1, a["a1"])
        _hip.call("bmx_pool2", ctx, a["a1"].ptr, *S1, 32, a["p1"].ptr)
        self._conv("enc2", a["p1"], S2, a["a2"])
        _hip.call("bmx_pool2", ctx, a["a2"].ptr, *S2, 64, a["p2"].ptr)
        self._conv("bott", a["p2"], S3, a["a3"])
        if grads:
            # training keeps the concatenated decoder inputs for the weight gradients
            _hip.call("bmx_up2", ctx, a["a3"].ptr, *S3, 128, a["u2"].ptr)
            _hip.call("bmx_concat", ctx, a["a2"].ptr, 64, a["u2"].ptr, 128, n2, a["c2"].ptr)
            self._conv("dec2", a["c2"], S2, a["d2"])
            _hip.call("bmx_up2", ctx, a["d2"].ptr, *S2, 64, a["u1"].ptr)
            _hip.call("bmx_concat", ctx, a["a1"].ptr, 32, a["u1"].ptr, 64, n1, a["c1"].ptr)
            self._conv("dec1", a["c1"], S1, a["d1"])
            _hip.call("bmx_up2", ctx, a["d1"].ptr, *S1, 32, a["u0"].ptr)
            _hip.call("bmx_concat", ctx, a["a0"].ptr, 16, a["u0"].ptr, 32, n0, a["c0"].ptr)
            self._conv("dec0", a["c0"], S0, a["d0"])
        else:
            # inference reads concat(skip, up2(coarser)) in place
            self._conv_cat("dec2", a["a2"], 64, a["a3"], 128, S2, a["d2"])
            self._conv_cat("dec1", a["a1"], 32, a["d2"], 64, S1, a["d1"])
            self._conv_cat("dec0", a["a0"], 16, a["d1"], 32, S0, a["d0"])
        self._last = (x, (N, D, H, W), a)
        return a["d0"]

    def _wgrad(self, name, x, shape, dy):
        w, b, ci, co = self.conv[name]
        _hip.call("bmx_conv3d_wgrad", self.eng.ctx, x.ptr, *shape, ci, dy.ptr, co, w.g.ptr, b.g.ptr)

    def _dgrad(self, name, dy, shape, xact, dx):
        w, b, ci, co = self.conv[name]
        _hip.call("bmx_conv3d_dgrad", self.eng.ctx, dy.ptr, *shape, co, w.bf.ptr,
                  None if xact is None else xact.ptr, dx.ptr, ci, self.scratch.ptr)

    def backward(self):
        """Back-propagate a["g_d0"] (gradient w.r.t. d0) into every parameter gradient."""
        x, (N, D, H, W), a = self._last
        ctx = self.eng.ctx
        S0, S1, S2, S3 = (N, D, H, W), (N, D // 2, H // 2, W // 2), (N, D // 4, H // 4, W // 4), \
            (N, D // 8, H // 8, W // 8)
        n0, n1, n2 = int(np.prod(S0)), int(np.prod(S1)), int(np.prod(S2))
        _hip.call("bmx_relu_mask", ctx, a["g_d0"].ptr, a["d0"].ptr, n0 * 16)
        self._wgrad("dec0", a["c0"], S0, a["g_d0"])
        self._dgrad("dec0", a["g_d0"], S0, a["c0"], a["g_c0"])
        _hip.call("bmx_split", ctx, a["g_c0"].ptr, n0, a["g_a0"].ptr, 16, a["g_u0"].ptr, 32, 0)
        _hip.call("bmx_up2_bwd", ctx, a["g_u0"].ptr, *S1, 32, a["g_d1"].ptr)
        self._wgrad("dec1", a["c1"], S1, a["g_d1"])
        self._dgrad("dec1", a["g_d1"], S1, a["c1"], a["g_c1"])
        _hip.call("bmx_split", ctx, a["g_c1"].ptr, n1, a["g_a1"].ptr, 32, a["g_u1"].ptr, 64, 0)
        _hip.call("bmx_up2_bwd", ctx, a["g_u1"].ptr, *S2, 64, a["g_d2"].ptr)
        self._wgrad("dec2", a["c2"], S2, a["g_d2"])
        self._dgrad("dec2", a["g_d2"], S2, a["c2"], a["g_c2"])
        _hip.call("bmx_split", ctx, a["g_c2"].ptr, n2, a["g_a2"].ptr, 64, a["g_u2"].ptr, 128, 0)
        _hip.call("bmx_up2_bwd", ctx, a["g_u2"].ptr, *S3, 128, a["g_a3"].ptr)
        self._wgrad("bott", a["p2"], S3, a["g_a3"])
        self._dgrad("bott", a["g_a3"], S3, None, a["g_p2"])
        _hip.call("bmx_pool2_bwd", ctx, a["g_p2"].ptr, *S2, 64, a["g_tmp2"].ptr)
        _hip.call("bmx_add_bf16", ctx, a["g_a2"].ptr, a["g_tmp2"].ptr, n2 * 64)
        _hip.call("bmx_relu_mask", ctx, a["g_a2"].ptr, a["a2"].ptr, n2 * 64)
        self._wgrad("enc2", a["p1"], S2, a["g_a2"])
        self._dgrad("enc2", a["g_a2"], S2, None, a["g_p1"])
        _hip.call("bmx_pool2_bwd", ctx, a["g_p1"].ptr, *S1, 32, a["g_tmp1"].ptr)
        _hip.call("bmx_add_bf16", ctx, a["g_a1"].ptr, a["g_tmp1"].ptr, n1 * 32)
        _hip.call("bmx_relu_mask", ctx, a["g_a1"].ptr, a["a1"].ptr, n1 * 32)
        self._wgrad("enc1", a["p0"], S1, a["g_a1"])
        self._dgrad("enc1", a["g_a1"], S1, None, a["g_p0"])
        _hip.call("bmx_pool2_bwd", ctx, a["g_p0"].ptr, *S0, 16, a["g_tmp0"].ptr)
        _hip.call("bmx_add_bf16", ctx, a["g_a0"].ptr, a["g_tmp0"].ptr, n0 * 16)
        _hip.call("bmx_relu_mask", ctx, a["g_a0"].ptr, a["a0"].ptr, n0 * 16)
        self._wgrad("enc0", x, S0, a["g_a0"])

    def zero_grad(self):
        for p in self.params():
            p.g.zero()

    def loss_and_grad(self, x, target, vmask, N, D, H, W, loss_buf, target_idx=None):
        """Forward, softmax cross-entropy against soft targets (float [.., L_MAX]) or
        label indices (uint8 target_idx), backward.  The summed loss accumulates
        into loss_buf (float)."""
        d0 = self.forward(x, N, D, H, W, grads=True)
        a = self._last[2]
        nvox = N * D * H * W
        _hip.call("bmx_head", self.eng.ctx, d0.ptr, nvox, L_MAX, self.hW.w.ptr, self.hb.w.ptr, None,
                  None if target is None else target.ptr, None if target_idx is None else target_idx.ptr,
                  vmask.ptr, _f(1.0 / nvox), a["g_d0"].ptr, self.hW.g.ptr, self.hb.g.ptr, loss_buf.ptr)
        self.backward()

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8, wd=0.0):
        self.step_count += 1
        self._head_inf = None
        for p in self.params():
            _hip.call("bmx_adam", self.eng.ctx, p.w.ptr, p.g.ptr, p.m.ptr, p.v.ptr,
                      None if p.bf is None else p.bf.ptr, p.n, _f(lr), _f(beta1), _f(beta2), _f(eps),
                      self.step_count, _f(wd))

    def _head_for_inference(self):
        """Head weights divided by the temperature: logits / T = (W/T) x + b/T."""
        if self._head_inf is None:
            self._head_inf = (self.eng.upload((self.hW.get() / self.T).astype(np.float32)),
                              self.eng.upload((self.hb.get() / self.T).astype(np.float32)))
        return self._head_inf

    def predict(self, x, dims, grid, nlabels, thr, pred, unc):
        """Labels (index) and uncertainty flags of a (Z,Y,X) volume whose input x lives on
        the padded grid (Zp,Yp,Xp): flag = top softmax(logits / T) probability < thr."""
        Zp, Yp, Xp = grid
        d0 = self.forward(x, 1, Zp, Yp, Xp)
        W, b = self._head_for_inference()
        _hip.call("bmx_head_predict", self.eng.ctx, d0.ptr, *dims, Zp, Yp, Xp, int(nlabels),
                  W.ptr, b.ptr, _f(thr), pred.ptr, unc.ptr)

    # Slab-wise inference: a slab of the volume plus HALO planes on each side
    # (more than the network's 29-voxel receptive field) gives exact predictions
    # for its core planes; volumes whose padded grid exceeds MAX_GRID voxels
    # are predicted slab by slab along z.
    HALO = 32
    MAX_GRID = 300_000_000

    def predict_volume(self, vol, flags, nlabels, thr, pred, unc, max_grid=None):
        """Labels and uncertainty flags of a DeviceVolume (flags: host uint8 plane
        flags [Z+Y+X]) into the device buffers pred / unc (uint8, V each)."""
        Z, Y, X = vol.dims
        Yp, Xp = pad8(Y), pad8(X)
        max_grid = self.MAX_GRID if max_grid is None else max_grid
        eng = self.eng
        if pad8(Z) * Yp * Xp <= max_grid:
            core = Z
        else:
            core = (max_grid // (Yp * Xp) - 2 * self.HALO) // 8 * 8
            if core < 8:
                raise ValueError("volume too wide for slab-wise inference (%d x %d)" % (Y, X))
        item = 1 if vol.img_dtype == _hip.IMG_U8 else 4
        flags = np.ascontiguousarray(flags, np.uint8)
        W, b = self._head_for_inference()
        for z0 in range(0, Z, core):
            zs, ze = max(0, z0 - self.HALO), min(Z, z0 + core + self.HALO)
            nz = ze - zs
            grid = (pad8(nz), Yp, Xp)
            x = self._get("input", int(np.prod(grid)) * CIN * 2)
            fs = eng.upload(np.concatenate([flags[zs:ze], flags[Z:Z + Y], flags[Z + Y:]]))
            _hip.call("bmx_net_input", eng.ctx, vol.img.ptr + zs * Y * X * item, vol.img_dtype,
                      vol.lab.ptr + zs * Y * X, fs.ptr, nz, Y, X, *grid, x.ptr)
            d0 = self.forward(x, 1, *grid)
            nc = min(core, Z - z0)
            off = (z0 - zs) * Yp * Xp * 16 * 2   # first core plane of the slab's decoder output
            _hip.call("bmx_head_predict", eng.ctx, d0.ptr + off, nc, Y, X, grid[0], Yp, Xp, int(nlabels),
                      W.ptr, b.ptr, _f(thr), pred.ptr + z0 * Y * X, unc.ptr + z0 * Y * X)

    def log_probs(self, x, grid, nlabels):
        """log softmax over the first nlabels logits (temperature 1) on the padded grid,
        float32 [Zp,Yp,Xp,nlabels] on the host."""
        Zp, Yp, Xp = grid
        d0 = self.forward(x, 1, Zp, Yp, Xp)
        n = Zp * Yp * Xp
        prob = self._get("prob", n * nlabels * 4)
        _hip.call("bmx_head", self.eng.ctx, d0.ptr, n, int(nlabels), self.hW.w.ptr, self.hb.w.ptr, prob.ptr,
                  None, None, None, _f(1.0), None, None, None, None)
        p = prob.download((Zp, Yp, Xp, nlabels), np.float32)
        return np.log(np.maximum(p, 1e-38))

    def flops_forward(self, N, D, H, W):
        v0 = N * D * H * W
        lv = {"enc0": v0, "enc1": v0 / 8, "enc2": v0 / 64, "bott": v0 / 512, "dec2": v0 / 64,
              "dec1": v0 / 8, "dec0": v0}
        f = sum(2.0 * 27 * ci * co * lv[n] for n, ci, co in LAYERS)
        return f + 2.0 * 16 * L_MAX * v0


_nets = {}


def net(device=0, weights=WEIGHTS_PATH):
    """Process-wide trained network on `device` (weights loaded from `weights`)."""
    key = (device, weights)
    n = _nets.get(key)
    if n is None:
        n = Net(_hip.engine(device))
        if not os.path.exists(weights):
            raise RuntimeError("surrogate weights not found at %s: train them with "
                               "`python -m biomedisa_amd.surrogate train`" % weights)
        n.load(weights)
        _nets[key] = n
    return n


# ---------------------------------------------------------------------------
# training data: seeded synthetic volumes walked by the plain walker
# ---------------------------------------------------------------------------
def training_volume(seed, size=64):
    """One seeded synthetic volume with sparse labels (structure count, sizes, noise,
    bias and label spacing drawn from the seed)."""
    from .synth import structures_volume
    rng = np.random.default_rng(seed)
    n_struct = int(rng.integers(2, 9)) if rng.random() < 0.7 else int(rng.integers(9, 17))
    every = int(rng.integers(6, 13))
    allaxis = bool(rng.random() < 0.7)
    return structures_volume(shape=(size, size, size), seed=int(seed) + 1000003, n_struct=n_struct,
                             radius=(3, max(6, size // 3)), mean_range=(30, 230),
                             sigma=float(rng.uniform(8, 20)), bias=float(rng.uniform(0, 15)),
                             every=every, allaxis=allaxis)


class Sample:
    """Device-resident training sample on a padded grid: input x, the walker's label
    index per voxel (tidx, hard targets) or its hit fractions (target, soft
    targets), and the in-volume mask."""

    def __init__(self, x, target, tidx, vmask, grid):
        self.x, self.target, self.tidx, self.vmask, self.grid = x, target, tidx, vmask, grid


def make_sample(eng, vol, nbrw=10, sorw=4000, key=1, hard=True):
    """Walk a synthetic volume with the plain walker and build its training sample.

    hard=True trains towards the one-hot of the walker's argmax, so the network's
    top probability estimates how likely the plain walker is to give that label;
    hard=False towards the walker's per-voxel hit fractions."""
    from . import helpers as H
    from .interpolation import DeviceVolume, run_pipeline
    bm = H.Biomedisa()
    bm.data, bm.labelData, bm.only, bm.ignore, bm.success = vol.img, vol.labels, "all", "none", True
    bm = H.pre_processing(bm)
    if not bm.success or bm.nol > L_MAX:
        return None
    img = np.pad(H.normalize_image(bm.data), 1)
    lab = np.pad(bm.labelData, 1)
    labidx = H.label_index_volume(lab, bm.allLabels)
    dv = DeviceVolume(eng, np.ascontiguousarray(img), labidx)
    out = run_pipeline(dv, bm.nol, bm.allLabels.astype(np.uint8), nbrw=nbrw, sorw=sorw,
                       allaxis=vol.allaxis, keep_hits=True, key=key)
    Z, Y, X = dv.dims
    grid = (pad8(Z), pad8(Y), pad8(X))
    ng = int(np.prod(grid))
    x = eng.alloc(ng * CIN * 2)
    flags = eng.upload(out["flags"])
    _hip.call("bmx_net_input", eng.ctx, dv.img.ptr, dv.img_dtype, dv.lab.ptr, flags.ptr, Z, Y, X, *grid, x.ptr)
    vmask = eng.alloc(ng)
    target = tidx = None
    if hard:
        tidx = eng.alloc(ng)
        _hip.call("bmx_hits_index", eng.ctx, out["hits"].ptr, bm.nol, Z, Y, X, *grid, tidx.ptr, vmask.ptr)
    else:
        target = eng.alloc(ng * L_MAX * 4)
        _hip.call("bmx_hits_target", eng.ctx, out["hits"].ptr, bm.nol, Z, Y, X, *grid, L_MAX, 0, target.ptr,
                  vmask.ptr)
    eng.sync()
    return Sample(x, target, tidx, vmask, grid)


def train(n_volumes=2000, size=64, iters=1500, batch=4, lr=2e-3, seed=0, out=WEIGHTS_PATH, log=print,
          device=0, hard=True):
    """Train the network on the GPU with the project's own kernels; save FP32 weights."""
    eng = _hip.engine(device)
    t0 = time.perf_counter()
    samples = []
    s = seed
    while len(samples) < n_volumes:
        smp = make_sample(eng, training_volume(s, size), hard=hard)
        s += 1
        if smp is not None:
            samples.append(smp)
        if len(samples) % 200 == 0 and smp is not None:
            log("training data: %d/%d volumes (%.0fs)" % (len(samples), n_volumes, time.perf_counter() - t0))
    grid = samples[0].grid
    ng = int(np.prod(grid))
    model = Net(eng, seed=seed)
    xb = eng.alloc(batch * ng * CIN * 2)
    tb = None if hard else eng.alloc(batch * ng * L_MAX * 4)
    ib = eng.alloc(batch * ng) if hard else None
    mb = eng.alloc(batch * ng)
    lossb = eng.alloc(4)
    rng = np.random.default_rng(seed)
    t1 = time.perf_counter()
    hist = []
    for it in range(iters):
        idx = rng.choice(len(samples), batch, replace=False)
        for k, i in enumerate(idx):
            smp = samples[i]
            _copy(eng, xb, k * ng * CIN * 2, smp.x, ng * CIN * 2)
            if hard:
                _copy(eng, ib, k * ng, smp.tidx, ng)
            else:
                _copy(eng, tb, k * ng * L_MAX * 4, smp.target, ng * L_MAX * 4)
            _copy(eng, mb, k * ng, smp.vmask, ng)
        model.zero_grad()
        lossb.zero()
        model.loss_and_grad(xb, tb, mb, batch, *grid, lossb, target_idx=ib)
        lr_t = lr * 0.5 * (1 + np.cos(np.pi * it / iters))
        model.adam(lr_t)
        if it % 50 == 0 or it == iters - 1:
            ls = float(lossb.download((1,), np.float32)[0]) / (batch * ng)
            hist.append((it, ls))
            log("iter %d loss %.4f (%.1fs)" % (it, ls, time.perf_counter() - t1))
    eng.sync()
    model.save(out)
    return model, hist, time.perf_counter() - t1


def validation_volume(seed, size=256):
    """Held-out volume of the 512^3 benchmark's kind scaled to `size` (8 structures,
    ellipsoids and tubes, means 25 apart over 30..230, noise 12, bias 15, labels on
    every 10th plane of all axes).  Calibration seeds are disjoint from the seed of
    the benchmark volume."""
    from .synth import structures_volume
    return structures_volume(shape=(size, size, size), seed=int(seed), n_struct=8,
                             radius=(5, max(8, 60 * size // 512)), every=10, allaxis=True)


def calibrate(net, n_volumes=2, size=512, seed0=7001, thr=0.9, brad=3, target_recall=0.95, min_count=500,
              log=print):
    """Fit the softmax temperature T used by predict on held-out volumes.

    For each volume the plain walker gives the reference labels; voxels where the
    network's argmax disagrees with them and that lie outside the predicted
    boundaries dilated by `brad` must be caught by the confidence rule
    (top softmax(logits/T) probability < thr).  A disagreement voxel counts for
    both labels involved.  T is the smallest value on a 0.05 grid for which, for
    every label with at least `min_count` disagreements, the rule plus the
    boundary band covers `target_recall` of them.  Sets net.T and returns
    (T, worst per-label recall, stats)."""
    from . import helpers as H
    from .interpolation import DeviceVolume, run_pipeline
    eng = net.eng
    per_vol = []       # (log-probs, labels involved) of disagreements outside the band
    n_dis = n_band = 0
    tot_l = {}         # label -> disagreements
    band_l = {}        # label -> disagreements inside the band
    for i in range(n_volumes):
        vol = validation_volume(seed0 + i, size)
        bm = H.Biomedisa()
        bm.data, bm.labelData, bm.only, bm.ignore, bm.success = vol.img, vol.labels, "all", "none", True
        bm = H.pre_processing(bm)
        img = np.pad(H.normalize_image(bm.data), 1)
        lab = np.pad(bm.labelData, 1)
        z0, z1, y0, y1, x0, x1 = H.predict_blocksize(lab)
        img = np.ascontiguousarray(img[z0:z1, y0:y1, x0:x1])
        labidx = H.label_index_volume(np.ascontiguousarray(lab[z0:z1, y0:y1, x0:x1]), bm.allLabels)
        nol = bm.nol
        dv = DeviceVolume(eng, img, labidx)
        out = run_pipeline(dv, nol, bm.allLabels.astype(np.uint8), nbrw=10, sorw=4000, allaxis=True,
                           keep_hits=True)
        Z, Y, X = dv.dims
        V = dv.V
        ref = eng.alloc(V)
        eng.argmax(out["hits"], nol, V, np.arange(nol, dtype=np.uint8), ref)
        A = ref.download(dv.dims, np.uint8)
        grid = (pad8(Z), pad8(Y), pad8(X))
        x = net._get("input", int(np.prod(grid)) * CIN * 2)
        flags = eng.upload(out["flags"])
        _hip.call("bmx_net_input", eng.ctx, dv.img.ptr, dv.img_dtype, dv.lab.ptr, flags.ptr, Z, Y, X, *grid,
                  x.ptr)
        lp = net.log_probs(x, grid, nol)[:Z, :Y, :X]
        pred = np.ascontiguousarray(lp.argmax(-1).astype(np.uint8))
        pb, zb = eng.upload(pred), eng.alloc(V)
        zb.zero()
        band, need, tmp = eng.alloc(V), eng.alloc(V), eng.alloc(V)
        _hip.call("bmx_guided_masks", eng.ctx, pb.ptr, zb.ptr, Z, Y, X, int(brad), 0, band.ptr, need.ptr, tmp.ptr)
        bnd = band.download(dv.dims, np.uint8).astype(bool)
        D = pred != A
        n_dis += int(D.sum())
        n_band += int((D & bnd).sum())
        vals = bm.allLabels
        for l in range(1, nol):
            m = D & ((A == l) | (pred == l))
            tot_l[int(vals[l])] = tot_l.get(int(vals[l]), 0) + int(m.sum())
            band_l[int(vals[l])] = band_l.get(int(vals[l]), 0) + int((m & bnd).sum())
        out_m = D & ~bnd
        per_vol.append((lp[out_m].astype(np.float64), vals[A[out_m]], vals[pred[out_m]]))
        log("calibration volume %d: %s, %d labels, %d disagreements (%d outside the band)"
            % (i, dv.dims, nol, int(D.sum()), int((D & ~bnd).sum())))
    labels = [l for l, c in tot_l.items() if c >= min_count]
    recall, T = 0.0, 1.0
    for T in np.arange(1.0, 40.0001, 0.05):
        caught = dict(band_l)
        for lp, la, lb in per_vol:
            if lp.size == 0:
                continue
            z = lp / T
            z -= z.max(-1, keepdims=True)
            flag = (1.0 / np.exp(z).sum(-1)) < thr
            for l in labels:
                caught[l] += int((flag & ((la == l) | (lb == l))).sum())
        recall = min([caught[l] / tot_l[l] for l in labels] or [1.0])
        if recall >= target_recall:
            break
    net.T = float(T)
    net._head_inf = None
    log("temperature %.2f: worst per-label recall %.4f over labels %s; %d disagreements, %d in the boundary band"
        % (T, recall, sorted(labels), n_dis, n_band))
    return net.T, recall, {"disagreements": n_dis, "in_band": n_band, "per_label": tot_l}


def _copy(eng, dst, dst_off, src, nbytes):
    """Device-to-device copy into dst at byte offset dst_off."""
    assert dst_off + nbytes <= dst.nbytes and nbytes <= src.nbytes
    _hip.call("bmx_copy", eng.ctx, dst.ptr + dst_off, src.ptr, nbytes)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "calibrate"])
    ap.add_argument("--volumes", type=int, default=2000)
    ap.add_argument("--iters", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--size", type=int, default=64)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=WEIGHTS_PATH)
    ap.add_argument("--target", choices=["soft", "hard"], default="hard")
    ap.add_argument("--weights", default=WEIGHTS_PATH, help="calibrate: weights to calibrate")
    ap.add_argument("--recall", type=float, default=0.95)
    a = ap.parse_args()
    lg = lambda m: print("[%s] %s" % (a.cmd, m), flush=True)  # noqa: E731
    if a.cmd == "train":
        train(a.volumes, a.size, a.iters, a.batch, a.lr, out=a.out, hard=a.target == "hard", log=lg)
    else:
        model = Net(_hip.engine(0))
        model.load(a.weights)
        calibrate(model, n_volumes=a.volumes, size=a.size, target_recall=a.recall, log=lg)
        model.save(a.out)
