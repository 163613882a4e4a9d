"""ORACLE — test infrastructure only.

NumPy restatement of the learned denoiser (the U-Net that fills the slot of the
reference's denoiseAndPresent, src/host/denoiser.cu:72-109, whose OptiX model
is closed): forward pass, L1 loss in log space, backward pass and AdamW, with
the same storage roundings as the HIP kernels when ``bf16=True`` (activations,
gradients and weights stored as bfloat16, round-to-nearest-even; sums in
float64 here, float32 on the GPU).  With ``bf16=False`` everything stays in
float64, which the tests pin against torch.nn.functional.conv2d, torch
autograd and torch.optim.AdamW on the CPU.

Layer table and tensor order follow bluetruthengine_amd/csrc/bte_denoise_api.cpp.
"""
import numpy as np

# (cin real, cin stored, cout, relu) in canonical order
LAYERS = [
    ("c0a", 9, 32, 32, True), ("c0b", 32, 32, 32, True), ("c1a", 32, 32, 64, True), ("c1b", 64, 64, 64, True),
    ("c2a", 64, 64, 128, True), ("c2b", 128, 128, 128, True), ("c3a", 128, 128, 128, True),
    ("c3b", 128, 128, 128, True), ("d2a", 256, 256, 128, True), ("d2b", 128, 128, 128, True),
    ("d1a", 192, 192, 64, True), ("d1b", 64, 64, 64, True), ("d0a", 96, 96, 32, True), ("d0b", 32, 32, 32, True),
    ("out", 32, 32, 3, False),
]
NAMES = [l[0] for l in LAYERS]
# forward tensors in the order of the library's tensor ids 0..17, gradients 18..38
FORWARD = ["x", "e0a", "s0", "p0", "e1a", "s1", "p1", "e2a", "s2", "p2", "e3a", "b3", "d2a", "d2b", "d1a", "d1b",
           "d0a", "d0b"]
GRADS = ["go", "gd0b", "gd0a", "gcat0", "gd1b", "gd1a", "gcat1", "gd2b", "gd2a", "gcat2", "gb3", "ge3a", "gp2",
         "gs2", "ge2a", "gp1", "gs1", "ge1a", "gp0", "gs0", "ge0a"]
OUTPUT_ID = len(FORWARD) + len(GRADS)


def bf16_round(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = np.isnan(a)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = u.astype(np.uint32).view(np.float32).reshape(a.shape)
    return np.where(nan, np.float32(np.nan), out)


FIREFLY_RATIO = np.float32(6.0)


def features(rgb, alb, nrm):
    """Render buffers (H, W, 3 each, float32) -> the network input (1, H, W, 9):
    log(1 + rgb'), albedo, normal, where rgb' = max(rgb, 0) with firefly
    suppression: a pixel whose luminance 0.2126 r + 0.7152 g + 0.0722 b
    exceeds FIREFLY_RATIO times the largest luminance among its in-image 8
    neighbours is scaled by bound / luminance (binary32 as in k_features)."""
    f = np.float32
    c = np.maximum(np.asarray(rgb, f), f(0))
    H, W, _ = c.shape
    lum = (f(0.2126) * c[..., 0] + f(0.7152) * c[..., 1]) + f(0.0722) * c[..., 2]
    nb = np.zeros((H, W), f)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dx == 0 and dy == 0:
                continue
            sh = np.zeros((H, W), f)      # out-of-image neighbours count as 0 (luminance is >= 0)
            ys, yd = (slice(dy, H), slice(0, H - dy)) if dy >= 0 else (slice(0, H + dy), slice(-dy, H))
            xs, xd = (slice(dx, W), slice(0, W - dx)) if dx >= 0 else (slice(0, W + dx), slice(-dx, W))
            sh[yd, xd] = lum[ys, xs]
            nb = np.maximum(nb, sh)
    bound = FIREFLY_RATIO * nb
    hit = lum > bound
    k = np.where(hit, bound / np.where(hit, lum, f(1)), f(1)).astype(f)
    c = np.where(hit[..., None], c * k[..., None], c).astype(f)
    x = np.concatenate([np.log1p(c), np.asarray(alb, f), np.asarray(nrm, f)], -1)
    return x[None].astype(f)


def param_count():
    return sum(co * 9 * cr + co for _, cr, _, co, _ in LAYERS)


def split_params(flat):
    """canonical flat vector -> {name: (W[cout,3,3,cin_real], b[cout])}"""
    out, k = {}, 0
    for name, cr, _, co, _ in LAYERS:
        n = co * 9 * cr
        w = np.asarray(flat[k:k + n]).reshape(co, 3, 3, cr)
        k += n
        b = np.asarray(flat[k:k + co])
        k += co
        out[name] = (w, b)
    assert k == len(flat)
    return out


def join_params(p):
    return np.concatenate([np.concatenate([p[n][0].reshape(-1), p[n][1].reshape(-1)]) for n in NAMES])


def init_params(seed, scale_out=0.01):
    rs = np.random.default_rng(seed)
    p = {}
    for name, cr, _, co, _ in LAYERS:
        sd = np.sqrt(2.0 / (9 * cr)) * (scale_out if name == "out" else 1.0)
        p[name] = (rs.normal(0.0, sd, (co, 3, 3, cr)), np.zeros(co))
    return p


class Net:
    """U-Net restatement.  dt = float64, or bf16 storage emulation."""

    def __init__(self, params, bf16=True):
        self.bf16 = bf16
        self.p = {}
        for name, (w, b) in params.items():
            w = np.asarray(w, np.float64)
            if bf16:
                self.p[name] = (bf16_round(w.astype(np.float32)).astype(np.float64), np.asarray(b, np.float32))
            else:
                self.p[name] = (w, np.asarray(b, np.float64))

    # storage rounding of a tensor written by a kernel
    def _store(self, x):
        return bf16_round(np.asarray(x, np.float32)).astype(np.float64) if self.bf16 else x

    def _f32(self, x):
        return np.asarray(x, np.float32) if self.bf16 else x

    @staticmethod
    def _patches(x):
        """(n, h, w, c) -> (n, h, w, 9, c) zero-padded 3x3 neighbourhoods, tap = ky*3 + kx."""
        n, h, w, c = x.shape
        xp = np.zeros((n, h + 2, w + 2, c), x.dtype)
        xp[:, 1:-1, 1:-1] = x
        return np.stack([xp[:, ky:ky + h, kx:kx + w] for ky in range(3) for kx in range(3)], axis=3)

    def conv_raw(self, x, w):
        """sum over taps and channels (float64): x (n,h,w,cin), w (cout,3,3,cin)."""
        pt = self._patches(np.asarray(x, np.float64))
        n, h, ww = x.shape[:3]
        return pt.reshape(n * h * ww, -1) @ np.asarray(w, np.float64).reshape(w.shape[0], -1).T

    def conv(self, name, x, relu, resid=None):
        w, b = self.p[name]
        acc = self.conv_raw(x, w).reshape(x.shape[:3] + (w.shape[0],))
        if self.bf16:
            v = acc.astype(np.float32) + b.astype(np.float32)
            if resid is not None:
                v = v + resid.astype(np.float32)
            if relu:
                v = np.maximum(v, np.float32(0))
            return v.astype(np.float64) if name == "out" else self._store(v)
        v = acc + b
        if resid is not None:
            v = v + resid
        return np.maximum(v, 0.0) if relu else v

    def pool(self, x):
        if self.bf16:
            x = np.asarray(x, np.float32)
            s = ((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + x[:, 1::2, 0::2]) + x[:, 1::2, 1::2]
            return self._store(np.float32(0.25) * s)
        return 0.25 * (x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2])

    @staticmethod
    def up(x):
        return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)

    def forward(self, x9):
        """x9 (n, h, w, 9) -> dict of tensors; 'o' (n, h, w, 3) is the log-space output."""
        t = {"x": self._store(np.asarray(x9, np.float64))}
        x = t["x"]
        t["e0a"] = self.conv("c0a", x, True)
        t["s0"] = self.conv("c0b", t["e0a"], True)
        t["p0"] = self.pool(t["s0"])
        t["e1a"] = self.conv("c1a", t["p0"], True)
        t["s1"] = self.conv("c1b", t["e1a"], True)
        t["p1"] = self.pool(t["s1"])
        t["e2a"] = self.conv("c2a", t["p1"], True)
        t["s2"] = self.conv("c2b", t["e2a"], True)
        t["p2"] = self.pool(t["s2"])
        t["e3a"] = self.conv("c3a", t["p2"], True)
        t["b3"] = self.conv("c3b", t["e3a"], True)
        t["d2a"] = self.conv("d2a", np.concatenate([self.up(t["b3"]), t["s2"]], -1), True)
        t["d2b"] = self.conv("d2b", t["d2a"], True)
        t["d1a"] = self.conv("d1a", np.concatenate([self.up(t["d2b"]), t["s1"]], -1), True)
        t["d1b"] = self.conv("d1b", t["d1a"], True)
        t["d0a"] = self.conv("d0a", np.concatenate([self.up(t["d1b"]), t["s0"]], -1), True)
        t["d0b"] = self.conv("d0b", t["d0a"], True)
        t["o"] = self.conv("out", t["d0b"], False, resid=x[..., :3])
        return t

    # ---- backward ---------------------------------------------------------
    def _dgrad(self, name, g, mask=None):
        """gradient w.r.t. the layer input: correlation of g with the flipped,
        transposed kernel; times (mask > 0) when the input is a ReLU output."""
        w = self.p[name][0]
        wt = np.transpose(w[:, ::-1, ::-1, :], (3, 1, 2, 0))   # (cin, 3, 3, cout)
        acc = self.conv_raw(g, wt).reshape(g.shape[:3] + (w.shape[3],))
        v = acc.astype(np.float32) if self.bf16 else acc
        if mask is not None:
            v = np.where(mask > 0, v, 0)
        return self._store(v)

    def _wgrad(self, g, x):
        """dW (cout, 3, 3, cin), db (cout) of one layer from its output gradient and input."""
        pt = self._patches(np.asarray(x, np.float64))
        n, h, w = x.shape[:3]
        gw = np.asarray(g, np.float64).reshape(n * h * w, -1)
        dw = gw.T @ pt.reshape(n * h * w, -1)
        return dw.reshape(g.shape[-1], 3, 3, x.shape[-1]), gw.sum(axis=0)

    @staticmethod
    def _up_bwd(g):
        return ((g[:, 0::2, 0::2] + g[:, 0::2, 1::2]) + g[:, 1::2, 0::2]) + g[:, 1::2, 1::2]

    def loss_grad(self, o, target, mode=0):
        """Training loss and its gradient w.r.t. the log-space output o.
        mode 0: mean |o - t|.  mode 1: mean (y - T)^2 / (max(y, 0)^2 + 0.01),
        y = e^o - 1, T = e^t - 1, the denominator held constant, gradient
        clipped to [-4, 4] before the 1 / count scale."""
        count = o.size
        inv = np.float32(1) / np.float32(count) if self.bf16 else 1.0 / count
        if mode == 0:
            d = o - target
            return np.abs(d).sum() / count, np.sign(d) * inv
        eo = np.exp(o)
        y = eo - 1.0
        T = np.expm1(target)
        den = np.maximum(y, 0.0) ** 2 + 0.01
        d = y - T
        g = np.clip(2.0 * d * eo / den, -4.0, 4.0)
        return (d * d / den).sum() / count, (g.astype(np.float32) * inv) if self.bf16 else g * inv

    def backward(self, t, target, loss_mode=0):
        """Loss (mean L1 in log space, or relMSE with loss_mode=1) and
        parameter gradients {name: (dW, db)}."""
        o = t["o"]
        loss, go = self.loss_grad(o, target, loss_mode)
        g = {"go": self._store(go)}
        grads = {}
        st = self._store
        f32 = self._f32

        def wg(name, gout, xin):
            dw, db = self._wgrad(gout, xin)
            grads[name] = (dw[..., :self.p[name][0].shape[3]], db)

        wg("out", g["go"], t["d0b"])
        g["gd0b"] = self._dgrad("out", g["go"], t["d0b"])
        wg("d0b", g["gd0b"], t["d0a"])
        g["gd0a"] = self._dgrad("d0b", g["gd0b"], t["d0a"])
        wg("d0a", g["gd0a"], np.concatenate([self.up(t["d1b"]), t["s0"]], -1))
        g["gcat0"] = self._dgrad("d0a", g["gd0a"])
        g["gd1b"] = st(np.where(t["d1b"] > 0, self._up_bwd(f32(g["gcat0"][..., :64])), 0))
        wg("d1b", g["gd1b"], t["d1a"])
        g["gd1a"] = self._dgrad("d1b", g["gd1b"], t["d1a"])
        wg("d1a", g["gd1a"], np.concatenate([self.up(t["d2b"]), t["s1"]], -1))
        g["gcat1"] = self._dgrad("d1a", g["gd1a"])
        g["gd2b"] = st(np.where(t["d2b"] > 0, self._up_bwd(f32(g["gcat1"][..., :128])), 0))
        wg("d2b", g["gd2b"], t["d2a"])
        g["gd2a"] = self._dgrad("d2b", g["gd2b"], t["d2a"])
        wg("d2a", g["gd2a"], np.concatenate([self.up(t["b3"]), t["s2"]], -1))
        g["gcat2"] = self._dgrad("d2a", g["gd2a"])
        g["gb3"] = st(np.where(t["b3"] > 0, self._up_bwd(f32(g["gcat2"][..., :128])), 0))
        wg("c3b", g["gb3"], t["e3a"])
        g["ge3a"] = self._dgrad("c3b", g["gb3"], t["e3a"])
        wg("c3a", g["ge3a"], t["p2"])
        g["gp2"] = self._dgrad("c3a", g["ge3a"])
        quarter = np.float32(0.25) if self.bf16 else 0.25

        def pool_bwd_add(gp, gcat, act):
            v = quarter * f32(self.up(gp)) + f32(gcat)
            return st(np.where(act > 0, v, 0))

        g["gs2"] = pool_bwd_add(g["gp2"], g["gcat2"][..., 128:], t["s2"])
        wg("c2b", g["gs2"], t["e2a"])
        g["ge2a"] = self._dgrad("c2b", g["gs2"], t["e2a"])
        wg("c2a", g["ge2a"], t["p1"])
        g["gp1"] = self._dgrad("c2a", g["ge2a"])
        g["gs1"] = pool_bwd_add(g["gp1"], g["gcat1"][..., 128:], t["s1"])
        wg("c1b", g["gs1"], t["e1a"])
        g["ge1a"] = self._dgrad("c1b", g["gs1"], t["e1a"])
        wg("c1a", g["ge1a"], t["p0"])
        g["gp0"] = self._dgrad("c1a", g["ge1a"])
        g["gs0"] = pool_bwd_add(g["gp0"], g["gcat0"][..., 64:], t["s0"])
        wg("c0b", g["gs0"], t["e0a"])
        g["ge0a"] = self._dgrad("c0b", g["gs0"], t["e0a"])
        wg("c0a", g["ge0a"], t["x"][..., :9])
        return loss, grads, g


def adamw(w, g, m, v, step, lr, wd, decay, beta1=0.9, beta2=0.999, eps=1e-8, dtype=np.float32):
    """One AdamW step in the library's order: w *= 1 - lr*wd (weights only),
    then w -= lr * mhat / (sqrt(vhat) + eps)."""
    f = dtype
    w, g, m, v = (np.asarray(a, f) for a in (w, g, m, v))
    one = f(1)
    m = f(beta1) * m + (one - f(beta1)) * g
    v = f(beta2) * v + (one - f(beta2)) * g * g
    if decay:
        w = w * (one - f(lr) * f(wd))
    # formed in binary64 from the binary32 betas, then rounded, as on the host
    bc1 = f(1 - float(f(beta1)) ** step)
    bc2 = f(1 - float(f(beta2)) ** step)
    w = w - f(lr) * (m / bc1) / (np.sqrt(v / bc2) + f(eps))
    return w, m, v
