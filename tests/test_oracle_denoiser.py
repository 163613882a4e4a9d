"""Pins the denoiser oracle (oracle/denoiser.py) against PyTorch on the CPU.

* a single 3x3 convolution equals torch.nn.functional.conv2d (float64);
* the whole U-Net forward pass equals the same network written with torch
  functional ops (conv2d, avg_pool2d, nearest interpolate, cat, relu);
* the hand-written backward pass equals torch autograd of the mean L1 loss;
* one AdamW step equals torch.optim.AdamW;
* bfloat16 rounding is round-to-nearest-even (known answers).
A planted mistake in the oracle (a wrong flip of the transposed kernel, a
missing ReLU mask, a skipped 1/4 of the pooling gradient) breaks the backward
comparison; the test at the end checks that.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import denoiser as od


def _torch_net(params, x9, target):
    P = {n: (torch.tensor(np.transpose(w, (0, 3, 1, 2)).copy(), requires_grad=True),
             torch.tensor(b.copy(), requires_grad=True)) for n, (w, b) in params.items()}
    x = torch.tensor(np.transpose(x9, (0, 3, 1, 2)).copy())

    def c(n, t, relu=True):
        y = F.conv2d(t, P[n][0], P[n][1], padding=1)
        return F.relu(y) if relu else y

    def up(t):
        return F.interpolate(t, scale_factor=2, mode="nearest")

    e0a = c("c0a", x)
    s0 = c("c0b", e0a)
    e1a = c("c1a", F.avg_pool2d(s0, 2))
    s1 = c("c1b", e1a)
    e2a = c("c2a", F.avg_pool2d(s1, 2))
    s2 = c("c2b", e2a)
    e3a = c("c3a", F.avg_pool2d(s2, 2))
    b3 = c("c3b", e3a)
    d2b = c("d2b", c("d2a", torch.cat([up(b3), s2], 1)))
    d1b = c("d1b", c("d1a", torch.cat([up(d2b), s1], 1)))
    d0b = c("d0b", c("d0a", torch.cat([up(d1b), s0], 1)))
    o = c("out", d0b, relu=False) + x[:, :3]
    loss = (o - torch.tensor(np.transpose(target, (0, 3, 1, 2)).copy())).abs().mean()
    loss.backward()
    grads = {n: (np.transpose(P[n][0].grad.numpy(), (0, 2, 3, 1)), P[n][1].grad.numpy()) for n in P}
    return np.transpose(o.detach().numpy(), (0, 2, 3, 1)), loss.item(), grads


def _case(seed=0, n=2, h=16, w=24):
    rs = np.random.default_rng(seed)
    params = od.init_params(seed, scale_out=1.0)
    for name in params:   # non-zero biases exercise the bias paths
        params[name] = (params[name][0], rs.normal(0, 0.05, params[name][1].shape))
    x9 = rs.normal(0, 1, (n, h, w, 9))
    target = rs.normal(0, 1, (n, h, w, 3))
    return params, x9, target


def test_bf16_rounding_known_answers():
    one = 1.0
    vals = np.array([one, one + 2 ** -8, one + 3 * 2 ** -8, one + 2 ** -8 + 2 ** -20, -2.5, 0.0, np.inf], np.float32)
    want = np.array([one, one, one + 2 ** -6, one + 2 ** -7, -2.5, 0.0, np.inf], np.float32)
    assert np.array_equal(od.bf16_round(vals), want)
    assert np.isnan(od.bf16_round(np.array([np.nan], np.float32)))[0]


def test_conv_matches_torch_conv2d():
    rs = np.random.default_rng(3)
    x = rs.normal(size=(2, 9, 11, 5))
    w = rs.normal(size=(7, 3, 3, 5))
    net = od.Net({}, bf16=False)
    got = net.conv_raw(x, w).reshape(2, 9, 11, 7)
    ref = F.conv2d(torch.tensor(x).permute(0, 3, 1, 2), torch.tensor(w).permute(0, 3, 1, 2), padding=1)
    assert np.allclose(got, ref.permute(0, 2, 3, 1).numpy(), rtol=1e-12, atol=1e-12)


def test_forward_and_backward_match_torch_autograd():
    params, x9, target = _case()
    net = od.Net(params, bf16=False)
    t = net.forward(x9)
    loss, grads, _ = net.backward(t, target)
    o_ref, loss_ref, g_ref = _torch_net(params, x9, target)
    assert np.allclose(t["o"], o_ref, rtol=1e-10, atol=1e-10)
    assert abs(loss - loss_ref) < 1e-12
    for name in od.NAMES:
        dw, db = grads[name]
        rw, rb = g_ref[name]
        scale = np.abs(rw).max() + 1e-30
        assert np.abs(dw - rw).max() <= 1e-9 * scale, name
        assert np.allclose(db, rb, rtol=1e-9, atol=1e-12 * scale), name


def test_relmse_loss_gradient_matches_torch():
    # the relMSE loss with the prediction held constant in the denominator;
    # values kept small enough that the [-4, 4] clip is inactive
    rs = np.random.default_rng(9)
    o = rs.uniform(0.5, 1.5, (1, 4, 8, 3))
    tgt = o + rs.normal(0, 0.02, o.shape)
    net = od.Net({}, bf16=False)
    loss, g = net.loss_grad(o, tgt, mode=1)
    ot = torch.tensor(o, requires_grad=True)
    y = torch.exp(ot) - 1
    den = torch.clamp(y.detach(), min=0) ** 2 + 0.01
    lt = ((y - torch.expm1(torch.tensor(tgt))) ** 2 / den).mean()
    lt.backward()
    assert abs(loss - lt.item()) < 1e-12
    assert np.all(np.abs(g * o.size) < 4)
    assert np.allclose(g, ot.grad.numpy(), rtol=1e-10, atol=1e-14)


def test_adamw_matches_torch():
    rs = np.random.default_rng(5)
    w0 = rs.normal(size=50)
    g = rs.normal(size=50)
    m0 = rs.normal(size=50) * 0.1
    v0 = np.abs(rs.normal(size=50)) * 0.01
    p = torch.tensor(w0.copy(), requires_grad=True)
    opt = torch.optim.AdamW([p], lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.05)
    st = opt.state[p]
    st["step"] = torch.tensor(6.0)
    st["exp_avg"] = torch.tensor(m0.copy())
    st["exp_avg_sq"] = torch.tensor(v0.copy())
    p.grad = torch.tensor(g.copy())
    opt.step()
    w1, _, _ = od.adamw(w0, g, m0, v0, 7, 1e-3, 0.05, True, dtype=np.float64)
    assert np.allclose(w1, p.detach().numpy(), rtol=1e-12, atol=1e-15)


def test_planted_backward_mistakes_are_caught(monkeypatch):
    params, x9, target = _case(seed=1, n=1, h=16, w=16)
    _, _, g_ref = _torch_net(params, x9, target)

    def worst(net):
        t = net.forward(x9)
        _, grads, _ = net.backward(t, target)
        return max(np.abs(grads[n][0] - g_ref[n][0]).max() / (np.abs(g_ref[n][0]).max() + 1e-30) for n in od.NAMES)

    assert worst(od.Net(params, bf16=False)) < 1e-9
    # 1: kernel not flipped in the input gradient
    orig = od.Net._dgrad

    def no_flip(self, name, g, mask=None):
        w = self.p[name][0]
        wt = np.transpose(w, (3, 1, 2, 0))
        acc = self.conv_raw(g, wt).reshape(g.shape[:3] + (w.shape[3],))
        return np.where(mask > 0, acc, 0) if mask is not None else acc

    monkeypatch.setattr(od.Net, "_dgrad", no_flip)
    assert worst(od.Net(params, bf16=False)) > 1e-3
    monkeypatch.setattr(od.Net, "_dgrad", orig)
    # 2: ReLU mask dropped
    monkeypatch.setattr(od.Net, "_dgrad", lambda self, name, g, mask=None: orig(self, name, g, None))
    assert worst(od.Net(params, bf16=False)) > 1e-3


def _nb_max_torch(lum):
    """largest luminance among the in-image 8 neighbours, through
    torch.nn.functional.unfold (zero padding; luminance is >= 0)"""
    t = torch.tensor(lum, dtype=torch.float64)[None, None]
    cols = F.unfold(t, 3, padding=1)[0]                   # (9, H*W), row 4 = the pixel itself
    return torch.cat([cols[:4], cols[5:]]).max(0).values.reshape(lum.shape).numpy()


def test_features_firefly_suppression():
    rs = np.random.default_rng(4)
    H, W = 12, 17
    rgb = np.full((H, W, 3), 0.5, np.float32) + rs.uniform(0, 0.1, (H, W, 3)).astype(np.float32)
    alb = rs.uniform(0, 1, (H, W, 3)).astype(np.float32)
    nrm = rs.normal(size=(H, W, 3)).astype(np.float32)
    rgb[5, 7] = [100.0, 50.0, 25.0]          # isolated firefly
    rgb[0, 0] = [40.0, 40.0, 40.0]           # corner firefly: 3 neighbours only
    rgb[9, 3] = rgb[9, 4] = [30.0, 30.0, 30.0]   # a 2-pixel highlight is no firefly
    rgb[2, 12] = [-1.0, 0.2, 0.3]            # negative radiance clamps to 0
    x = od.features(rgb, alb, nrm)
    assert x.shape == (1, H, W, 9) and x.dtype == np.float32
    assert np.array_equal(x[0, ..., 3:6], alb) and np.array_equal(x[0, ..., 6:], nrm)
    out = np.expm1(x[0, ..., :3].astype(np.float64))
    c = np.maximum(rgb.astype(np.float64), 0)
    lum = c @ [0.2126, 0.7152, 0.0722]
    nb = _nb_max_torch(lum)
    for (y, xx) in ((5, 7), (0, 0)):
        lo = out[y, xx] @ [0.2126, 0.7152, 0.0722]
        assert np.isclose(lo, 6.0 * nb[y, xx], rtol=1e-5), (y, xx)
        assert np.allclose(out[y, xx] / out[y, xx].sum(), c[y, xx] / c[y, xx].sum(), rtol=1e-5)   # colour kept
    # every other pixel is unchanged (up to log1p / expm1 rounding)
    keep = lum <= 6.0 * nb
    assert keep.sum() == H * W - 2
    assert np.allclose(out[keep], c[keep], rtol=1e-5, atol=1e-7)
    assert np.allclose(out[9, 3], 30.0, rtol=1e-5) and out[2, 12, 0] == 0.0
    # exactly at the bound stays (strict inequality)
    g = np.ones((3, 3, 3), np.float32)
    g[1, 1] = 6.0
    assert np.allclose(np.expm1(od.features(g, g, g)[0, 1, 1, :3]), 6.0, rtol=1e-6)
