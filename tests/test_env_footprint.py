"""The environment lookup reads one footprint record per sample: the four
texels of the bilinear 2x2 around (u, v) and their pdf factors.  That is exact
only if the pdf's nearest texel (floor(u W), floor(v H), clamped) is always
one of those four corners (the bilinear base is floor(u W - 0.5) wrapped in
x, floor(v H - 0.5) clamped in y).  This checks that property in binary32
arithmetic, the kernels' own, over random and edge coordinates in [0, 1] (the
range the equirect mapping and the CDF inversion produce) and several map
sizes; the kernel falls back to the CDFs when it does not hold, so a
failure here would mean that fallback is reachable, not a wrong image."""
import numpy as np
import pytest

f32 = np.float32


def footprint(u, v, W, H):
    x = f32(u) * f32(W) - f32(0.5)
    y = f32(v) * f32(H) - f32(0.5)
    x = np.where(np.isnan(x), f32(0.0), x)
    y = np.where(np.isnan(y), f32(0.0), y)
    ix = np.floor(x).astype(np.int64)
    iy = np.floor(y).astype(np.int64)
    ix0 = np.mod(ix, W)
    ix1 = np.where(ix0 + 1 == W, 0, ix0 + 1)
    iy0 = np.clip(iy, 0, H - 1)
    iy1 = np.clip(iy + 1, 0, H - 1)
    return ix0, ix1, iy0, iy1


def pdf_texel(u, v, W, H):
    fx = f32(u) * f32(W)
    fy = f32(v) * f32(H)
    ix = np.clip(np.minimum(np.maximum(fx, f32(0.0)), f32(W)).astype(np.int64), 0, W - 1)
    iy = np.clip(np.minimum(np.maximum(fy, f32(0.0)), f32(H)).astype(np.int64), 0, H - 1)
    return ix, iy


def edge_values():
    one = f32(1.0)
    tiny = np.nextafter(f32(0.0), one)
    below_one = np.nextafter(one, f32(0.0))
    return np.array([0.0, tiny, 1e-7, 0.5, below_one, 1.0], np.float32)


@pytest.mark.parametrize("W,H", [(2048, 1024), (512, 256), (7, 3), (1, 1), (4096, 2048), (3, 1)])
def test_pdf_texel_is_a_footprint_corner(W, H):
    rs = np.random.default_rng(W * 131 + H)
    u = rs.random(200000, dtype=np.float32)
    v = rs.random(200000, dtype=np.float32)
    # texel borders and centres, where floor() changes
    k = np.arange(W + 1, dtype=np.float32)
    u = np.concatenate([u, k / f32(W), (k[:-1] + f32(0.5)) / f32(W), edge_values()])
    j = np.arange(H + 1, dtype=np.float32)
    v_extra = np.concatenate([j / f32(H), (j[:-1] + f32(0.5)) / f32(H), edge_values()])
    uu = np.concatenate([u, np.resize(u, v_extra.size)])
    vv = np.concatenate([np.resize(v, u.size), v_extra])
    ix0, ix1, iy0, iy1 = footprint(uu, vv, W, H)
    px, py = pdf_texel(uu, vv, W, H)
    ok = ((px == ix0) | (px == ix1)) & ((py == iy0) | (py == iy1))
    bad = np.flatnonzero(~ok)
    assert bad.size == 0, f"pdf texel outside the footprint at u={uu[bad[0]]!r} v={vv[bad[0]]!r}"


def test_a_shifted_base_would_break_the_property():
    # the check above is not vacuous: with the bilinear base taken one texel
    # to the right (floor(u W + 0.5)) the nearest texel falls outside the
    # footprint for about half of the samples
    W, H = 64, 32
    rs = np.random.default_rng(1)
    u = rs.random(10000, dtype=np.float32)
    v = rs.random(10000, dtype=np.float32)
    x = u * f32(W) + f32(0.5)
    ix0 = np.mod(np.floor(x).astype(np.int64), W)
    ix1 = np.where(ix0 + 1 == W, 0, ix0 + 1)
    px, _ = pdf_texel(u, v, W, H)
    outside = np.mean((px != ix0) & (px != ix1))
    assert outside > 0.3
