"""Pins the oracle's texture lookup (the texture unit's bilinear filter with
wrap addressing and normalized-float reads, src/host/scene_upload.cu:250-259)
against a float64 restatement of the same definition: texel centres at
(i + 0.5) / size, value = byte / 255, weights (1-a)(1-b), a(1-b), (1-a)b, ab.
Known answers: a texel centre returns that texel exactly; u = 0 lies halfway
between the last and the first column (wrap, not clamp); u and u + k give the
same value.  A planted clamp-addressing lookup fails the wrap case."""
import numpy as np

import oracle


def _ref(img, uv, wrap=True):
    img = img.astype(np.float64) / 255.0
    H, W = img.shape[:2]
    u = uv[:, 0].astype(np.float64)
    v = uv[:, 1].astype(np.float64)
    u = u - np.floor(u)
    v = v - np.floor(v)
    x = u * W - 0.5
    y = v * H - 0.5
    fx, fy = np.floor(x), np.floor(y)
    a, b = (x - fx)[:, None], (y - fy)[:, None]
    ix0, iy0 = fx.astype(int), fy.astype(int)
    if wrap:
        ix0 %= W
        iy0 %= H
        ix1, iy1 = (ix0 + 1) % W, (iy0 + 1) % H
    else:
        ix1, iy1 = np.clip(ix0 + 1, 0, W - 1), np.clip(iy0 + 1, 0, H - 1)
        ix0, iy0 = np.clip(ix0, 0, W - 1), np.clip(iy0, 0, H - 1)
    return ((1 - a) * (1 - b) * img[iy0, ix0] + a * (1 - b) * img[iy0, ix1] + (1 - a) * b * img[iy1, ix0]
            + a * b * img[iy1, ix1])


def test_texture_lookup_matches_bilinear_wrap():
    rs = np.random.default_rng(4)
    img = rs.integers(0, 256, (7, 11, 4), dtype=np.uint8)
    uv = rs.uniform(-3.0, 3.0, (20000, 2)).astype(np.float32)
    got = oracle.texture_sample(img, uv)
    assert np.abs(got - _ref(img, uv)).max() < 2e-6


def test_texture_known_answers():
    rs = np.random.default_rng(5)
    img = rs.integers(0, 256, (4, 8, 4), dtype=np.uint8)
    H, W = img.shape[:2]
    # texel centres
    ii, jj = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    uv = np.stack([(jj.ravel() + 0.5) / W, (ii.ravel() + 0.5) / H], 1).astype(np.float32)
    got = oracle.texture_sample(img, uv)
    assert np.array_equal(got, img.reshape(-1, 4).astype(np.float32) / np.float32(255.0))
    # u = 0 at a row centre: halfway between the last and the first column
    uv0 = np.array([[0.0, 0.5 / H]], np.float32)
    want = (img[0, W - 1].astype(np.float64) + img[0, 0]) / 2 / 255.0
    assert np.abs(oracle.texture_sample(img, uv0)[0] - want).max() < 1e-6
    # a planted clamp lookup gives the first column instead
    assert np.abs(_ref(img, uv0, wrap=False)[0] - want).max() > 1e-3 or np.array_equal(img[0, W - 1], img[0, 0])
    # periodic in both axes
    uvs = rs.uniform(0, 1, (500, 2)).astype(np.float32)
    shifted = (uvs + np.array([2.0, -3.0], np.float32)).astype(np.float32)
    assert np.abs(oracle.texture_sample(img, uvs) - oracle.texture_sample(img, shifted)).max() < 1e-5
