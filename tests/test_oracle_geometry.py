"""Pins the oracle's ray casting (closest hit / occlusion) and camera.

* Closest hit: an independent binary64 Moller-Trumbore brute force over all
  triangles decides the reference answer; rays are kept when that answer is
  unambiguous (barycentric margin 1e-4, next-closest t gap 1e-4 relative), the
  "non-degenerate rays" of the north star.
* The oracle's own BVH agrees with its O(T) loop on every ray (tie rule: equal
  t resolves to the lower triangle index).
* Camera: rt_camera.cuh:15-37 in binary64 within 1e-6.
"""
import numpy as np

import oracle
from bluetruthengine_amd import scenes


def mt_all(rays, tris):
    """binary64 Moller-Trumbore: t [R, T] (inf on miss) and min barycentric."""
    o = rays[:, None, :3].astype(np.float64)
    d = rays[:, None, 3:].astype(np.float64)
    v0 = tris["v0"].astype(np.float64)[None]
    e1 = tris["v1"].astype(np.float64)[None] - v0
    e2 = tris["v2"].astype(np.float64)[None] - v0
    p = np.cross(d, e2)
    det = np.sum(e1 * p, axis=-1)
    inv = 1.0 / np.where(det == 0, np.inf, det)
    s = o - v0
    u = np.sum(s * p, axis=-1) * inv
    q = np.cross(s, e1)
    v = np.sum(d * q, axis=-1) * inv
    t = np.sum(e2 * q, axis=-1) * inv
    w = 1 - u - v
    bmin = np.minimum(np.minimum(u, v), w)
    hit = (bmin >= 0) & (t >= 1e-4) & (det != 0)
    return np.where(hit, t, np.inf), bmin, u, v


def _rays(n, seed):
    rs = np.random.default_rng(seed)
    o = rs.uniform((-0.95, 0.05, -0.95), (0.95, 1.95, 0.95), (n, 3))
    d = rs.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)


def test_closest_hit_vs_binary64_brute_force():
    sd = scenes.cornell_c1(width=32, height=32)
    ref = oracle.OracleScene(sd)
    rays = _rays(3000, 1)
    tri, t, u, v = ref.trace_closest(rays)
    T, bmin, _, _ = mt_all(rays, sd.triangles)
    best = np.argmin(T, axis=1)
    tb = T[np.arange(len(rays)), best]
    Ts = np.sort(T, axis=1)
    with np.errstate(invalid="ignore"):
        gap = np.where(np.isfinite(Ts[:, 1]), (Ts[:, 1] - Ts[:, 0]) / np.maximum(Ts[:, 0], 1e-9), np.inf)
    clear = np.isfinite(tb) & (bmin[np.arange(len(rays)), best] > 1e-4) & (gap > 1e-4)
    miss = ~np.isfinite(tb)
    # misses that pass no triangle edge within the margin
    clear_miss = miss & ~np.any((bmin > -1e-4) & (bmin < 1e-4), axis=1)
    assert clear.mean() > 0.8
    assert np.array_equal(tri[clear], best[clear])
    assert np.allclose(t[clear], tb[clear], rtol=1e-5)
    assert np.all(tri[clear_miss] == -1)


def test_bvh_equals_brute_force_bitwise():
    sd = scenes.small_envlit(n_spheres=20)
    ref = oracle.OracleScene(sd)
    rs = np.random.default_rng(2)
    o = rs.uniform((-4, 0.01, -4), (4, 0.6, 4), (20000, 3))
    d = rs.normal(size=(20000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1).astype(np.float32)
    a = ref.trace_closest(rays)
    b = ref.trace_closest(rays, brute=True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert 0.1 < np.mean(a[0] >= 0) < 0.99


def test_occlusion_consistent_with_closest_hit():
    sd = scenes.cornell_c1(width=32, height=32)
    ref = oracle.OracleScene(sd)
    rays = _rays(20000, 3)
    tri, t, _, _ = ref.trace_closest(rays)
    tmax = np.random.default_rng(4).uniform(0.01, 3.0, len(rays)).astype(np.float32)
    occ = ref.trace_occluded(rays, tmax)
    expect = (tri >= 0) & (t <= tmax)
    assert np.array_equal(occ.astype(bool), expect)


def test_primary_ray_closed_form():
    sd = scenes.cornell_c1(width=37, height=23)
    cam = sd.camera
    L = oracle.lib()
    rs = np.random.default_rng(9)
    for _ in range(200):
        px, py = int(rs.integers(0, 37)), int(rs.integers(0, 23))
        jx, jy = rs.uniform(-0.5, 0.5, 2).astype(np.float32)
        out = np.zeros(3, np.float32)
        L.ora_primary_ray(oracle._ptr(np.ascontiguousarray(cam)), px, py, 37, 23, float(jx), float(jy),
                          oracle._ptr(out))
        fx = (px + 0.5 + float(jx)) / 37
        fy = (py + 0.5 + float(jy)) / 23
        th = np.tan(0.5 * float(cam["fovYRadians"]))
        d = cam["forward"].astype(np.float64) + cam["right"] * ((2 * fx - 1) * float(cam["aspect"]) * th) \
            + cam["up"] * ((2 * fy - 1) * th)
        d /= np.linalg.norm(d)
        assert np.allclose(out, d, atol=1e-6)
    # centre of an odd image with zero jitter looks straight down the forward axis
    sd1 = scenes.cornell_c1(width=33, height=33)
    L.ora_primary_ray(oracle._ptr(np.ascontiguousarray(sd1.camera)), 16, 16, 33, 33, 0.0, 0.0, oracle._ptr(out))
    assert np.allclose(out, sd1.camera["forward"], atol=1e-7)


def _near_tie_scene():
    """Triangles at nearly equal distances along many rays: 60 overlapping
    coplanar quads at y = 1.98 (like the terrain room's ceiling lights) and 40
    mutually intersecting icospheres resting on a floor quad."""
    from bluetruthengine_amd import pod
    sb = scenes.SceneBuilder()
    b = sb.add_bsdf(pod.bsdf_diffuse((0.5, 0.5, 0.5)))
    rs = np.random.default_rng(21)
    for _ in range(60):
        sx, sz = rs.uniform(0.2, 0.5, 2)
        cx, cz = rs.uniform(-0.7, 0.7, 2)
        sb.add_mesh(scenes.ceiling_light(cx, cz, sx, sz, 1.98), None, b)
    sb.add_mesh(scenes.quad((-2, 0, -2), (-2, 0, 2), (2, 0, 2), (2, 0, -2)), None, b)
    v, f = scenes.icosphere(2)
    for _ in range(40):
        r = rs.uniform(0.1, 0.3)
        c = np.array([rs.uniform(-0.8, 0.8), r, rs.uniform(-0.8, 0.8)])
        p, n = scenes.mesh_arrays(v * r + c, f, v)
        sb.add_mesh(p, n, b)
    cam = scenes.make_camera((0, 1, 3.6), (0, 1, 0), (0, 1, 0), 40.0, 16, 16)
    return sb.build("near_ties", 16, 16, 1, 4, cam)


def test_bvh_equals_brute_force_on_near_ties():
    # the box test must not prune a box whose triangle's computed distance is
    # within rounding of the box entry: with overlapping coplanar quads the
    # closest of two nearly equal distances must not depend on the tree
    sd = _near_tie_scene()
    ref = oracle.OracleScene(sd)
    rs = np.random.default_rng(22)
    n = 200_000
    o = rs.uniform((-0.9, 0.05, -0.9), (0.9, 1.2, 0.9), (n, 3))
    tgt = np.stack([rs.uniform(-0.9, 0.9, n), np.full(n, 1.98), rs.uniform(-0.9, 0.9, n)], 1)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    up = np.concatenate([o, d], 1).astype(np.float32)
    down = np.concatenate([o, rs.normal(size=(n, 3))], 1).astype(np.float32)
    down[:, 3:] /= np.linalg.norm(down[:, 3:], axis=1, keepdims=True)
    for rays in (up, down):
        a = ref.trace_closest(rays)
        b = ref.trace_closest(rays, brute=True)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        # any hit with tMax equal to the closest distance: the closest triangle itself occludes
        hit = a[0] >= 0
        occ = ref.trace_occluded(rays[hit], a[1][hit])
        assert occ.all()
