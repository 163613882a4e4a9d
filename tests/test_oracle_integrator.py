"""Pins the oracle integrator as a whole (NEE + MIS + BSDF sampling + Russian
roulette) with furnace scenes whose expected pixel values are closed forms:

* closed box, every wall Lambertian (albedo rho) and emissive (Le), camera
  inside:  E[L] = Le / (1 - rho)  (area NEE and emitter-hit MIS must sum to 1);
* convex Lambertian sphere under a constant environment Le:  E[L] = rho Le
  (environment NEE and environment-hit MIS must sum to 1);
* mirror sphere: L = tint * Le exactly, every sample;
* glass sphere (IOR 1.5) under a constant environment:  E[L] = Le (radiance
  scaling by eta^2 on refraction and the eta-aware roulette must cancel);
* a scene without lights renders exactly black.
Tolerances are several standard errors of the Monte Carlo estimate.
"""
import numpy as np

import oracle
from bluetruthengine_amd import pod, scenes


def _closed_box(rho, le):
    sb = scenes.SceneBuilder()
    b = sb.add_bsdf(pod.bsdf_diffuse((rho, rho, rho)))
    walls = list(scenes.room_quads().reshape(5, 2, 3, 3))
    walls.append(scenes.quad((-1, 0, 1), (-1, 2, 1), (1, 2, 1), (1, 0, 1)))
    for w in walls:
        sb.add_mesh(w, None, b, emission=(le, le, le))
    cam = scenes.make_camera((0, 1, 0), (0, 1, -1), (0, 1, 0), 60.0, 24, 24)
    return sb.build("furnace_box", 24, 24, 64, 200, cam)


def test_closed_emissive_box_furnace():
    sd = _closed_box(0.5, 1.0)
    ref = oracle.OracleScene(sd)
    acc = ref.render(24, 24, 0, 64, 200)
    m = acc.mean()
    assert abs(m - 2.0) < 0.04, m
    assert np.allclose(acc[..., 0], acc[..., 1]) and np.allclose(acc[..., 0], acc[..., 2])


def _sphere_in_env(bsdf, le=1.0, n=24):
    sb = scenes.SceneBuilder()
    b = sb.add_bsdf(bsdf)
    v, f = scenes.icosphere(4)
    p, nn = scenes.mesh_arrays(v * 0.5, f, v)
    sb.add_mesh(p, nn, b)
    env = np.full((32, 64, 4), le, np.float32)
    cam = scenes.make_camera((0, 0, 3), (0, 0, 0), (0, 1, 0), 12.0, n, n)
    return sb.build("env_furnace", n, n, 1, 200, cam, env=env)


def _sphere_pixels(ref, sd):
    pix = np.arange(sd.width * sd.height, dtype=np.int32)
    _, tris = ref.trace_samples(pix, sd.width, sd.height, 0, sd.max_depth, max_rec=1)
    return tris[:, 0] >= 0


def test_lambertian_sphere_env_furnace():
    sd = _sphere_in_env(pod.bsdf_diffuse((0.6, 0.6, 0.6)))
    ref = oracle.OracleScene(sd)
    on = _sphere_pixels(ref, sd)
    acc = ref.render(sd.width, sd.height, 0, 256, sd.max_depth).reshape(-1, 3)
    assert on.mean() > 0.5
    assert abs(acc[on].mean() - 0.6) < 0.01, acc[on].mean()
    assert np.allclose(acc[~on], 1.0)


def test_mirror_sphere_is_exact():
    sd = _sphere_in_env(pod.bsdf_mirror((0.9, 0.8, 0.7)), le=2.0)
    ref = oracle.OracleScene(sd)
    on = _sphere_pixels(ref, sd)
    acc = ref.render(sd.width, sd.height, 0, 4, sd.max_depth).reshape(-1, 3)
    assert np.allclose(acc[on], [1.8, 1.6, 1.4], rtol=1e-6)


def test_glass_sphere_env_furnace():
    sd = _sphere_in_env(pod.bsdf_dielectric(1.5, 1.0))
    ref = oracle.OracleScene(sd)
    on = _sphere_pixels(ref, sd)
    acc = ref.render(sd.width, sd.height, 0, 128, sd.max_depth).reshape(-1, 3)
    assert abs(acc[on].mean() - 1.0) < 0.02, acc[on].mean()


def test_no_lights_is_black():
    sb = scenes.SceneBuilder()
    b = sb.add_bsdf(pod.bsdf_diffuse((0.7, 0.7, 0.7)))
    sb.add_mesh(scenes.room_quads(), None, b)
    cam = scenes.make_camera((0, 1, 3.6), (0, 1, 0), (0, 1, 0), 40.0, 16, 16)
    sd = sb.build("dark", 16, 16, 2, 8, cam)
    acc = oracle.OracleScene(sd).render(16, 16, 0, 2, 8)
    assert np.all(acc == 0.0)


def test_render_is_deterministic_and_frame_indexed():
    sd = scenes.cornell_c1(width=16, height=16)
    ref = oracle.OracleScene(sd)
    a = ref.render(16, 16, 0, 3, 8)
    b = ref.render(16, 16, 0, 2, 8)
    b = ref.render(16, 16, 2, 1, 8, accum=b)
    assert np.array_equal(a, b)
