"""Scene generators and flattening (src/main.cpp:686-758 semantics)."""
import json
import os

import numpy as np

from bluetruthengine_amd import pod, scenes

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "scene_digests.json")))


def test_generator_digests_are_stable():
    assert scenes.cornell_c1().digest() == GOLD["c1_cornell_256"]
    assert scenes.small_envlit().digest() == GOLD["c3_small_128x72"]
    assert scenes.terrain_c2(width=96, height=54, spp=2, freq=20).digest() == GOLD["c2_terrain_freq20"]


def test_c1_emitter_table_closed_form():
    sd = scenes.cornell_c1()
    assert sd.triangles.shape[0] == 10 + 2 + 2 * 1280
    e = sd.emitters
    assert e.shape[0] == 1 and e[0]["triCount"] == 2 and e[0]["areaSum"] == np.float32(0.25)
    assert np.array_equal(sd.emitter_tri_cdf, np.array([0.0, 0.125, 0.25], np.float32))
    assert np.array_equal(sd.scene_emitter_cdf, np.array([0.0, 0.25 * 15.0], np.float32))
    flags = sd.tri_emitter_flags
    assert flags.sum() == 2 and np.all(sd.tri_emission[flags == 1] == 15.0)
    # emitter faces down: geometric normal (v1-v0)x(v2-v0) has negative y
    t = sd.triangles[sd.emitter_tri_indices]
    n = np.cross(t["v1"] - t["v0"], t["v2"] - t["v0"])
    assert np.all(n[:, 1] < 0)


def test_camera_basis():
    sd = scenes.cornell_c1()
    c = sd.camera
    for k in ("forward", "right", "up"):
        assert abs(np.linalg.norm(c[k]) - 1) < 1e-6
    assert abs(np.dot(c["forward"], c["right"])) < 1e-6 and abs(np.dot(c["forward"], c["up"])) < 1e-6
    assert np.allclose(c["forward"], [0, 0, -1]) and np.allclose(c["right"], [1, 0, 0])
    assert c["aspect"] == np.float32(1.0)


def test_geodesic_terrain_is_watertight_and_sized():
    v, f = scenes.geodesic_sphere(20)
    assert f.shape[0] == 20 * 20 * 20
    edges = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)   # closed 2-manifold: every edge shared by two faces


def test_c3_full_size_counts():
    v, f = scenes.uv_sphere(50, 21)
    assert f.shape[0] == 2000
    pv = v[f]
    area = np.linalg.norm(np.cross(pv[:, 1] - pv[:, 0], pv[:, 2] - pv[:, 0]), axis=1)
    assert area.min() > 0


def test_pod_layouts():
    assert pod.TRIANGLE.itemsize == 96 and pod.CAMERA.itemsize == 56 and pod.EMITTER.itemsize == 32
    assert pod.SPOTLIGHT.itemsize == 48 and pod.BSDF.itemsize == 68
    b = pod.bsdf_microfacet((0.2, 0.5, 0.3), alpha=0.2)
    assert b["p1"][3] == np.float32(1.0) - np.float32(0.5)
