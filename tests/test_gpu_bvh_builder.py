"""The BVH built on the GPU (bte_bvh_gpu.hip, the default builder) against the
host's binned-SAH builder and the oracle.

Hits do not depend on the tree (a hit's distance is clamped to its triangle's
own box interval, so any BVH finds what a loop over all triangles finds): the
GPU-built BVH2, BVH4 and BVH8 must give the oracle's closest and any hits and
images bit-identical to the host-built trees'.  The rest of the GPU suite runs
on GPU-built trees too.
"""
import numpy as np
import pytest

import oracle
from bluetruthengine_amd import renderer, scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _rays(n, seed, lo, hi):
    rs = np.random.default_rng(seed)
    o = rs.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rs.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d.astype(np.float32)], 1)


@pytest.mark.parametrize("name", ["terrain_freq60", "envlit_small", "cornell"])
def test_gpu_built_bvh_hits_match_oracle(name):
    sd = {"terrain_freq60": lambda: scenes.terrain_c2(width=64, height=36, spp=1, max_depth=6, freq=60),
          "envlit_small": lambda: scenes.small_envlit(n_spheres=40),
          "cornell": lambda: scenes.cornell_c1(width=32, height=32, spp=1, max_depth=6)}[name]()
    _say(f"case: GPU-built BVH, {name} ({sd.triangles.shape[0]} tris), 300k closest + 100k any-hit rays, seed 51")
    renderer.set_accel_builder(1)
    r = renderer.Renderer(sd)
    r.build_accel()
    st = r.accel_stats()
    _say(f"accel: {st}")
    assert st["builder"] == 1 and st["triangles"] == sd.triangles.shape[0]
    assert 1 <= st["depth4"] <= st["depth2"] <= 60 and st["nodes4"] <= st["nodes2"] < sd.triangles.shape[0]
    lo, hi = sd.triangles["v0"].min(0), sd.triangles["v0"].max(0)
    rays = _rays(300_000, 51, lo, hi)
    tri, t, u, v = renderer.trace_rays(rays)
    ref = oracle.OracleScene(sd)
    rtri, rt, ru, rv = ref.trace_closest(rays)
    mism = np.where(tri != rtri)[0]
    _say(f"closest: {mism.size} mismatches, hits {(rtri >= 0).mean():.3f}")
    assert mism.size == 0
    assert np.array_equal(t, rt) and np.array_equal(u, ru) and np.array_equal(v, rv)
    rays2 = rays[:100_000]
    tmax = np.random.default_rng(52).uniform(0.01, 2.0, rays2.shape[0]).astype(np.float32)
    occ = renderer.trace_rays(rays2, any_hit=True, tmax_per_ray=tmax)
    assert np.array_equal(occ.astype(np.uint8), ref.trace_occluded(rays2, tmax))


@pytest.mark.parametrize("width", [2, 4, 8])
def test_gpu_and_host_trees_render_identically(width):
    _say(f"case: terrain freq 40 + env-lit small, 3 frames, BVH{width} from the GPU builder vs the host builder")
    for sd in (scenes.terrain_c2(width=96, height=54, spp=3, max_depth=10, freq=40), scenes.small_envlit()):
        out = []
        for builder in (0, 1):
            renderer.set_accel_builder(builder)
            r = renderer.Renderer(sd, bvh_width=width)
            r.enable_accum64(True)
            r.render(3)
            out.append((r.accum().copy(), r.accum64().copy(), r.accel_stats()))
        renderer.set_accel_builder(1)
        _say(f"{sd.name}: host {out[0][2]['nodes2']} nodes depth {out[0][2]['depth2']}, "
             f"GPU {out[1][2]['nodes2']} nodes depth {out[1][2]['depth2']}")
        assert out[0][2]["builder"] == 0 and out[1][2]["builder"] == 1
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_gpu_builder_on_the_benchmark_scenes():
    # build time and tree size at 1M triangles and over the 64 serving scenes; hits
    # on the 1080p frame are covered by tests/test_gpu_fullsize.py (default builder)
    sd = scenes.envlit_c3(spp=1)
    _say(f"case: GPU BVH build of the env-lit frame ({sd.triangles.shape[0]} tris) and of 16 serving scenes")
    renderer.set_accel_builder(1)
    r = renderer.Renderer(sd)
    r.build_accel()
    r.build_accel()
    g = r.accel_stats()
    renderer.set_accel_builder(0)
    r = renderer.Renderer(sd)
    r.build_accel()
    h = r.accel_stats()
    renderer.set_accel_builder(1)
    _say(f"1M tris: GPU {g['buildMs']:.1f} ms ({g['nodes2']} BVH2 / {g['nodes4']} BVH4 nodes, depth {g['depth2']}), "
         f"host {h['buildMs']:.1f} ms ({h['nodes2']} / {h['nodes4']}, depth {h['depth2']})")
    assert g["builder"] == 1 and g["depth2"] <= 60
    rays = _rays(200_000, 53, (-4, 0.01, -4), (4, 1.0, 4))
    a = renderer.trace_rays(rays)
    renderer.set_accel_builder(0)
    renderer.Renderer(sd).build_accel()
    b = renderer.trace_rays(rays)
    renderer.set_accel_builder(1)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
