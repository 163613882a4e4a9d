"""GPU: the tail pass (one persistent kernel that finishes every path from a
given bounce on) against the per-bounce passes and the oracle.

The per-bounce passes are held to the oracle bit for bit elsewhere; the tail
pass evaluates every path with the same functions in the same order, so every
output must be identical to theirs: per-sample radiance, the hit triangle of
every bounce, AOVs, running means, binary64 sums and the ray counters (per
bounce, hits, shadow rays).  Scenes cover every light class (area lights,
spotlights, sky), every BSDF type, textures, BVH2/BVH4/BVH8 and multi-scene
batches.
"""
import numpy as np
import pytest

import oracle
from bluetruthengine_amd import renderer, scenes
from test_gpu_parity import _compare, _many_spots, _sample_parity, _spot_scene, _textured_scene

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _run(sd, frames, max_rec, **cfg):
    r = renderer.Renderer(sd, **cfg)
    r.set_hit_capture(max_rec)
    r.enable_aov(True)
    r.enable_accum64(True)
    r.render(frames)
    out = {"rad": r.sample_radiance(frames), "hits": r.hit_capture(frames, max_rec), "accum": r.accum(),
           "sum64": r.accum64(), "aov": r.aov()}
    st = r.stats().as_dict()
    r.set_hit_capture(0)
    r.enable_aov(False)
    return out, st


def _same(a, b, label):
    for k in ("rad", "hits", "accum", "sum64"):
        assert np.array_equal(a[k], b[k]), f"{label}: {k} differs"
    for x, y in zip(a["aov"], b["aov"]):
        assert np.array_equal(x, y), f"{label}: AOV differs"


def _cases():
    return [
        (scenes.cornell_c1(width=64, height=64, spp=3, max_depth=8), {}),
        (scenes.small_envlit(), {}),
        (scenes.terrain_c2(width=96, height=54, spp=2, max_depth=10, freq=20), {}),
        (_spot_scene(), {}),
        (_many_spots(_textured_scene(), 6, 19), {}),
        (scenes.small_envlit(), {"bvh_width": 2}),
        (scenes.terrain_c2(width=96, height=54, spp=2, max_depth=10, freq=20), {"bvh_width": 8}),
    ]


def test_tail_pass_equals_per_bounce_passes():
    for sd, cfg in _cases():
        frames = 2
        base, st0 = _run(sd, frames, sd.max_depth, tail_bounce=0, **cfg)
        for tb in (1, 3, sd.max_depth - 1):
            _say(f"case: {sd.name} {sd.width}x{sd.height} depth {sd.max_depth} {cfg or ''}, tail from bounce {tb}")
            out, st = _run(sd, frames, sd.max_depth, tail_bounce=tb, **cfg)
            _same(out, base, f"{sd.name} tail {tb}")
            for k in ("extensionRays", "hits", "shadowRays", "pathsTraced", "raysPerBounce"):
                assert st[k] == st0[k], f"{sd.name} tail {tb}: {k} {st[k]} vs {st0[k]}"
        # the tail pass really replaced the per-bounce passes: with kernel timing on,
        # bounces 0..2 are per-bounce launches and 3.. one tail launch
        r = renderer.Renderer(sd, tail_bounce=3, **cfg)
        r.set_instrumentation(renderer.BTE_INSTR_KERNEL_TIMES)
        r.render(1)
        st = r.stats()
        assert (st.extendLaunches, st.tailLaunches) == (3, 1) and st.tailMs > 0.0
        r.set_instrumentation(0)
        _say(f"{sd.name}: identical; extension rays {st0['extensionRays']}, shadow rays {st0['shadowRays']}, "
             f"rays per bounce {st0['raysPerBounce']}")


def test_tail_pass_matches_oracle():
    _say("case: terrain freq 20 + Disney + 4 lights 96x54, 2 frames, depth 10, tail from bounce 2")
    sd = scenes.terrain_c2(width=96, height=54, spp=2, max_depth=10, freq=20)
    _, _, exact = _sample_parity(sd, 2, max_rec=10, tail_bounce=2)
    assert min(exact) == 1.0
    _say("case: reduced env-lit 128x72, 2 frames, depth 12, tail from bounce 1")
    _, _, exact = _sample_parity(scenes.small_envlit(), 2, max_rec=12, tail_bounce=1)
    assert min(exact) == 1.0


def test_tail_pass_with_graphs_across_passes():
    # captured segments ending in the tail pass replay across passes of 3 frames
    sd = scenes.cornell_c1(width=64, height=64, spp=7, max_depth=8)
    _say("case: Cornell 64x64, 7 frames in passes of 3, graphs on, tail from bounce 2, binary64 sums vs oracle")
    r = renderer.Renderer(sd, frames_in_flight=3, tail_bounce=2)
    r.enable_accum64(True)
    r.render(7)
    assert r.stats().graphLaunches > 0
    s64 = np.zeros((sd.height, sd.width, 3), np.float64)
    acc = oracle.OracleScene(sd).render(sd.width, sd.height, 0, 7, sd.max_depth, sum64=s64)
    _compare(r.accum(), acc, "accum, tail from bounce 2")
    assert np.array_equal(r.accum64(), s64) and np.array_equal(r.accum(), acc)


def test_batch_tail_pass_equals_per_bounce_passes():
    sds = [scenes.cornell_c1(width=48, height=40, spp=2, max_depth=8), scenes.small_envlit(),
           scenes.terrain_c2(width=64, height=36, spp=2, max_depth=10, freq=12), _spot_scene()]
    outs = {}
    for tb in (0, 2, 5):
        _say(f"case: batch of {len(sds)} scenes, 3 frames, depth 8, tail from bounce {tb}")
        b = renderer.Batch(max_depth=8, frames_in_flight=3, tail_bounce=tb)
        for sd in sds:
            b.add(sd)
        b.set_hit_capture(8)
        b.enable_accum64(True)
        b.render(3)
        outs[tb] = ([(b.sample_radiance(i, 3), b.hit_capture(i, 3, 8), b.accum(i), b.accum64(i))
                     for i in range(len(sds))], b.stats().as_dict())
        b.set_hit_capture(0)
    for tb in (2, 5):
        for i in range(len(sds)):
            for x, y in zip(outs[tb][0][i], outs[0][0][i]):
                assert np.array_equal(x, y), f"batch slot {i}, tail {tb}"
        for k in ("extensionRays", "hits", "shadowRays", "raysPerBounce"):
            assert outs[tb][1][k] == outs[0][1][k], f"batch tail {tb}: {k}"
    _say(f"batch: identical; extension rays {outs[0][1]['extensionRays']}")
