"""Multi-scene batches (bte_batch_*): scenes of different sizes, resolutions,
BSDF types and light classes traced together in one persistent wavefront pass.

Every slot must accumulate exactly what a Renderer of that scene alone
accumulates: the binary32 running means, the binary64 sums, the display
buffer and the denoised image are compared bit for bit.  The single-scene
path is pinned to the CPU oracle by the parity tests; the first case also
checks one slot's per-pixel means against the oracle directly.
"""
import numpy as np
import pytest

import oracle
from bluetruthengine_amd import denoiser, renderer, scenes
from test_gpu_parity import _spot_scene

pytestmark = pytest.mark.gpu

DEPTH = 8


def _say(msg):
    print(msg, flush=True)


def _scene_set():
    """Four scenes covering every light class and BSDF type: area light only
    (Cornell, mirror + dielectric), area + two spotlights (microfacet walls),
    four area lights over a Disney terrain, and an env map over Disney metal
    and dielectric spheres (no area light)."""
    return [scenes.cornell_c1(width=64, height=48, spp=1, max_depth=DEPTH),
            _spot_scene(),
            scenes.terrain_c2(width=96, height=54, spp=1, max_depth=DEPTH, freq=20),
            scenes.small_envlit(width=72, height=40, n_spheres=10, env_size=(128, 64))]


def _alone(sd, frames, aov=False, **cfg):
    r = renderer.Renderer(sd, max_depth=DEPTH, **cfg)
    r.enable_accum64(True)
    if aov:
        r.enable_aov(True)
    r.render(frames)
    return r.accum().copy(), r.accum64().copy(), r.framebuffer().copy()


def _batched(sds, frames, calls=1, aov=False, **cfg):
    b = renderer.Batch(max_depth=DEPTH, aov=aov, **cfg)
    for sd in sds:
        b.add(sd)
    b.enable_accum64(True)
    for _ in range(calls):
        b.render(frames)
    return b


def _slots(b, n):
    """Every slot's outputs, read before anything else touches the library
    (a Renderer starts from bte_cleanup, which frees the batch)."""
    return [(b.accum(i).copy(), b.accum64(i).copy(), b.framebuffer(i).copy()) for i in range(n)]


def _check(got, sds, frames, label, **cfg):
    for i, sd in enumerate(sds):
        ref = _alone(sd, frames, **cfg)
        same = tuple(np.array_equal(x, y) for x, y in zip(got[i], ref))
        _say(f"{label}: slot {i} {sd.name} {sd.width}x{sd.height}: accum / accum64 / framebuffer equal {same}")
        assert all(same), f"{label}: slot {i} ({sd.name}) differs from the scene rendered alone"


def test_batch_slots_equal_scenes_rendered_alone():
    _say("case: 4-scene batch (area, area+spots, 4 area lights, env), 3 frames, depth 8, graphs on")
    sds = _scene_set()
    b = _batched(sds, 3)
    assert len(b) == 4 and b.frame_index() == 3
    st = b.stats()
    _say(f"batch stats: {st.pathsTraced} paths, {st.extensionRays} extension rays, {st.shadowRays} shadow rays")
    assert st.pathsTraced == 3 * sum(sd.width * sd.height for sd in sds)
    got = _slots(b, len(sds))
    # one slot against the CPU oracle's per-pixel means
    sd = sds[1]
    ref = oracle.OracleScene(sd).render(sd.width, sd.height, 0, 3, DEPTH)
    rel = np.abs(got[1][0].astype(np.float64) - ref) / np.maximum(np.abs(ref.astype(np.float64)), 1e-6)
    _say(f"slot 1 vs oracle: max rel {rel.max():.3g}, bit-exact {(got[1][0] == ref).mean():.4f}")
    assert rel.max() <= 1e-4
    _check(got, sds, 3, "batch")


def test_batch_uneven_passes_and_plain_launches():
    # frames in flight 2 over 5 frames: passes of 2, 2 and 1 frames (three
    # path-id layouts), accumulated over two calls; graphs off, then on
    sds = _scene_set()
    for graphs in (0, 1):
        _say(f"case: 4-scene batch, framesInFlight 2, 2 calls x 5 frames, graphs {graphs}")
        b = _batched(sds, 5, calls=2, frames_in_flight=2, graphs=graphs)
        assert b.frame_index() == 10
        _check(_slots(b, len(sds)), sds, 10, f"uneven passes graphs={graphs}")


def test_batch_bvh8_equals_bvh4_alone():
    # batch passes over 8-wide nodes against each scene rendered alone over the default BVH4
    _say("case: 4-scene batch, BVH8 extend/connect, 2 frames, vs scenes alone on BVH4")
    sds = _scene_set()
    b = _batched(sds, 2, bvh_width=8)
    _check(_slots(b, len(sds)), sds, 2, "batch bvh8")


def test_batch_reset_and_clear():
    _say("case: batch reset restarts accumulation; clear empties the batch")
    sds = _scene_set()[:2]
    b = _batched(sds, 2)
    b.reset()
    assert b.frame_index() == 0
    b.render(2)
    got = _slots(b, len(sds))
    with pytest.raises(renderer.BteError):
        renderer._call("bte_batch_read_accum", 5, None, 0)
    b.clear()
    assert renderer.lib().bte_batch_scene_count() == 0
    with pytest.raises(renderer.BteError):
        b.render(1)
    _check(got, sds, 2, "after reset")


def test_batch_denoise_equals_single_scene_denoise():
    _say("case: 3-scene batch with AOVs, 4 frames, batched U-Net vs per-scene bte_denoise")
    sds = [scenes.small_envlit(width=72, height=40, n_spheres=10, env_size=(128, 64)),
           scenes.cornell_c1(width=64, height=48, spp=1, max_depth=DEPTH),
           scenes.terrain_c2(width=96, height=54, spp=1, max_depth=DEPTH, freq=20)]
    b = _batched(sds, 4, aov=True)
    denoiser.load(denoiser.WEIGHTS_PATH)
    b.denoise()
    got = [(b.denoised(i).copy(), b.framebuffer(i).copy()) for i in range(len(sds))]
    for i, sd in enumerate(sds):
        r = renderer.Renderer(sd, max_depth=DEPTH)
        r.enable_aov(True)
        r.render(4)
        denoiser.denoise()
        same = np.array_equal(got[i][0], denoiser.denoised(sd.width, sd.height))
        same_fb = np.array_equal(got[i][1], r.framebuffer())
        _say(f"slot {i} {sd.name}: denoised equal {same}, display buffer equal {same_fb}")
        assert same and same_fb


def test_batch_autotune_keeps_images():
    _say("case: batch autotune over 2-frame renders of 2 scenes, then a render against the scenes alone")
    sds = _scene_set()[2:]
    b = _batched(sds, 1)
    ms, cfg = b.autotune(2)
    _say(f"batch autotune: {ms:.2f} ms with {cfg}")
    assert ms > 0 and cfg["bvhWidth"] in (2, 4, 8) and cfg["traversalStack"] in (8, 16, 32)
    assert b.frame_index() == 0
    b.render(3)
    _check(_slots(b, len(sds)), sds, 3, "after autotune")
