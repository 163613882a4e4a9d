"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bar (north_star): hit triangle IDs identical; per-sample radiance within 1e-4
relative in binary32 (the kernels and the oracle perform the same binary32
operation sequence, so the observed difference is expected to be zero and the
bit-exact fraction is asserted as well).
"""
import numpy as np
import pytest

import oracle
from bluetruthengine_amd import renderer, scenes

pytestmark = pytest.mark.gpu

REL_TOL = 1e-4
ABS_FLOOR = 1e-6


def _say(msg):
    print(msg, flush=True)


def _compare(gpu, cpu, what):
    gpu = np.asarray(gpu, np.float32).reshape(-1, 3)
    cpu = np.asarray(cpu, np.float32).reshape(-1, 3)
    both_nan = np.isnan(gpu) & np.isnan(cpu)
    diff = np.where(both_nan, 0.0, np.abs(gpu.astype(np.float64) - cpu.astype(np.float64)))
    rel = diff / np.maximum(np.abs(cpu.astype(np.float64)), ABS_FLOOR)
    exact = np.all((gpu == cpu) | both_nan, axis=1)
    _say(f"{what}: {gpu.shape[0]} samples, bit-exact {exact.mean():.6f}, max rel err {rel.max():.3g}")
    bad = np.where(rel.max(axis=1) > REL_TOL)[0]
    assert bad.size == 0, f"{what}: {bad.size} samples exceed {REL_TOL} (first {bad[:5]}: gpu {gpu[bad[:3]]} cpu {cpu[bad[:3]]})"
    return exact.mean()


def _random_rays(n, seed, lo=(-0.95, 0.05, -0.95), hi=(0.95, 1.95, 0.95)):
    rs = np.random.default_rng(seed)
    o = rs.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rs.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d.astype(np.float32)], 1)


@pytest.fixture(scope="module")
def c1():
    return scenes.cornell_c1(width=96, height=96, spp=3, max_depth=8)


def test_closest_hit_matches_oracle(c1):
    _say("case: closest-hit rays, Cornell C1 geometry, 200k rays seed 11")
    renderer.Renderer(c1)
    rays = _random_rays(200_000, 11)
    tri, t, u, v = renderer.trace_rays(rays)
    ref = oracle.OracleScene(c1)
    rtri, rt, ru, rv = ref.trace_closest(rays)
    btri, bt, _, _ = ref.trace_closest(rays[:20000], brute=True)
    assert np.array_equal(rtri[:20000], btri) and np.array_equal(rt[:20000], bt)
    mism = np.where(tri != rtri)[0]
    _say(f"hit-id mismatches {mism.size} / {rays.shape[0]}; misses {(tri < 0).mean():.3f}")
    assert mism.size == 0, f"first mismatches {mism[:10]}: gpu {tri[mism[:5]]} cpu {rtri[mism[:5]]}"
    hit = tri >= 0
    assert np.array_equal(t[hit], rt[hit]) and np.array_equal(u[hit], ru[hit]) and np.array_equal(v[hit], rv[hit])


def test_any_hit_matches_oracle(c1):
    _say("case: any-hit rays, Cornell C1 geometry, 100k rays seed 12")
    renderer.Renderer(c1)
    rays = _random_rays(100_000, 12)
    tmax = np.random.default_rng(13).uniform(0.05, 3.0, rays.shape[0]).astype(np.float32)
    occ = renderer.trace_rays(rays, any_hit=True, tmax_per_ray=tmax)
    ref = oracle.OracleScene(c1).trace_occluded(rays, tmax)
    assert np.array_equal(occ.astype(np.uint8), ref)


def _sample_parity(sd, frames, max_rec=4, env_from_gpu=True, **cfg):
    r = renderer.Renderer(sd, **cfg)
    r.set_hit_capture(max_rec)
    r.render(frames)
    gpu = r.sample_radiance(frames)
    hits = r.hit_capture(frames, max_rec)
    r.set_hit_capture(0)
    env_cdf = r.env_cdf() if (env_from_gpu and sd.env_rgba is not None) else None
    ref = oracle.OracleScene(sd, env_cdf=env_cdf)
    W, H = sd.width, sd.height
    pixels = np.arange(W * H, dtype=np.int32)
    exact = []
    for f in range(frames):
        cpu, tris = ref.trace_samples(pixels, W, H, f, sd.max_depth, max_rec=max_rec)
        mism = np.where(np.any(hits[f] != tris, axis=1))[0]
        _say(f"{sd.name} frame {f}: hit-id mismatching paths {mism.size} / {pixels.size}")
        assert mism.size == 0, f"paths {mism[:5]}: gpu {hits[f][mism[:3]]} cpu {tris[mism[:3]]}"
        exact.append(_compare(gpu[f], cpu, f"{sd.name} frame {f}"))
    return r, ref, exact


def test_cornell_samples_match_oracle(c1):
    _say("case: Cornell C1 96x96, 3 frames, depth 8")
    _, _, exact = _sample_parity(c1, 3)
    assert min(exact) == 1.0


def test_envlit_samples_match_oracle():
    _say("case: reduced env-lit scene 128x72, 2 frames, depth 12")
    sd = scenes.small_envlit()
    _, _, exact = _sample_parity(sd, 2)
    assert min(exact) == 1.0


def test_bvh2_traversal_samples_match_oracle():
    # the other cases run the default BVH4; this one the BVH2 it is collapsed from
    _say("case: BVH2 extend/connect, terrain freq 20 + reduced env-lit, 2 frames each")
    for sd in (scenes.terrain_c2(width=96, height=54, spp=2, max_depth=10, freq=20), scenes.small_envlit()):
        _, _, exact = _sample_parity(sd, 2, bvh_width=2)
        assert min(exact) == 1.0


def test_bvh8_traversal_samples_match_oracle():
    # 8-wide nodes collapsed from the same BVH2: closest and any hit through the 8-child sorting network
    _say("case: BVH8 extend/connect, terrain freq 20 + reduced env-lit, 2 frames each")
    for sd in (scenes.terrain_c2(width=96, height=54, spp=2, max_depth=10, freq=20), scenes.small_envlit()):
        _, _, exact = _sample_parity(sd, 2, bvh_width=8)
        assert min(exact) == 1.0


def test_terrain_disney_samples_match_oracle():
    _say("case: terrain freq 20 (8000 tris) + 64 Disney slots + 4 lights, 96x54, 2 frames, depth 10")
    sd = scenes.terrain_c2(width=96, height=54, spp=2, max_depth=10, freq=20)
    _, _, exact = _sample_parity(sd, 2)
    assert min(exact) == 1.0


def _spot_scene():
    """Cornell room + ceiling light + two spotlights (cone falloff) + a microfacet
    sphere: exercises spotlight NEE (rt_direct_lighting.cuh:175-230) and three
    shadow-ray classes in one path (area, spot x2)."""
    from bluetruthengine_amd import pod
    sd = scenes.cornell_c1(width=80, height=80, spp=2, max_depth=8)
    sp = np.zeros(2, pod.SPOTLIGHT)
    sp[0]["position"] = (-0.5, 1.8, 0.5)
    sp[0]["direction"] = np.array([0.3, -1.0, -0.3]) / np.linalg.norm([0.3, -1.0, -0.3])
    sp[0]["radiance"] = (1.0, 0.8, 0.6)
    sp[0]["intensity"] = 4.0
    sp[0]["innerConeCosine"] = np.cos(np.radians(15.0))
    sp[0]["outerConeCosine"] = np.cos(np.radians(35.0))
    sp[1]["position"] = (0.6, 1.5, 0.8)
    sp[1]["direction"] = (0.0, -1.0, 0.0)
    sp[1]["radiance"] = (0.5, 0.7, 1.0)
    sp[1]["intensity"] = 2.0
    sp[1]["innerConeCosine"] = np.cos(np.radians(5.0))
    sp[1]["outerConeCosine"] = np.cos(np.radians(60.0))
    sd.spotlights = sp
    sd.bsdfs[0] = pod.bsdf_microfacet((0.5, 0.4, 0.3), alpha=0.2)   # walls: microfacet
    sd.name = "cornell_spot"
    return sd


def _textured_scene():
    """Cornell room with texture maps (rt_material.cuh:26-56): albedo maps on a
    diffuse wall pair and a microfacet wall (ks recomputed from the map), albedo
    and metallic-roughness maps on a Disney wall, texture coordinates well
    outside [0, 1] (wrap addressing), untextured materials in between."""
    from bluetruthengine_amd import pod
    sd = scenes.cornell_c1(width=80, height=80, spp=2, max_depth=8)
    tri = sd.triangles.copy()
    pos = np.stack([tri["v0"], tri["v1"], tri["v2"]], 1)
    uv = np.stack([pos[..., 0] * 1.7 + pos[..., 2] * 0.3, pos[..., 1] * 1.3 - pos[..., 2] * 0.9], -1)
    tri["uv0"], tri["uv1"], tri["uv2"] = uv[:, 0], uv[:, 1], uv[:, 2]
    sd.triangles = tri
    mat = np.full(tri.shape[0], 3, np.int32)     # the room's 5 quads are the first 10 triangles
    mat[0:2], mat[2:4], mat[4:6], mat[6:8] = 0, 3, 1, 2
    sd.tri_material_ids = mat
    n0 = sd.bsdfs.shape[0]
    extra = np.array([pod.bsdf_microfacet((0.6, 0.5, 0.4), alpha=0.3),
                      pod.bsdf_disney(base_color=(0.8, 0.7, 0.6), roughness=0.5, metallic=0.8)], pod.BSDF)
    sd.bsdfs = np.concatenate([sd.bsdfs, extra])
    ids = sd.tri_bsdf_ids.copy()
    ids[6:8], ids[4:6] = n0, n0 + 1
    sd.tri_bsdf_ids = ids
    rs = np.random.default_rng(8)
    checker = np.zeros((16, 16, 4), np.uint8)
    checker[(np.add.outer(np.arange(16), np.arange(16)) // 4) % 2 == 0] = (230, 60, 40, 255)
    checker[(np.add.outer(np.arange(16), np.arange(16)) // 4) % 2 == 1] = (40, 90, 220, 255)
    sd.textures = [checker, rs.integers(0, 256, (5, 9, 4), dtype=np.uint8),
                   rs.integers(0, 256, (32, 8, 4), dtype=np.uint8), None]
    sd.mr_textures = [None, rs.integers(0, 256, (6, 6, 4), dtype=np.uint8), None, None]
    sd.name = "cornell_textured"
    return sd


def test_textured_samples_match_oracle():
    _say("case: Cornell with albedo / metallic-roughness texture maps 80x80, 2 frames, depth 8")
    sd = _textured_scene()
    _, _, exact = _sample_parity(sd, 2)
    assert min(exact) == 1.0


def test_textures_change_the_image():
    # the same scene without maps renders differently where the maps are seen
    sd = _textured_scene()
    r = renderer.Renderer(sd)
    r.render(2)
    with_maps = r.accum().copy()
    sd.textures = sd.mr_textures = None
    r = renderer.Renderer(sd)
    r.render(2)
    diff = np.abs(with_maps - r.accum()).max(axis=-1)
    _say(f"pixels changed by the maps: {(diff > 1e-3).mean():.3f}")
    assert (diff > 1e-3).mean() > 0.2


def test_spotlight_samples_match_oracle():
    _say("case: Cornell + 2 spotlights + microfacet walls 80x80, 2 frames, depth 8")
    _, _, exact = _sample_parity(_spot_scene(), 2)
    assert min(exact) == 1.0


def test_env_cdf_binary64_matches_oracle():
    _say("case: env CDF 2048x1024 sky, GPU binary64 scan vs oracle sequential binary64")
    sd = scenes.small_envlit(env_size=(2048, 1024))
    r = renderer.Renderer(sd)
    marg, cond, total = r.env_cdf()
    om, oc, ot = oracle.env_build_cdf(sd.env_rgba, 1)
    for a, b, name in ((marg, om, "marginal"), (cond, oc, "conditional")):
        ulps = np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))
        _say(f"{name}: max ulp diff {ulps.max()}, differing entries {(ulps > 0).sum()} / {a.size}")
        assert ulps.max() <= 1
    assert abs(float(total) - float(ot)) <= 1e-6 * float(ot)


def test_accumulation_and_display_match_oracle():
    _say("case: Cornell 64x64, 4 frames: Welford mean and gamma-packed display")
    sd = scenes.cornell_c1(width=64, height=64, spp=4, max_depth=8)
    r = renderer.Renderer(sd)
    r.render(4)
    acc = r.accum()
    fb = r.framebuffer()
    ref = oracle.OracleScene(sd)
    racc, rfb = ref.render(64, 64, 0, 4, sd.max_depth, rgba8=True)
    _compare(acc, racc, "accum")
    assert np.array_equal(acc, racc)
    assert np.array_equal(fb, rfb)
    assert r.frame_index() == 4


def test_frames_in_flight_do_not_change_result():
    _say("case: Cornell 48x48, 4 frames with 1, 3 and 4 frames in flight")
    sd = scenes.cornell_c1(width=48, height=48, spp=4, max_depth=8)
    outs = []
    for fif in (1, 3, 4):
        r = renderer.Renderer(sd, frames_in_flight=fif)
        r.render(4)
        outs.append(r.accum())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_megakernel_one_spp_equals_wavefront():
    _say("case: Cornell 48x48 megakernel (1 path/frame) vs wavefront, 2 frames")
    sd = scenes.cornell_c1(width=48, height=48, spp=2, max_depth=8)
    w = renderer.Renderer(sd)
    w.render(2)
    a = w.accum()
    m = renderer.Renderer(sd, mode=renderer.BTE_MODE_MEGAKERNEL, megakernel_spp=1)
    m.render(2)
    b = m.accum()
    _compare(b, a, "megakernel vs wavefront")
    assert np.array_equal(a, b)


def test_megakernel_five_paths_per_frame_matches_oracle():
    _say("case: Cornell 48x48 megakernel mode, 5 paths/pixel/frame sharing one RNG stream, 2 frames")
    sd = scenes.cornell_c1(width=48, height=48, spp=2, max_depth=8)
    m = renderer.Renderer(sd, mode=renderer.BTE_MODE_MEGAKERNEL, megakernel_spp=5)
    m.render(2)
    ref = oracle.OracleScene(sd).render_megakernel(48, 48, 0, 2, 5, sd.max_depth)
    _compare(m.accum(), ref, "megakernel spp5")
    assert np.array_equal(m.accum(), ref)


def test_error_paths_are_reported():
    _say("case: C-ABI error reporting")
    renderer.cudaCleanup()
    renderer._call("bte_denoiser_release")
    with pytest.raises(renderer.BteError):
        renderer.cudaRender(16, 16)
    with pytest.raises(renderer.BteError):
        renderer.cudaToggleDenoiser()


def test_aov_means_match_oracle():
    for sd, what in ((scenes.small_envlit(), "reduced env-lit 128x72"),
                     (scenes.cornell_c1(width=64, height=64, spp=3, max_depth=8), "Cornell 64x64")):
        _say(f"case: albedo/normal AOVs, {what}, 3 frames")
        r = renderer.Renderer(sd)
        r.enable_aov(True)
        r.render(3)
        alb, nrm = r.aov()
        env_cdf = r.env_cdf() if sd.env_rgba is not None else None
        ralb, rnrm = oracle.OracleScene(sd, env_cdf=env_cdf).render_aov(sd.width, sd.height, 0, 3, sd.max_depth)
        assert np.array_equal(alb, ralb) and np.array_equal(nrm, rnrm)
        plain = renderer.Renderer(sd)
        plain.render(3)
        assert np.array_equal(r.accum(), plain.accum())     # AOVs leave the radiance untouched


def test_accum64_matches_oracle_across_passes():
    # the binary64 sums behind every converged reference image: 7 frames in
    # passes of 3 (3 + 3 + 1) with graphs on, against the oracle's per-frame
    # sums (rt_path_integrator.cuh:72-87 plus the plain binary64 sum)
    for sd in (scenes.cornell_c1(width=64, height=64, spp=7, max_depth=8), scenes.small_envlit()):
        _say(f"case: {sd.name} {sd.width}x{sd.height}, 7 frames in passes of 3, graphs on: binary64 sums")
        r = renderer.Renderer(sd, frames_in_flight=3)
        r.enable_accum64(True)
        r.render(7)
        assert r.stats().graphLaunches > 0
        ref = oracle.OracleScene(sd, env_cdf=r.env_cdf() if sd.env_rgba is not None else None)
        s64 = np.zeros((sd.height, sd.width, 3), np.float64)
        acc = ref.render(sd.width, sd.height, 0, 7, sd.max_depth, sum64=s64)
        same = np.array_equal(r.accum64(), s64), np.array_equal(r.accum(), acc)
        _say(f"{sd.name}: binary64 sums equal {same[0]}, running means equal {same[1]}")
        assert all(same)


def _many_spots(sd, n, seed):
    """n spotlights near the ceiling aimed down at random (cone angles from
    narrow to wide), in addition to the scene's other lights."""
    from bluetruthengine_amd import pod
    rs = np.random.default_rng(seed)
    sp = np.zeros(n, pod.SPOTLIGHT)
    for i in range(n):
        d = rs.normal(size=3) * 0.5 + np.array([0.0, -1.0, 0.0])
        sp[i]["position"] = rs.uniform((-0.9, 1.6, -0.9), (0.9, 1.95, 0.9))
        sp[i]["direction"] = d / np.linalg.norm(d)
        sp[i]["radiance"] = rs.uniform(0.3, 1.0, 3)
        sp[i]["intensity"] = rs.uniform(0.5, 3.0)
        inner, outer = np.sort(rs.uniform(5, 80, 2))
        sp[i]["innerConeCosine"] = np.cos(np.radians(inner))
        sp[i]["outerConeCosine"] = np.cos(np.radians(outer))
    sd.spotlights = sp
    sd.name += f"_{n}spots"
    return sd


def test_twenty_spotlights_match_oracle():
    # any number of spotlights (cudaInitSpotlights takes any count): one spot
    # pass traces each path's spotlights in index order; textured diffuse,
    # microfacet and Disney walls, BVH4 and BVH2
    _say("case: textured Cornell + ceiling light + 20 spotlights 64x64, 2 frames, depth 8, BVH4 and BVH2")
    for width in (4, 2):
        sd = _many_spots(_textured_scene(), 20, 17)
        sd.width = sd.height = 64
        _, _, exact = _sample_parity(sd, 2, bvh_width=width)
        assert min(exact) == 1.0
    r = renderer.Renderer(_many_spots(scenes.cornell_c1(width=32, height=32, spp=1, max_depth=8), 20, 18))
    r.render(1)
    st = r.stats()
    _say(f"shadow rays {st.shadowRays} for {st.pathsTraced} paths")
    assert st.shadowRays > 5 * st.pathsTraced      # counts every traced spotlight ray, not one per request
