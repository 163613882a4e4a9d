"""GPU parity at the full sizes of the benchmark workloads.

Three workloads run on the HIP path at the sizes the benchmark and its
configurations name, with hit capture on:
  * the 1080p env-lit frame: 1920x1080, depth 12, 1,000,002 triangles,
    2048x1024 sky, 16 frames in one wavefront pass;
  * the 720p terrain room: 1280x720, depth 10, 200,018 triangles, 64 Disney
    slots, 4 area lights, 64 frames in one pass;
  * the serving batch: 64 scenes of 256^2..1024^2 pixels and 10k..500k
    triangles (7.5 M in all), 8 frames, depth 10, in passes of several frames.
A seeded subset of at least 100,000 pixel-samples at the first and the last
frame of a pass (every frame of the 1080p pass) is traced by the oracle and
compared bit for bit: the hit triangle of every bounce and the sample's radiance (bar: 1e-4
relative; the observed difference is zero and bit-exactness is asserted).  At
those pixels the binary32 running mean (Welford, rt_path_integrator.cuh:72-87)
and the binary64 sums over every frame are recomputed from the oracle's
samples and compared bit for bit as well.
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


def _welford(samples):
    """Running mean of frames 0..F-1 in binary32, the Welford count of frame f
    being f + 1 (rt_path_integrator.cuh:72-87), and the binary64 plain sum in
    frame order."""
    mean = np.zeros(samples.shape[1:], np.float32)
    s64 = np.zeros(samples.shape[1:], np.float64)
    for f in range(samples.shape[0]):
        mean = mean + (samples[f] - mean) / np.float32(f + 1)
        s64 = s64 + samples[f].astype(np.float64)
    return mean, s64


def _check(label, gpu_rad, gpu_hits, cpu_rad, cpu_hits, pix=None, strict=True):
    """Hit ids and radiance of the same samples; returns the failing sample
    indices (asserts none when strict)."""
    gpu_rad = np.asarray(gpu_rad, np.float32).reshape(-1, 3)
    cpu_rad = np.asarray(cpu_rad, np.float32).reshape(-1, 3)
    mism = np.where(np.any(gpu_hits != cpu_hits, axis=1))[0]
    _say(f"{label}: {cpu_hits.shape[0]} samples, hit-id mismatching paths {mism.size}, "
         f"misses at bounce 0 {(cpu_hits[:, 0] == -1).mean():.3f}, paths past bounce 3 "
         f"{(cpu_hits[:, 3] != -2).mean():.3f}, past bounce {cpu_hits.shape[1] - 1} {(cpu_hits[:, -1] != -2).mean():.4f}")
    diff = np.abs(gpu_rad.astype(np.float64) - cpu_rad.astype(np.float64))
    rel = diff / np.maximum(np.abs(cpu_rad.astype(np.float64)), ABS_FLOOR)
    exact = np.all(gpu_rad == cpu_rad, axis=1)
    _say(f"{label}: radiance bit-exact {exact.mean():.6f}, max rel err {rel.max():.3g}")
    bad = np.union1d(mism, np.where(~exact)[0])
    for i in bad[:8]:
        _say(f"{label}: sample {i}" + (f" (pixel {pix[i]})" if pix is not None else "") +
             f": gpu {gpu_rad[i]} cpu {cpu_rad[i]}, gpu hits {gpu_hits[i]}, cpu hits {cpu_hits[i]}")
    if strict:
        assert mism.size == 0, f"{label}: {mism.size} paths with different hit ids"
        assert np.all(rel <= REL_TOL), f"{label}: {(rel.max(axis=1) > REL_TOL).sum()} samples exceed {REL_TOL}"
        assert exact.all()
    return bad


def _check_means(label, mean, s64, samples):
    want_mean, want_s64 = _welford(samples)
    same_mean = np.array_equal(mean, want_mean)
    same_s64 = np.array_equal(s64, want_s64)
    _say(f"{label}: running mean equal {same_mean}, binary64 sums equal {same_s64} "
         f"({samples.shape[1]} pixels x {samples.shape[0]} frames)")
    if not (same_mean and same_s64):
        bad = np.where(np.any(mean != want_mean, axis=1) | np.any(s64 != want_s64, axis=1))[0]
        _say(f"{label}: {bad.size} pixels differ; first {bad[:5]}: gpu mean {mean[bad[:3]]} want {want_mean[bad[:3]]}; "
             f"gpu sum {s64[bad[:3]]} want {want_s64[bad[:3]]}; nan samples {np.isnan(samples).any(axis=(0, 2)).sum()}")
    assert same_mean and same_s64


def test_envlit_1080p_fullsize_matches_oracle():
    _say("case: env-lit 1920x1080, 1,000,002 tris, 2048x1024 sky, depth 12, 16 frames in one pass, seed 31")
    sd = scenes.envlit_c3(spp=16)
    W, H = sd.width, sd.height
    r = renderer.Renderer(sd)
    r.enable_accum64(True)
    r.set_hit_capture(sd.max_depth)
    r.render(16)
    st = r.stats()
    _say(f"frames in flight {st.framesInFlight}, extension rays {st.extensionRays}, shadow rays {st.shadowRays}")
    assert st.framesInFlight == 16
    gpu = r.sample_radiance(16).reshape(16, W * H, 3)
    hits = r.hit_capture(16, sd.max_depth)
    mean, s64 = r.accum().reshape(-1, 3), r.accum64().reshape(-1, 3)
    r.set_hit_capture(0)
    ref = oracle.OracleScene(sd, env_cdf=r.env_cdf())
    pix = np.sort(np.random.default_rng(31).choice(W * H, 100_000, replace=False)).astype(np.int32)
    samples = np.empty((16, pix.size, 3), np.float32)
    failing = 0
    for f in range(16):
        cpu, tris = ref.trace_samples(pix, W, H, f, sd.max_depth, max_rec=sd.max_depth)
        failing += _check(f"env-lit 1080p frame {f}", gpu[f][pix], hits[f][pix], cpu, tris, pix, strict=False).size
        samples[f] = cpu
    assert failing == 0, f"{failing} of {16 * pix.size} samples differ"
    _check_means("env-lit 1080p", mean[pix], s64[pix], samples)


def test_terrain_720p_fullsize_matches_oracle():
    _say("case: terrain room 1280x720, 200,018 tris, 64 Disney slots, 4 area lights, depth 10, 64 frames in one "
         "pass, seed 32")
    sd = scenes.terrain_c2()
    W, H = sd.width, sd.height
    r = renderer.Renderer(sd)
    r.enable_accum64(True)
    r.set_hit_capture(sd.max_depth)
    r.render(64)
    st = r.stats()
    _say(f"frames in flight {st.framesInFlight}, extension rays {st.extensionRays}, shadow rays {st.shadowRays}")
    assert st.framesInFlight == 64
    gpu = r.sample_radiance(64).reshape(64, W * H, 3)
    hits = r.hit_capture(64, sd.max_depth)
    mean, s64 = r.accum().reshape(-1, 3), r.accum64().reshape(-1, 3)
    r.set_hit_capture(0)
    ref = oracle.OracleScene(sd)
    rs = np.random.default_rng(32)
    pix = np.sort(rs.choice(W * H, 100_000, replace=False)).astype(np.int32)
    for f in (0, 63):
        cpu, tris = ref.trace_samples(pix, W, H, f, sd.max_depth, max_rec=sd.max_depth)
        _check(f"terrain 720p frame {f}", gpu[f][pix], hits[f][pix], cpu, tris)
    # means over all 64 frames on a smaller subset
    sub = pix[:: 5]
    samples = np.stack([ref.trace_samples(sub, W, H, f, sd.max_depth)[0] for f in range(64)])
    _check_means("terrain 720p", mean[sub], s64[sub], samples)


def test_serving_batch_fullsize_matches_oracle():
    _say("case: serving batch, 64 scenes 256^2..1024^2, 10k..500k tris, 8 frames, depth 10, default passes, seed 33")
    sds = scenes.serve_c5(64)
    _say(f"generated: {sum(sd.triangles.shape[0] for sd in sds)} triangles, "
         f"{sum(sd.width * sd.height for sd in sds)} pixels")
    b = renderer.Batch(max_depth=10)
    for sd in sds:
        b.add(sd)
    b.set_hit_capture(10)
    b.enable_accum64(True)
    b.render(8)
    F = b.stats().framesInFlight
    first = F * ((8 - 1) // F)          # first frame of the last pass
    nlast = 8 - first
    _say(f"frames per pass {F}: last pass holds frames {first}..7")
    outs = []
    for i, sd in enumerate(sds):
        outs.append((b.sample_radiance(i, nlast).reshape(nlast, -1, 3), b.hit_capture(i, nlast, 10),
                     b.accum(i).reshape(-1, 3), b.accum64(i).reshape(-1, 3),
                     b.env_cdf(i, sd.env_rgba.shape[:2])))
    b.set_hit_capture(0)
    rs = np.random.default_rng(33)
    got = {"rad": [], "hit": [], "crad": [], "chit": []}
    for i, sd in enumerate(sds):
        W, H = sd.width, sd.height
        rad, hits, mean, s64, env_cdf = outs[i]
        ref = oracle.OracleScene(sd, env_cdf=env_cdf)
        pix = np.sort(rs.choice(W * H, 1600, replace=False)).astype(np.int32)
        samples = np.empty((8, pix.size, 3), np.float32)
        for f in range(8):
            if f in (first, 7):
                cpu, tris = ref.trace_samples(pix, W, H, f, sd.max_depth, max_rec=sd.max_depth)
                got["rad"].append(rad[f - first][pix])
                got["hit"].append(hits[f - first][pix])
                got["crad"].append(cpu)
                got["chit"].append(tris)
            else:
                cpu, _ = ref.trace_samples(pix, W, H, f, sd.max_depth)
            samples[f] = cpu
        want_mean, want_s64 = _welford(samples)
        ok = np.array_equal(mean[pix], want_mean) and np.array_equal(s64[pix], want_s64)
        if i % 8 == 7 or not ok:
            _say(f"slot {i} {W}x{H}, {sd.triangles.shape[0]} tris: means and sums equal {ok}")
        assert ok, f"slot {i}: running mean or binary64 sums differ from the oracle's"
    _check("serving batch, first and last frame of the last pass", np.concatenate(got["rad"]),
           np.concatenate(got["hit"]), np.concatenate(got["crad"]), np.concatenate(got["chit"]))
