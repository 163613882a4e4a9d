"""Pins the oracle's albedo / normal AOVs (the denoiser's guide channels).

Definition: per sample, the BSDF base colour (p0.xyz) and the interpolated
shading normal at the path's first non-delta hit (diffuse, microfacet or
Disney), zeros when the path ends before one.  The checks recompute the
expected values from the oracle's own per-bounce hit record and the scene
tables, independently of the AOV code path:

* the recorded albedo is the p0 of the first hit whose BSDF type is non-delta;
* the recorded normal equals the face normal on flat (all-vertex-normals-equal)
  triangles and has unit length everywhere;
* paths that miss or only meet mirrors / glass record zeros;
* the per-pixel means of render_aov equal the plain means of the per-sample
  values (Welford vs two-pass, within float rounding).
"""
import numpy as np

import oracle
from bluetruthengine_amd import pod, scenes

NON_DELTA = (pod.BSDF_DIFFUSE, pod.BSDF_MICROFACET, pod.BSDF_DISNEY)


def _expected(sd, tris):
    n = tris.shape[0]
    alb = np.zeros((n, 3), np.float32)
    first = np.full(n, -1, np.int64)
    types = sd.bsdfs["type"][sd.tri_bsdf_ids]
    for i in range(n):
        for k in range(tris.shape[1]):
            t = tris[i, k]
            if t < 0:
                break
            if types[t] in NON_DELTA:
                first[i] = t
                alb[i] = sd.bsdfs["p0"][sd.tri_bsdf_ids[t]][:3]
                break
    return alb, first


def _check_scene(sd, frame):
    ref = oracle.OracleScene(sd)
    pix = np.arange(sd.width * sd.height, dtype=np.int32)
    _, tris, aov = ref.trace_samples(pix, sd.width, sd.height, frame, sd.max_depth, max_rec=sd.max_depth, aov=True)
    alb, first = _expected(sd, tris)
    assert np.array_equal(aov[:, :3], alb)
    none = first < 0
    assert np.all(aov[none] == 0.0)
    nrm = aov[~none, 3:]
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=2e-6)
    # flat triangles: all three vertex normals equal -> the face's normal
    tri = sd.triangles[first[~none]]
    n0, n1, n2 = (np.asarray(tri[k], np.float32) for k in ("n0", "n1", "n2"))
    flat = np.all(n0 == n1, axis=1) & np.all(n0 == n2, axis=1)
    assert flat.sum() > 0
    assert np.allclose(nrm[flat], n0[flat] / np.linalg.norm(n0[flat], axis=1, keepdims=True), atol=2e-6)
    return none.mean()


def test_cornell_aov_follows_first_non_delta_hit():
    sd = scenes.cornell_c1(width=48, height=48, spp=2, max_depth=8)
    frac_none = _check_scene(sd, 0)
    assert frac_none < 0.5      # the closed room: only emitter-free escapes record nothing


def test_envlit_aov_follows_first_non_delta_hit():
    sd = scenes.small_envlit(width=64, height=36)
    frac_none = _check_scene(sd, 1)
    assert 0.0 < frac_none < 1.0   # sky pixels record zeros


def test_render_aov_is_running_mean_of_samples():
    sd = scenes.cornell_c1(width=24, height=24, spp=3, max_depth=8)
    ref = oracle.OracleScene(sd)
    pix = np.arange(24 * 24, dtype=np.int32)
    per = [ref.trace_samples(pix, 24, 24, f, sd.max_depth, aov=True)[2] for f in range(3)]
    mean = np.mean(np.stack(per), axis=0)
    alb, nrm = ref.render_aov(24, 24, 0, 3, sd.max_depth)
    assert np.allclose(alb.reshape(-1, 3), mean[:, :3], atol=1e-6)
    assert np.allclose(nrm.reshape(-1, 3), mean[:, 3:], atol=1e-6)
