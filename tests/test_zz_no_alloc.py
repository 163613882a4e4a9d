"""The allocation-free ray-tracing entry point (bte_trace_rays_noalloc) against
bte_trace_rays, the entry point it mirrors: closest hits (triangle, t, u, v)
and any hits must be identical bit for bit, on two calls in a row of which the
second is smaller (the static scratch array keeps the previous call's
contents, and its work counter must start from zero again), then on an
any-hit call.  Invalid calls (BVH not built, host buffers, too many rays) must
return errors without launching anything.  All buffers of the new entry point
are device memory allocated here, by the caller.
"""
import ctypes

import numpy as np
import pytest

from bluetruthengine_amd import denoiser, renderer, scenes

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


class _Dev:
    """Device buffers of the caller (the HIP runtime libbte is linked against)."""

    def __init__(self):
        self.hip = denoiser._hip()
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(int(nbytes), 4)) == 0
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _rays(n, seed):
    rs = np.random.default_rng(seed)
    o = rs.uniform((-0.95, 0.05, -0.95), (0.95, 1.95, 0.95), (n, 3))
    d = rs.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)


def _noalloc(dev, rays, any_hit=False, tmax_per_ray=None):
    n = rays.shape[0]
    dr = dev.put(rays)
    dtm = dev.put(tmax_per_ray.astype(np.float32)) if tmax_per_ray is not None else None
    dtri, dt, du, dv = (dev.alloc(4 * n) for _ in range(4))
    renderer._call("bte_trace_rays_noalloc", dr, dtm, n, 1e-4, 1e30, int(any_hit), dtri, dt, du, dv)
    tri = dev.get(dtri, n, np.int32)
    if any_hit:
        return tri
    return tri, dev.get(dt, n, np.float32), dev.get(du, n, np.float32), dev.get(dv, n, np.float32)


def test_noalloc_trace_equals_trace_rays():
    _say("case: bte_trace_rays_noalloc vs bte_trace_rays, terrain room freq 40 (32k tris), "
         "closest 300k then 70k rays, any hit 50k rays, seeds 41-43")
    sd = scenes.terrain_c2(width=64, height=36, spp=1, max_depth=4, freq=40)
    renderer.Renderer(sd).build_accel()
    dev = _Dev()
    try:
        for n, seed in ((300_000, 41), (70_000, 42)):
            rays = _rays(n, seed)
            want = renderer.trace_rays(rays)
            got = _noalloc(dev, rays)
            same = [np.array_equal(a, b) for a, b in zip(got, want)]
            _say(f"closest hit, {n} rays: tri/t/u/v identical {same}, hits {(want[0] >= 0).mean():.3f}")
            assert all(same)
        rays = _rays(50_000, 43)
        tmax = np.random.default_rng(44).uniform(0.05, 3.0, rays.shape[0]).astype(np.float32)
        want = renderer.trace_rays(rays, any_hit=True, tmax_per_ray=tmax)
        got = _noalloc(dev, rays, any_hit=True, tmax_per_ray=tmax)
        _say(f"any hit, 50000 rays: identical {np.array_equal(got, want)}, occluded {want.mean():.3f}")
        assert np.array_equal(got, want)
    finally:
        dev.free()


def test_noalloc_trace_rejects_invalid_calls():
    _say("case: bte_trace_rays_noalloc error paths (host buffer, too many rays, BVH not built)")
    sd = scenes.cornell_c1(width=16, height=16, spp=1, max_depth=4)
    renderer.Renderer(sd).build_accel()
    dev = _Dev()
    try:
        rays = _rays(1000, 45)
        host = np.zeros(1000, np.int32)
        dr = dev.put(rays)
        dtri, dt, du, dv = (dev.alloc(4000) for _ in range(4))
        with pytest.raises(renderer.BteError, match="not device memory|not memory of device"):
            renderer._call("bte_trace_rays_noalloc", dr, None, 1000, 1e-4, 1e30, 0,
                           host.ctypes.data_as(ctypes.c_void_p), dt, du, dv)
        with pytest.raises(renderer.BteError, match="exceed the maximum"):
            renderer._call("bte_trace_rays_noalloc", dr, None, 1 << 31, 1e-4, 1e30, 0, dtri, dt, du, dv)
        # new triangles: the BVH is stale until bte_build_accel, and this entry point does not build it
        renderer.cudaInitScene(sd.triangles, np.zeros((1, 3), np.float32), np.zeros(sd.triangles.shape[0], np.int32),
                               sd.width, sd.height)
        with pytest.raises(renderer.BteError, match="not built"):
            renderer._call("bte_trace_rays_noalloc", dr, None, 1000, 1e-4, 1e30, 0, dtri, dt, du, dv)
        renderer.cudaInitOptix()
        renderer._call("bte_trace_rays_noalloc", dr, None, 1000, 1e-4, 1e30, 0, dtri, dt, du, dv)
        tri = dev.get(dtri, 1000, np.int32)
        assert np.array_equal(tri, renderer.trace_rays(rays)[0])
    finally:
        dev.free()
