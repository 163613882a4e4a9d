"""ORACLE — test infrastructure only.

ctypes wrapper around ``oracle/_build/libbte_oracle.so``, the CPU restatement
of the reference integrator (see ``oracle/bte_oracle.h``).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
module; the product (``bluetruthengine_amd``) never does.
"""
import ctypes
import os
import subprocess

import numpy as np

from bluetruthengine_amd import pod

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libbte_oracle.so")
_lib = None

_f = ctypes.c_float
_i = ctypes.c_int
_l = ctypes.c_long
_p = ctypes.c_void_p
_u32 = ctypes.c_uint32


_V4_PATH = os.path.join(_HERE, "_build", "libbte_oracle_v4.so")
_V4_FLAGS = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl", "avx2", "bmi2", "fma", "movbe"}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def cpu_info():
    """CPU model, logical CPUs, physical cores and the flags of this host (/proc/cpuinfo)."""
    model, flags, cores = "unknown", set(), set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name":
                    model = v
                elif k == "flags" and not flags:
                    flags = set(v.split())
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    logical = os.cpu_count() or 1
    return {"model": model, "logical_cpus": logical, "physical_cores": len(cores) or logical, "flags": flags}


def baseline_library():
    """(path, build) of the oracle build the CPU baseline should time: the
    x86-64-v4 build where this CPU has its features, else the portable one."""
    if _V4_FLAGS <= cpu_info()["flags"] and os.path.exists(_V4_PATH):
        return _V4_PATH, "g++ -O3 -march=x86-64-v4 -fopenmp -ffp-contract=off"
    return _LIB_PATH, "g++ -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off"


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _load(_LIB_PATH)
    return _lib


class using_library:
    """Context manager: every oracle call inside uses the library at `path`
    (a copy built with a planted mistake, oracle/Makefile `planted`); objects
    created inside must be released inside."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        global _lib
        self.saved = lib()
        _lib = _load(self.path)
        return _lib

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def build_planted():
    subprocess.run(["make", "-s", "-j", "5", "-C", _HERE, "planted"], check=True)
    return [os.path.join(_HERE, "_build", "planted", f"libbte_oracle_plant{k}.so") for k in range(1, 6)]


def _load(path):
    """Loads an oracle library and declares its entry points."""
    L = ctypes.CDLL(path)
    for name in ("ora_sin", "ora_cos", "ora_tan", "ora_exp", "ora_log", "ora_acos"):
        getattr(L, name).restype = _f
        getattr(L, name).argtypes = [_f]
    L.ora_pow.restype = _f
    L.ora_pow.argtypes = [_f, _f]
    L.ora_atan2.restype = _f
    L.ora_atan2.argtypes = [_f, _f]
    L.ora_math_batch.argtypes = [_i, _p, _p, _p, _l]
    L.ora_hash.restype = _u32
    L.ora_hash.argtypes = [_u32]
    L.ora_texture_sample.argtypes = [_p, _p, _p, _l, _p]
    L.ora_texture_sample.restype = None
    L.ora_rng_u32.argtypes = [_u32, _p, _i]
    L.ora_rng_f01.argtypes = [_u32, _p, _i]
    L.ora_pixel_seed.restype = _u32
    L.ora_pixel_seed.argtypes = [_u32, _u32]
    L.ora_bsdf_eval.argtypes = [_p, _p, _p, _p]
    L.ora_bsdf_pdf.restype = _f
    L.ora_bsdf_pdf.argtypes = [_p, _p, _p]
    L.ora_bsdf_sample.argtypes = [_p, _p, _f, _f, _f, _p]
    L.ora_bsdf_batch_sample.argtypes = [_p, _p, _p, _l, _p]
    L.ora_bsdf_batch_eval.argtypes = [_p, _p, _p, _l, _p]
    L.ora_fresnel_dielectric.restype = _f
    L.ora_fresnel_dielectric.argtypes = [_f, _f, _f]
    L.ora_frame.argtypes = [_p, _p, _p, _p]
    L.ora_primary_ray.argtypes = [_p, _i, _i, _i, _i, _f, _f, _p]
    L.ora_env_build_cdf.restype = _f
    L.ora_env_build_cdf.argtypes = [_p, _i, _i, _i, _p, _p]
    L.ora_scene_create.restype = _p
    L.ora_scene_create.argtypes = [_p]
    L.ora_scene_destroy.argtypes = [_p]
    L.ora_trace_closest.argtypes = [_p, _p, _l, _f, _f, _i, _p, _p, _p, _p]
    L.ora_trace_occluded.argtypes = [_p, _p, _p, _l, _p]
    L.ora_env_eval.argtypes = [_p, _p, _p]
    L.ora_env_pdf.restype = _f
    L.ora_env_pdf.argtypes = [_p, _p]
    L.ora_env_sample.argtypes = [_p, _f, _f, _p]
    L.ora_dir_to_equirect.argtypes = [_p, _f, _p]
    L.ora_sample_emitter.argtypes = [_p, _p, _f, _f, _p]
    L.ora_render.argtypes = [_p, _i, _i, _i, _i, _i, _p, _p, _p, _i]
    L.ora_trace_samples.argtypes = [_p, _p, _l, _i, _i, _i, _i, _p, _p, _i, _p]
    L.ora_render_megakernel.argtypes = [_p, _i, _i, _i, _i, _i, _i, _p]
    L.ora_render_aov.argtypes = [_p, _i, _i, _i, _i, _i, _p, _p]
    L.ora_pack_gamma.restype = _u32
    L.ora_pack_gamma.argtypes = [_f, _f, _f]
    L.ora_num_threads.restype = _i
    L.ora_spot_nee.argtypes = [_p, _p, _p, _p, _p, _p, _i, _p]
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


MATH_FN = {"sin": 0, "cos": 1, "tan": 2, "exp": 3, "log": 4, "pow": 5, "atan2": 6, "acos": 7}


def math_batch(fn, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, np.float32)
    out = np.empty_like(x)
    lib().ora_math_batch(MATH_FN[fn], _ptr(x), _ptr(y), _ptr(out), x.size)
    return out


def rng_u32(seed, n):
    out = np.empty(n, np.uint32)
    lib().ora_rng_u32(seed, _ptr(out), n)
    return out


def rng_f01(seed, n):
    out = np.empty(n, np.float32)
    lib().ora_rng_f01(seed, _ptr(out), n)
    return out


def bsdf_eval_pdf(bsdf, wi, wo):
    """Batched eval (3) + pdf (1) of one BsdfData for local wi/wo arrays [n,3]."""
    b = np.ascontiguousarray(np.asarray(bsdf, pod.BSDF).reshape(()))
    wi = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
    out = np.empty((wi.shape[0], 4), np.float32)
    lib().ora_bsdf_batch_eval(_ptr(b), _ptr(wi), _ptr(wo), wi.shape[0], _ptr(out))
    return out


def bsdf_sample(bsdf, wi, rnd):
    """Batched sample: rnd [n,3] = (rndExtra, u0, u1). Returns [n,9]
    = weight(3), wo(3), eta, measure, pdf."""
    b = np.ascontiguousarray(np.asarray(bsdf, pod.BSDF).reshape(()))
    wi = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    rnd = np.ascontiguousarray(rnd, np.float32).reshape(-1, 3)
    out = np.empty((wi.shape[0], 9), np.float32)
    lib().ora_bsdf_batch_sample(_ptr(b), _ptr(wi), _ptr(rnd), wi.shape[0], _ptr(out))
    return out


def env_build_cdf(rgba, mode):
    rgba = np.ascontiguousarray(rgba, np.float32)
    H, W = rgba.shape[:2]
    marg = np.empty(H + 1, np.float32)
    cond = np.empty(H * (W + 1), np.float32)
    total = lib().ora_env_build_cdf(_ptr(rgba), W, H, mode, _ptr(marg), _ptr(cond))
    return marg, cond, np.float32(total)


class _SceneDesc(ctypes.Structure):
    _fields_ = [("triangles", _p), ("triangleCount", _i),
                ("bsdfs", _p), ("bsdfCount", _i),
                ("triangleBsdfIds", _p),
                ("triangleEmission", _p),
                ("triangleEmitterFlags", _p),
                ("emitters", _p), ("emitterCount", _i),
                ("emitterTriIndices", _p), ("emitterTriIndexCount", _i),
                ("emitterTriCdf", _p), ("emitterTriCdfCount", _i),
                ("sceneEmitterCdf", _p),
                ("spotlights", _p), ("spotlightCount", _i),
                ("envRGBA", _p), ("envWidth", _i), ("envHeight", _i),
                ("envIntensity", _f), ("envRotationDeg", _f),
                ("envMarginalCdf", _p), ("envConditionalCdf", _p),
                ("envTotalPower", _f),
                ("camera", ctypes.c_byte * 56),
                ("triangleMaterialIds", _p),
                ("texels", _p), ("texInfo", _p), ("texCount", _i),
                ("mrTexels", _p), ("mrInfo", _p), ("mrCount", _i)]


def pack_textures(images):
    """A texture set (list of (H, W, 4) uint8 arrays or None, one per material)
    as the back-to-back RGBA8 texels (uint32) and (offset, width, height, 0) rows."""
    info = np.zeros((len(images), 4), np.int32)
    parts, off = [], 0
    for i, im in enumerate(images):
        if im is None:
            continue
        im = np.ascontiguousarray(im, np.uint8)
        h, w = im.shape[:2]
        info[i] = (off, w, h, 0)
        parts.append(im.reshape(-1, 4).view(np.uint32).reshape(-1))
        off += w * h
    texels = np.concatenate(parts) if parts else np.zeros(1, np.uint32)
    return texels, info


def texture_sample(image, uv):
    """Bilinear wrap lookup of one (H, W, 4) uint8 texture at (n, 2) normalized uv -> (n, 4)."""
    texels, info = pack_textures([image])
    uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
    out = np.empty((uv.shape[0], 4), np.float32)
    lib().ora_texture_sample(_ptr(texels), _ptr(info), _ptr(uv), uv.shape[0], _ptr(out))
    return out


class OracleScene:
    """Oracle view of a flattened scene (bluetruthengine_amd.scenes.SceneData).

    ``env_cdf`` = (marginal, conditional, totalPower) lets a test feed the
    exact CDF arrays the GPU built; by default the oracle's binary64 build is
    used.
    """

    def __init__(self, sd, env_cdf=None):
        L = lib()
        self._keep = []

        def keep(a, dtype=None):
            if a is None:
                return None
            a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype))
            self._keep.append(a)
            return a

        d = _SceneDesc()
        tris = keep(sd.triangles, pod.TRIANGLE)
        d.triangles = _ptr(tris).value
        d.triangleCount = tris.shape[0]
        b = keep(sd.bsdfs, pod.BSDF)
        d.bsdfs = _ptr(b).value
        d.bsdfCount = b.shape[0]
        d.triangleBsdfIds = _ptr(keep(sd.tri_bsdf_ids, np.int32)).value
        d.triangleEmission = _ptr(keep(sd.tri_emission, np.float32)).value if sd.tri_emission is not None else None
        d.triangleEmitterFlags = _ptr(keep(sd.tri_emitter_flags, np.uint8)).value if sd.tri_emitter_flags is not None else None
        ne = 0 if sd.emitters is None else len(sd.emitters)
        if ne:
            e = keep(sd.emitters, pod.EMITTER)
            d.emitters = _ptr(e).value
            d.emitterCount = ne
            ti = keep(sd.emitter_tri_indices, np.int32)
            d.emitterTriIndices = _ptr(ti).value
            d.emitterTriIndexCount = ti.size
            tc = keep(sd.emitter_tri_cdf, np.float32)
            d.emitterTriCdf = _ptr(tc).value
            d.emitterTriCdfCount = tc.size
            d.sceneEmitterCdf = _ptr(keep(sd.scene_emitter_cdf, np.float32)).value
        ns = 0 if sd.spotlights is None else len(sd.spotlights)
        if ns:
            d.spotlights = _ptr(keep(sd.spotlights, pod.SPOTLIGHT)).value
            d.spotlightCount = ns
        if sd.env_rgba is not None:
            env = keep(sd.env_rgba, np.float32)
            H, W = env.shape[:2]
            d.envRGBA = _ptr(env).value
            d.envWidth, d.envHeight = W, H
            d.envIntensity = sd.env_intensity
            d.envRotationDeg = sd.env_rotation_deg
            if env_cdf is None:
                env_cdf = env_build_cdf(env, 1)
            marg, cond, total = env_cdf
            d.envMarginalCdf = _ptr(keep(marg, np.float32)).value
            d.envConditionalCdf = _ptr(keep(cond, np.float32)).value
            d.envTotalPower = float(total)
        cam = keep(np.asarray(sd.camera, pod.CAMERA).reshape(()))
        ctypes.memmove(d.camera, cam.ctypes.data, 56)
        textures = getattr(sd, "textures", None)
        if textures:
            d.triangleMaterialIds = _ptr(keep(sd.tri_material_ids, np.int32)).value
            tx, ti = pack_textures(textures)
            d.texels, d.texInfo, d.texCount = _ptr(keep(tx)).value, _ptr(keep(ti)).value, len(textures)
            mr = getattr(sd, "mr_textures", None)
            if mr:
                mx, mi = pack_textures(mr)
                d.mrTexels, d.mrInfo, d.mrCount = _ptr(keep(mx)).value, _ptr(keep(mi)).value, len(mr)
        self._desc = d
        self.handle = L.ora_scene_create(ctypes.byref(d))
        self.sd = sd

    def __del__(self):
        if getattr(self, "handle", None):
            lib().ora_scene_destroy(self.handle)
            self.handle = None

    def trace_closest(self, rays, tmin=1e-4, tmax=1e30, brute=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = rays.shape[0]
        tri = np.empty(n, np.int32)
        t = np.empty(n, np.float32)
        u = np.empty(n, np.float32)
        v = np.empty(n, np.float32)
        lib().ora_trace_closest(self.handle, _ptr(rays), n, tmin, tmax, int(brute),
                                _ptr(tri), _ptr(t), _ptr(u), _ptr(v))
        return tri, t, u, v

    def trace_occluded(self, rays, tmax):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        tmax = np.ascontiguousarray(tmax, np.float32).reshape(-1)
        occ = np.empty(rays.shape[0], np.uint8)
        lib().ora_trace_occluded(self.handle, _ptr(rays), _ptr(tmax), rays.shape[0], _ptr(occ))
        return occ

    def env_eval(self, d):
        d = np.ascontiguousarray(d, np.float32)
        out = np.empty(3, np.float32)
        lib().ora_env_eval(self.handle, _ptr(d), _ptr(out))
        return out

    def env_pdf(self, d):
        d = np.ascontiguousarray(d, np.float32)
        return lib().ora_env_pdf(self.handle, _ptr(d))

    def env_sample(self, u0, u1):
        out = np.empty(7, np.float32)
        lib().ora_env_sample(self.handle, u0, u1, _ptr(out))
        return out

    def sample_emitter(self, origin, u0, u1):
        o = np.ascontiguousarray(origin, np.float32)
        out = np.empty(11, np.float32)
        lib().ora_sample_emitter(self.handle, _ptr(o), u0, u1, _ptr(out))
        return out

    def spot_nee(self, p, n, d_in, thr, bsdf, index):
        """Spotlight NEE record: (ok, direction (3), tMax, contribution (3))."""
        f = lambda a: np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(np.asarray(bsdf, pod.BSDF).reshape(()))
        out = np.empty(8, np.float32)
        args = [f(p), f(n), f(d_in), f(thr)]
        lib().ora_spot_nee(self.handle, *[_ptr(a) for a in args], _ptr(b), int(index), _ptr(out))
        return bool(out[0]), out[1:4], out[4], out[5:8]

    def render(self, W, H, frame0, nframes, max_depth, accum=None, sum64=None, rgba8=False, threads=0):
        if accum is None:
            accum = np.zeros((H, W, 3), np.float32)
        out8 = np.empty((H, W), np.uint32) if rgba8 else None
        lib().ora_render(self.handle, W, H, frame0, nframes, max_depth, _ptr(accum), _ptr(sum64),
                         _ptr(out8), threads)
        return (accum, out8) if rgba8 else accum

    def render_aov(self, W, H, frame0, nframes, max_depth):
        """Albedo / normal running means (H, W, 3) of the first non-delta hit."""
        alb = np.zeros((H, W, 3), np.float32)
        nrm = np.zeros((H, W, 3), np.float32)
        lib().ora_render_aov(self.handle, W, H, frame0, nframes, max_depth, _ptr(alb), _ptr(nrm))
        return alb, nrm

    def render_megakernel(self, W, H, frame0, nframes, spp, max_depth, accum=None):
        if accum is None:
            accum = np.zeros((H, W, 3), np.float32)
        lib().ora_render_megakernel(self.handle, W, H, frame0, nframes, spp, max_depth, _ptr(accum))
        return accum

    def trace_samples(self, pixels, W, H, frame, max_depth, max_rec=0, aov=False):
        pixels = np.ascontiguousarray(pixels, np.int32)
        n = pixels.size
        rad = np.empty((n, 3), np.float32)
        tris = np.empty((n, max_rec), np.int32) if max_rec else None
        aov6 = np.empty((n, 6), np.float32) if aov else None
        lib().ora_trace_samples(self.handle, _ptr(pixels), n, W, H, frame, max_depth, _ptr(rad),
                                _ptr(tris), max_rec, _ptr(aov6))
        return (rad, tris, aov6) if aov else (rad, tris)
