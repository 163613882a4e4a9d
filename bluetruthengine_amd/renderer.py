"""Python mirror of the C ABI in ``include/bte_api.h`` (libbte.so).

The functions keep the reference's names and argument meaning
(``include/render_kernel.h:96-211``: cudaInitScene, cudaInitCamera, ...,
cudaRender); each raises ``BteError`` on a non-zero status instead of exiting,
so callers and tests can observe failures.  ``Renderer`` uploads a flattened
``scenes.SceneData`` in the reference's call order (``src/main.cpp:846-935``).

The library is loaded from ``bluetruthengine_amd/_lib/libbte.so``; there is no
CPU fallback — a missing library raises immediately.
"""
import ctypes
import os

import numpy as np

from . import pod

_HERE = os.path.dirname(os.path.abspath(__file__))
# BTE_LIB selects another build of the same library (e.g. a tuning variant).
LIB_PATH = os.environ.get("BTE_LIB") or os.path.join(_HERE, "_lib", "libbte.so")

BTE_MODE_MEGAKERNEL = 0
BTE_MODE_WAVEFRONT = 1
BTE_INSTR_KERNEL_TIMES = 1        # HIP events around every extend / shade / connect launch
BTE_INSTR_TRAVERSAL_COUNTS = 2    # BVH node and triangle-test counts

_p = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_ll = ctypes.c_longlong
_sz = ctypes.c_size_t


class BteError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f"{fn} failed with status {code}: {msg}")
        self.code = code


class BteFloat3(ctypes.Structure):
    _fields_ = [("x", _f), ("y", _f), ("z", _f)]


class BteConfig(ctypes.Structure):
    _fields_ = [("maxDepth", _i), ("megakernelSpp", _i), ("framesInFlight", _i), ("memoryBudgetMB", _i),
                ("traversalStack", _i), ("graphs", _i), ("workChunk", _i), ("traversalBlocksPerCU", _i),
                ("bvhWidth", _i), ("tailBounce", _i)]


def default_config():
    """The library's default configuration (bte_get_default_config) as a dict."""
    cfg = BteConfig()
    _call("bte_get_default_config", ctypes.byref(cfg))
    return {k: getattr(cfg, k) for k, _ in cfg._fields_}


def _tail_bounce(value):
    """The tail-pass bounce to configure: `value`, or the library default for None."""
    return int(value) if value is not None else default_config()["tailBounce"]


class BteSceneInfo(ctypes.Structure):
    _fields_ = [("width", _i), ("height", _i), ("triangleCount", _i), ("materialCount", _i), ("bsdfCount", _i),
                ("emitterCount", _i), ("emitterTriCount", _i), ("emitterCdfCount", _i), ("spotlightCount", _i),
                ("envWidth", _i), ("envHeight", _i), ("envIntensity", _f), ("envRotationDeg", _f),
                ("camera", ctypes.c_byte * 56)]


class BteSceneArrays(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "triangles", "materials", "triangleMaterialIds", "triangleBsdfIds", "bsdfs", "triangleEmission",
        "triangleEmitterFlags", "emitters", "emitterTriIndices", "emitterTriCdf", "sceneEmitterCdf", "spotlights",
        "envRGBA", "albedoTextures", "albedoTextureWidths", "albedoTextureHeights", "mrTextures", "mrTextureWidths",
        "mrTextureHeights")]


class BteAccelStats(ctypes.Structure):
    _fields_ = [("buildMs", ctypes.c_double), ("builder", _i), ("triangles", _i), ("nodes2", _i), ("nodes4", _i),
                ("nodes8", _i), ("depth2", _i), ("depth4", _i), ("depth8", _i)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BteStats(ctypes.Structure):
    _fields_ = [("renderMs", ctypes.c_double), ("extensionRays", _ll), ("hits", _ll), ("shadowRays", _ll),
                ("pathsTraced", _ll), ("bounces", _i), ("framesInFlight", _i), ("bvhNodes", _i), ("bvhDepth", _i),
                ("raysPerBounce", _ll * 64),
                ("extendMs", ctypes.c_double), ("shadeMs", ctypes.c_double), ("connectMs", ctypes.c_double),
                ("extendLaunches", _i), ("connectLaunches", _i),
                ("extendNodes", _ll), ("extendTris", _ll), ("connectNodes", _ll), ("connectTris", _ll),
                ("graphLaunches", _i), ("tailMs", ctypes.c_double), ("tailLaunches", _i)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "raysPerBounce"}
        d["raysPerBounce"] = [int(x) for x in self.raysPerBounce[: max(1, self.bounces)]]
        return d


# (name, argtypes) for every entry point of include/bte_api.h
API = {
    "bte_init_scene": [_p, _i, _p, _i, _p, _i, _i],
    "bte_init": [_p, BteFloat3, _i, _i],
    "bte_init_camera": [_p],
    "bte_init_triangle_emission": [_p, _i],
    "bte_init_triangle_emitter_flags": [_p, _i],
    "bte_init_emitter_table": [_p, _i, _p, _i, _p, _i, _p],
    "bte_init_bsdfs": [_p, _i, _p, _i],
    "bte_init_spotlights": [_p, _i],
    "bte_init_textures": [_p, _p, _p, _i],
    "bte_init_mr_textures": [_p, _p, _p, _i],
    "bte_init_env_map": [_p, _i, _i, _f, _f],
    "bte_build_accel": [],
    "bte_set_accel_builder": [_i],
    "bte_get_accel_stats": [_p],
    "bte_register_framebuffer": [_p],
    "bte_reset_accumulation": [_i, _i],
    "bte_render": [_i, _i],
    "bte_render_frames": [_i, _i, _i],
    "bte_toggle_denoiser": [],
    "bte_set_render_mode": [_i],
    "bte_get_render_mode": [],
    "bte_cleanup": [],
    "bte_set_config": [_p],
    "bte_get_config": [_p],
    "bte_get_default_config": [_p],
    "bte_get_stats": [_p],
    "bte_set_instrumentation": [_i],
    "bte_last_error": [],
    "bte_frame_index": [],
    "bte_set_frame_index": [_i],
    "bte_read_framebuffer": [_p, _sz],
    "bte_read_accum": [_p, _sz],
    "bte_enable_accum64": [_i],
    "bte_read_accum64": [_p, _sz],
    "bte_accum_device_ptr": [_p],
    "bte_read_sample_radiance": [_p, _sz],
    "bte_set_hit_capture": [_i],
    "bte_read_hit_capture": [_p, _sz],
    "bte_read_env_cdf": [_p, _p, _p],
    "bte_trace_rays": [_p, _p, _ll, _f, _f, _i, _p, _p, _p, _p],
    "bte_trace_rays_noalloc": [_p, _p, _ll, _f, _f, _i, _p, _p, _p, _p],
    "bte_autotune": [_i, _i, _i, _p],
    "bte_enable_aov": [_i],
    "bte_scene_load_xml": [ctypes.c_char_p, ctypes.c_char_p, _p],
    "bte_scene_last_error": [],
    "bte_scene_info": [_p, _p],
    "bte_scene_arrays": [_p, _p],
    "bte_scene_upload": [_p],
    "bte_scene_free": [_p],
    "bte_read_aov": [_p, _p, ctypes.c_size_t],
    "bte_batch_add_scene": [_i, _i],
    "bte_batch_autotune": [_i, _p],
    "bte_batch_scene_count": [],
    "bte_batch_render": [_i],
    "bte_batch_reset": [],
    "bte_batch_clear": [],
    "bte_batch_frame_index": [],
    "bte_batch_read_accum": [_i, _p, _sz],
    "bte_batch_read_framebuffer": [_i, _p, _sz],
    "bte_batch_read_accum64": [_i, _p, _sz],
    "bte_batch_read_sample_radiance": [_i, _p, _sz],
    "bte_batch_read_hit_capture": [_i, _p, _sz],
    "bte_batch_read_env_cdf": [_i, _p, _p, _p],
    "bte_batch_denoise": [],
    "bte_batch_read_denoised": [_i, _p, _sz],
    "bte_denoise": [],
    "bte_read_denoised": [_p, _sz],
    "bte_denoiser_param_count": [],
    "bte_denoiser_init": [ctypes.c_uint],
    "bte_denoiser_release": [],
    "bte_denoiser_set_weights": [_p, _ll],
    "bte_denoiser_load_file": [ctypes.c_char_p],
    "bte_denoiser_get_weights": [_p, _ll],
    "bte_denoiser_get_grads": [_p, _ll],
    "bte_denoiser_set_loss": [_i],
    "bte_denoiser_set_flips": [_i],
    "bte_denoiser_set_conv_kernel": [_i],
    "bte_denoise_image": [_p, _p, _p, _i, _i, _p],
    "bte_denoiser_get_flips": [],
    "bte_denoiser_dataset_reserve": [_i, _i, _i],
    "bte_denoiser_capture": [_i, _i],
    "bte_denoiser_train": [_i, _i, _f, _f, ctypes.c_uint, _p],
    "bte_denoiser_forward_host": [_p, _i, _i, _i, _p],
    "bte_denoiser_train_step_host": [_p, _p, _i, _i, _i, _f, _f, _p],
    "bte_denoiser_tensor_shape": [_i, _p],
    "bte_denoiser_read_tensor": [_i, _p, _ll],
}
RESTYPE = {"bte_last_error": ctypes.c_char_p, "bte_scene_last_error": ctypes.c_char_p,
           "bte_denoiser_param_count": _ll, "bte_scene_free": None}

_lib = None


def lib():
    """Load libbte.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in API.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = RESTYPE.get(name, _i)
        _lib = L
    return _lib


def _call(name, *args):
    rc = getattr(lib(), name)(*args)
    if rc < 0:
        raise BteError(name, rc, lib().bte_last_error().decode(errors="replace"))
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ---- reference-named wrappers (include/render_kernel.h:103-211) -------------
def cudaInitScene(triangles, materials, triangle_material_ids, width, height):
    tris = np.ascontiguousarray(triangles, pod.TRIANGLE)
    mats = np.ascontiguousarray(materials, np.float32).reshape(-1, 3)
    ids = np.ascontiguousarray(triangle_material_ids, np.int32)
    return _call("bte_init_scene", _ptr(tris), tris.shape[0], _ptr(mats), mats.shape[0], _ptr(ids), width, height)


def cudaInitCamera(camera):
    cam = np.ascontiguousarray(np.asarray(camera, pod.CAMERA).reshape(()))
    return _call("bte_init_camera", _ptr(cam))


def cudaInitTriangleEmission(emission):
    if emission is None:
        return _call("bte_init_triangle_emission", None, 0)
    e = np.ascontiguousarray(emission, np.float32).reshape(-1, 3)
    return _call("bte_init_triangle_emission", _ptr(e), e.shape[0])


def cudaInitTriangleEmitterFlags(flags):
    if flags is None:
        return _call("bte_init_triangle_emitter_flags", None, 0)
    f = np.ascontiguousarray(flags, np.uint8)
    return _call("bte_init_triangle_emitter_flags", _ptr(f), f.size)


def cudaInitEmitterTable(emitters, tri_indices, tri_cdf, scene_cdf):
    if emitters is None or len(emitters) == 0:
        return _call("bte_init_emitter_table", None, 0, None, 0, None, 0, None)
    e = np.ascontiguousarray(emitters, pod.EMITTER)
    ti = np.ascontiguousarray(tri_indices, np.int32)
    tc = np.ascontiguousarray(tri_cdf, np.float32)
    sc = np.ascontiguousarray(scene_cdf, np.float32)
    return _call("bte_init_emitter_table", _ptr(e), e.shape[0], _ptr(ti), ti.size, _ptr(tc), tc.size, _ptr(sc))


def cudaInitBsdfs(bsdfs, tri_bsdf_ids):
    b = np.ascontiguousarray(bsdfs, pod.BSDF)
    ids = np.ascontiguousarray(tri_bsdf_ids, np.int32)
    return _call("bte_init_bsdfs", _ptr(b), b.shape[0], _ptr(ids), ids.size)


def cudaInitSpotlights(spotlights):
    if spotlights is None or len(spotlights) == 0:
        return _call("bte_init_spotlights", None, 0)
    s = np.ascontiguousarray(spotlights, pod.SPOTLIGHT)
    return _call("bte_init_spotlights", _ptr(s), s.shape[0])


def cudaInitEnvMap(rgba, intensity=1.0, rotation_deg=0.0):
    if rgba is None:
        return _call("bte_init_env_map", None, 0, 0, 0.0, 0.0)
    e = np.ascontiguousarray(rgba, np.float32)
    H, W = e.shape[:2]
    return _call("bte_init_env_map", _ptr(e), W, H, float(intensity), float(rotation_deg))


def _texture_set(name, images):
    """cudaInitTextures / cudaInitMrTextures: one (H, W, 4) uint8 RGBA array or None per material."""
    if not images:
        return _call(name, None, None, None, 0)
    n = len(images)
    keep = [None if im is None else np.ascontiguousarray(im, np.uint8) for im in images]
    ptrs = (ctypes.c_void_p * n)(*[None if im is None else im.ctypes.data for im in keep])
    ws = np.array([0 if im is None else im.shape[1] for im in keep], np.int32)
    hs = np.array([0 if im is None else im.shape[0] for im in keep], np.int32)
    return _call(name, ptrs, _ptr(ws), _ptr(hs), n)


def cudaInitTextures(images):
    return _texture_set("bte_init_textures", images)


def cudaInitMrTextures(images):
    return _texture_set("bte_init_mr_textures", images)


def cudaInitOptix():
    return _call("bte_build_accel")


def cudaResetAccumulation(width, height):
    return _call("bte_reset_accumulation", width, height)


def cudaRender(width, height):
    return _call("bte_render", width, height)


def cudaSetRenderMode(mode):
    return _call("bte_set_render_mode", int(mode))


def cudaGetRenderMode():
    return _call("bte_get_render_mode")


def cudaToggleDenoiser():
    return _call("bte_toggle_denoiser")


def cudaCleanup():
    return _call("bte_cleanup")


# ---- headless renderer -----------------------------------------------------------
def _upload(scene):
    """The reference's upload sequence (src/main.cpp:846-935) for one SceneData."""
    cudaInitScene(scene.triangles, np.zeros((1, 3), np.float32), scene.tri_material_ids, scene.width,
                  scene.height)
    cudaInitCamera(scene.camera)
    cudaInitTriangleEmission(scene.tri_emission)
    cudaInitTriangleEmitterFlags(scene.tri_emitter_flags)
    cudaInitEmitterTable(scene.emitters, scene.emitter_tri_indices, scene.emitter_tri_cdf, scene.scene_emitter_cdf)
    cudaInitBsdfs(scene.bsdfs, scene.tri_bsdf_ids)
    cudaInitSpotlights(scene.spotlights)
    cudaInitEnvMap(scene.env_rgba, scene.env_intensity, scene.env_rotation_deg)
    cudaInitTextures(scene.textures)
    cudaInitMrTextures(scene.mr_textures)


def _rgb(flat, w, h):
    return flat.reshape(h, w, 3)


class Renderer:
    """Uploads a SceneData through the C ABI and renders into device buffers."""

    def __init__(self, scene, max_depth=None, frames_in_flight=0, memory_budget_mb=0, mode=BTE_MODE_WAVEFRONT,
                 megakernel_spp=5, traversal_stack=0, graphs=1, work_chunk=0, traversal_blocks_per_cu=0, bvh_width=4,
                 tail_bounce=None):
        self.scene = scene
        lib()
        cudaCleanup()
        _upload(scene)
        self.set_config(max_depth if max_depth is not None else scene.max_depth, frames_in_flight,
                        memory_budget_mb, megakernel_spp, traversal_stack, graphs, work_chunk,
                        traversal_blocks_per_cu, bvh_width, tail_bounce)
        cudaSetRenderMode(mode)
        cudaResetAccumulation(scene.width, scene.height)

    def set_config(self, max_depth, frames_in_flight=0, memory_budget_mb=0, megakernel_spp=5, traversal_stack=0,
                   graphs=1, work_chunk=0, traversal_blocks_per_cu=0, bvh_width=4, tail_bounce=None):
        """tail_bounce None: the library default."""
        cfg = BteConfig(int(max_depth), int(megakernel_spp), int(frames_in_flight), int(memory_budget_mb),
                        int(traversal_stack), int(graphs), int(work_chunk), int(traversal_blocks_per_cu),
                        int(bvh_width), _tail_bounce(tail_bounce))
        _call("bte_set_config", ctypes.byref(cfg))

    def config(self):
        cfg = BteConfig()
        _call("bte_get_config", ctypes.byref(cfg))
        return {k: getattr(cfg, k) for k, _ in cfg._fields_}

    def set_tail_bounce(self, bounce):
        """Sets BteConfig.tailBounce (0 = per-bounce passes to the end), the rest unchanged; returns the old value."""
        cfg = BteConfig()
        _call("bte_get_config", ctypes.byref(cfg))
        old = cfg.tailBounce
        cfg.tailBounce = int(bounce)
        _call("bte_set_config", ctypes.byref(cfg))
        return old

    def autotune(self, frames):
        """Picks the fastest launch shape for `frames`-frame passes at the scene's resolution (resets the
        accumulation); returns (best ms, config dict)."""
        ms = ctypes.c_double(0.0)
        _call("bte_autotune", self.scene.width, self.scene.height, int(frames), ctypes.byref(ms))
        return ms.value, self.config()

    def build_accel(self):
        cudaInitOptix()

    def accel_stats(self):
        """Builder, build time, node counts and depths of the last BVH build."""
        return accel_stats()

    def reset(self):
        cudaResetAccumulation(self.scene.width, self.scene.height)

    def render(self, frames=1):
        _call("bte_render_frames", self.scene.width, self.scene.height, int(frames))

    def set_instrumentation(self, flags):
        _call("bte_set_instrumentation", int(flags))

    def stats(self):
        s = BteStats()
        _call("bte_get_stats", ctypes.byref(s))
        return s

    def frame_index(self):
        return _call("bte_frame_index")

    def set_frame_index(self, index):
        _call("bte_set_frame_index", int(index))

    def accum(self):
        W, H = self.scene.width, self.scene.height
        out = np.empty((H, W, 3), np.float32)
        _call("bte_read_accum", _ptr(out), out.size)
        return out

    def framebuffer(self):
        W, H = self.scene.width, self.scene.height
        out = np.empty((H, W), np.uint32)
        _call("bte_read_framebuffer", _ptr(out), out.size)
        return out

    def enable_accum64(self, on=True):
        _call("bte_enable_accum64", int(on))

    def accum64(self):
        W, H = self.scene.width, self.scene.height
        out = np.empty((H, W, 3), np.float64)
        _call("bte_read_accum64", _ptr(out), out.size)
        return out

    def enable_aov(self, on=True):
        _call("bte_enable_aov", int(on))

    def aov(self):
        """(albedo, normal) running means, each (H, W, 3) float32, row 0 = bottom."""
        W, H = self.scene.width, self.scene.height
        alb = np.empty((H, W, 3), np.float32)
        nrm = np.empty((H, W, 3), np.float32)
        _call("bte_read_aov", _ptr(alb), _ptr(nrm), alb.size)
        return alb, nrm

    def sample_radiance(self, frames):
        W, H = self.scene.width, self.scene.height
        out = np.empty((frames, H, W, 3), np.float32)
        _call("bte_read_sample_radiance", _ptr(out), out.size)
        return out

    def set_hit_capture(self, max_rec):
        _call("bte_set_hit_capture", int(max_rec))

    def hit_capture(self, frames, max_rec):
        W, H = self.scene.width, self.scene.height
        out = np.empty((frames, H * W, max_rec), np.int32)
        _call("bte_read_hit_capture", _ptr(out), out.size)
        return out

    def env_cdf(self):
        H, W = self.scene.env_rgba.shape[:2]
        marg = np.empty(H + 1, np.float32)
        cond = np.empty(H * (W + 1), np.float32)
        total = ctypes.c_float(0)
        _call("bte_read_env_cdf", _ptr(marg), _ptr(cond), ctypes.byref(total))
        return marg, cond, np.float32(total.value)


class Batch:
    """Scenes of uneven size and resolution traced together (bte_batch_*): every
    bounce of every scene runs in one persistent wavefront launch; each slot
    accumulates exactly what a Renderer of that scene alone would."""

    def __init__(self, max_depth=10, frames_in_flight=0, memory_budget_mb=0, traversal_stack=0, graphs=1,
                 work_chunk=0, traversal_blocks_per_cu=0, bvh_width=4, aov=False, tail_bounce=None):
        lib()
        cudaCleanup()
        cfg = BteConfig(int(max_depth), 5, int(frames_in_flight), int(memory_budget_mb), int(traversal_stack),
                        int(graphs), int(work_chunk), int(traversal_blocks_per_cu), int(bvh_width),
                        _tail_bounce(tail_bounce))
        _call("bte_set_config", ctypes.byref(cfg))
        cudaSetRenderMode(BTE_MODE_WAVEFRONT)
        _call("bte_enable_aov", 1 if aov else 0)
        self.sizes = []

    def add(self, scene):
        """Uploads `scene` and moves it into a new slot; returns the slot index."""
        _upload(scene)
        slot = _call("bte_batch_add_scene", int(scene.width), int(scene.height))
        self.sizes.append((int(scene.width), int(scene.height)))
        return slot

    def __len__(self):
        return len(self.sizes)

    def render(self, frames=1):
        _call("bte_batch_render", int(frames))

    def autotune(self, frames):
        """Launch-shape search over `frames`-frame batch renders (resets the accumulation); returns
        (best ms, config dict)."""
        ms = ctypes.c_double(0.0)
        _call("bte_batch_autotune", int(frames), ctypes.byref(ms))
        cfg = BteConfig()
        _call("bte_get_config", ctypes.byref(cfg))
        return ms.value, {k: getattr(cfg, k) for k, _ in cfg._fields_}

    def reset(self):
        _call("bte_batch_reset")

    def clear(self):
        _call("bte_batch_clear")
        self.sizes = []

    def frame_index(self):
        return _call("bte_batch_frame_index")

    def stats(self):
        s = BteStats()
        _call("bte_get_stats", ctypes.byref(s))
        return s

    def accum(self, slot):
        w, h = self.sizes[slot]
        out = np.empty(w * h * 3, np.float32)
        _call("bte_batch_read_accum", int(slot), _ptr(out), out.size)
        return _rgb(out, w, h)

    def enable_accum64(self, on=True):
        """Binary64 sample sums per slot (takes effect from the next render, which then restarts the batch)."""
        _call("bte_enable_accum64", int(on))

    def accum64(self, slot):
        w, h = self.sizes[slot]
        out = np.empty(w * h * 3, np.float64)
        _call("bte_batch_read_accum64", int(slot), _ptr(out), out.size)
        return _rgb(out, w, h)

    def framebuffer(self, slot):
        w, h = self.sizes[slot]
        out = np.empty(w * h, np.uint32)
        _call("bte_batch_read_framebuffer", int(slot), _ptr(out), out.size)
        return out.reshape(h, w)

    def set_hit_capture(self, max_rec):
        """Record the hit triangle of the first max_rec bounces of every path of the next pass (0 = off)."""
        _call("bte_set_hit_capture", int(max_rec))

    def sample_radiance(self, slot, frames):
        """Per-sample radiance of slot's scene in the last pass: (frames, H, W, 3)."""
        w, h = self.sizes[slot]
        out = np.empty((frames, h, w, 3), np.float32)
        _call("bte_batch_read_sample_radiance", int(slot), _ptr(out), out.size)
        return out

    def hit_capture(self, slot, frames, max_rec):
        w, h = self.sizes[slot]
        out = np.empty((frames, h * w, max_rec), np.int32)
        _call("bte_batch_read_hit_capture", int(slot), _ptr(out), out.size)
        return out

    def env_cdf(self, slot, env_shape):
        """(marginal, conditional, total) of slot's env map as built on the GPU; env_shape = (H, W)."""
        H, W = env_shape
        marg = np.empty(H + 1, np.float32)
        cond = np.empty(H * (W + 1), np.float32)
        total = ctypes.c_float(0)
        _call("bte_batch_read_env_cdf", int(slot), _ptr(marg), _ptr(cond), ctypes.byref(total))
        return marg, cond, np.float32(total.value)

    def denoise(self):
        _call("bte_batch_denoise")

    def denoised(self, slot):
        w, h = self.sizes[slot]
        out = np.empty(w * h * 3, np.float32)
        _call("bte_batch_read_denoised", int(slot), _ptr(out), out.size)
        return _rgb(out, w, h)


def accel_stats():
    """Builder, build time, node counts and depths of the last BVH build (bte_get_accel_stats)."""
    s = BteAccelStats()
    _call("bte_get_accel_stats", ctypes.byref(s))
    return s.as_dict()


def set_accel_builder(builder):
    """0: binned SAH on the host, 1 (default): on the GPU (takes effect at the next BVH build)."""
    _call("bte_set_accel_builder", int(builder))


def trace_rays(rays, tmin=1e-4, tmax=1e30, any_hit=False, tmax_per_ray=None):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = rays.shape[0]
    tri = np.empty(n, np.int32)
    t = np.empty(n, np.float32)
    u = np.empty(n, np.float32)
    v = np.empty(n, np.float32)
    tm = None if tmax_per_ray is None else np.ascontiguousarray(tmax_per_ray, np.float32)
    _call("bte_trace_rays", _ptr(rays), _ptr(tm), n, float(tmin), float(tmax), int(any_hit), _ptr(tri), _ptr(t),
          _ptr(u), _ptr(v))
    return (tri, t, u, v) if not any_hit else tri
