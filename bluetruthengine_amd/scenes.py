"""Procedural scene generators and scene flattening.

``SceneBuilder`` flattens meshes into the reference's per-triangle side tables
and emitter tables exactly as ``src/main.cpp:686-758`` does (one emitter per
emissive mesh, binary32 running area CDF, power weight = area x Rec.709
luminance, binary32 running scene CDF).  The generators build the benchmark
workloads:

* ``cornell_c1``   256x256 room + two icospheres (mirror, glass), 16 spp, depth 8
* ``terrain_c2``   1280x720 room + 200k-triangle displaced icosphere terrain,
                   64 random Disney slots, 4 quad lights, 64 spp, depth 10
* ``envlit_c3``    1920x1080, 500 displaced UV spheres (1M triangles), GGX /
                   dielectric materials, 2048x1024 procedural HDR sky, 16 spp,
                   depth 12

All randomness comes from ``numpy.random.default_rng(seed)`` with fixed seeds,
so every buffer is reproducible (tests pin their SHA-256).
"""
from dataclasses import dataclass, field
import hashlib
import os

import numpy as np

from . import pod

F32 = np.float32


@dataclass
class SceneData:
    name: str
    width: int
    height: int
    spp: int
    max_depth: int
    triangles: np.ndarray
    bsdfs: np.ndarray
    tri_bsdf_ids: np.ndarray
    tri_material_ids: np.ndarray
    tri_emission: np.ndarray
    tri_emitter_flags: np.ndarray
    emitters: np.ndarray = None
    emitter_tri_indices: np.ndarray = None
    emitter_tri_cdf: np.ndarray = None
    scene_emitter_cdf: np.ndarray = None
    spotlights: np.ndarray = None
    env_rgba: np.ndarray = None
    env_intensity: float = 1.0
    env_rotation_deg: float = 0.0
    camera: np.ndarray = None
    # texture maps per material (tri_material_ids): (H, W, 4) uint8 RGBA or None
    textures: list = None
    mr_textures: list = None
    extra: dict = field(default_factory=dict)

    def digest(self):
        h = hashlib.sha256()
        for a in (self.triangles, self.bsdfs, self.tri_bsdf_ids, self.tri_emission, self.tri_emitter_flags,
                  self.emitters, self.emitter_tri_indices, self.emitter_tri_cdf, self.scene_emitter_cdf,
                  self.env_rgba, self.camera):
            if a is not None:
                h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()


# ---------------------------------------------------------------------------
# camera (src/main.cpp:533-564: forward, right = forward x up, up' = right x forward)
# ---------------------------------------------------------------------------
def _normalize32(v):
    v = np.asarray(v, F32)
    d = F32(v[0] * v[0] + v[1] * v[1]) + F32(v[2] * v[2])
    return (v * (F32(1.0) / np.sqrt(F32(d)))).astype(F32)


def _cross32(a, b):
    a = np.asarray(a, F32)
    b = np.asarray(b, F32)
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], F32)


def make_camera(eye, target, up, fov_y_deg, width, height):
    eye = np.asarray(eye, F32)
    fwd = _normalize32(np.asarray(target, F32) - eye)
    right = _normalize32(_cross32(fwd, up))
    true_up = _cross32(right, fwd)
    c = np.zeros((), pod.CAMERA)
    c["origin"] = eye
    c["forward"] = fwd
    c["right"] = right
    c["up"] = true_up
    c["fovYRadians"] = F32(fov_y_deg) * F32(0.017453292519943295)
    c["aspect"] = F32(width) / F32(height)
    return c


def crop_camera(eye, target, up, fov_y_deg, full_w, full_h, cx, cy, size):
    """A size x size view aimed at pixel (cx, cy) of the full_w x full_h view
    (eye, target, up, fov_y_deg), with that window's field of view: training
    crops keep the pixel footprint of the full-resolution frame."""
    base = make_camera(eye, target, up, fov_y_deg, full_w, full_h)
    th = np.tan(np.radians(fov_y_deg) / 2.0)
    ndx = 2.0 * (cx + 0.5) / full_w - 1.0
    ndy = 2.0 * (cy + 0.5) / full_h - 1.0
    d = (base["forward"].astype(np.float64) + base["right"] * (ndx * full_w / full_h * th) + base["up"] * (ndy * th))
    fov = np.degrees(2.0 * np.arctan(th * size / full_h))
    return make_camera(eye, np.asarray(eye, np.float64) + d, up, fov, size, size)


# ---------------------------------------------------------------------------
# flattening (src/main.cpp:686-758)
# ---------------------------------------------------------------------------
def triangle_area32(v0, v1, v2):
    """triangleAreaHost (src/main.cpp:475-489) in binary32, vectorised."""
    a = (v1 - v0).astype(F32)
    b = (v2 - v0).astype(F32)
    cx = a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1]
    cy = a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2]
    cz = a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]
    ln = np.sqrt(((cx * cx + cy * cy) + cz * cz).astype(F32))
    return (F32(0.5) * ln).astype(F32)


class SceneBuilder:
    def __init__(self):
        self.tris = []
        self.bsdf_ids = []
        self.emission = []
        self.flags = []
        self.bsdfs = []
        self.emitters = []
        self.em_tri_idx = []
        self.em_tri_cdf = []
        self.scene_cdf = [F32(0.0)]
        self.weight_sum = F32(0.0)
        self.count = 0

    def add_bsdf(self, b):
        self.bsdfs.append(np.asarray(b, pod.BSDF))
        return len(self.bsdfs) - 1

    def add_mesh(self, pos, nrm, bsdf_id, emission=None):
        """pos, nrm: [n,3,3] float32 (vertex-major).  bsdf_id: int or [n] ints."""
        pos = np.asarray(pos, F32).reshape(-1, 3, 3)
        n = pos.shape[0]
        if nrm is None:
            fn = np.cross(pos[:, 1] - pos[:, 0], pos[:, 2] - pos[:, 0]).astype(F32)
            fn /= np.maximum(np.linalg.norm(fn, axis=1, keepdims=True), F32(1e-30))
            nrm = np.repeat(fn[:, None, :], 3, axis=1)
        nrm = np.asarray(nrm, F32).reshape(-1, 3, 3)
        t = np.zeros(n, pod.TRIANGLE)
        t["v0"], t["v1"], t["v2"] = pos[:, 0], pos[:, 1], pos[:, 2]
        t["n0"], t["n1"], t["n2"] = nrm[:, 0], nrm[:, 1], nrm[:, 2]
        first = self.count
        self.tris.append(t)
        self.bsdf_ids.append(np.broadcast_to(np.asarray(bsdf_id, np.int32), (n,)).copy())
        if emission is not None:
            emc = np.asarray(emission, F32)
            self.emission.append(np.repeat(emc[None, :], n, axis=0))
            self.flags.append(np.ones(n, np.uint8))
            # one emitter per emissive mesh, running binary32 sums
            tri_off = len(self.em_tri_idx)
            cdf_off = len(self.em_tri_cdf)
            self.em_tri_cdf.append(F32(0.0))
            areas = triangle_area32(pos[:, 0], pos[:, 1], pos[:, 2])
            acc = F32(0.0)
            for i in range(n):
                self.em_tri_idx.append(first + i)
                acc = F32(acc + areas[i])
                self.em_tri_cdf.append(acc)
            lum = F32(F32(F32(0.2126) * emc[0]) + F32(F32(0.7152) * emc[1])) + F32(F32(0.0722) * emc[2])
            weight = F32(acc * lum)
            self.weight_sum = F32(self.weight_sum + weight)
            self.scene_cdf.append(self.weight_sum)
            e = np.zeros((), pod.EMITTER)
            e["triIndexOffset"] = tri_off
            e["triCount"] = n
            e["cdfOffset"] = cdf_off
            e["radiance"] = emc
            e["areaSum"] = acc
            e["powerWeight"] = weight
            self.emitters.append(e)
        else:
            self.emission.append(np.zeros((n, 3), F32))
            self.flags.append(np.zeros(n, np.uint8))
        self.count += n
        return first

    def build(self, name, width, height, spp, max_depth, camera, env=None, env_intensity=1.0,
              env_rotation_deg=0.0, spotlights=None):
        tris = np.concatenate(self.tris)
        sd = SceneData(
            name=name, width=width, height=height, spp=spp, max_depth=max_depth,
            triangles=tris,
            bsdfs=np.array(self.bsdfs, pod.BSDF).reshape(-1),
            tri_bsdf_ids=np.concatenate(self.bsdf_ids).astype(np.int32),
            tri_material_ids=np.zeros(tris.shape[0], np.int32),
            tri_emission=np.concatenate(self.emission).astype(F32),
            tri_emitter_flags=np.concatenate(self.flags).astype(np.uint8),
            camera=camera, env_rgba=env, env_intensity=env_intensity, env_rotation_deg=env_rotation_deg,
            spotlights=spotlights)
        if self.emitters:
            sd.emitters = np.array(self.emitters, pod.EMITTER).reshape(-1)
            sd.emitter_tri_indices = np.array(self.em_tri_idx, np.int32)
            sd.emitter_tri_cdf = np.array(self.em_tri_cdf, F32)
            sd.scene_emitter_cdf = np.array(self.scene_cdf, F32)
        return sd


# ---------------------------------------------------------------------------
# geometry helpers
# ---------------------------------------------------------------------------
def quad(c0, c1, c2, c3):
    """Two triangles (c0,c1,c2), (c0,c2,c3); geometric normal = (c1-c0)x(c2-c0)."""
    c = [np.asarray(x, F32) for x in (c0, c1, c2, c3)]
    return np.array([[c[0], c[1], c[2]], [c[0], c[2], c[3]]], F32)


def room_quads():
    """[-1,1] x [0,2] x [-1,1] open at +z; geometric normals point inward."""
    floor = quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1))
    ceil = quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1))
    back = quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1))
    left = quad((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1))
    right = quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1))
    return np.concatenate([floor, ceil, back, left, right])


def ceiling_light(cx, cz, sx, sz, y):
    """Downward-facing quad (geometric normal -y) centred at (cx, y, cz)."""
    x0, x1 = cx - sx / 2, cx + sx / 2
    z0, z1 = cz - sz / 2, cz + sz / 2
    return quad((x0, y, z0), (x1, y, z0), (x1, y, z1), (x0, y, z1))


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return v, f


def geodesic_sphere(freq):
    """Class-I geodesic sphere: each icosahedron face split into freq^2
    triangles, vertices projected to the unit sphere and de-duplicated.
    Returns (vertices [V,3] float64, faces [20 freq^2, 3] int64), outward winding."""
    iv, ifc = _icosahedron()
    verts = []
    faces = []
    base = 0
    for a, b, c in ifc:
        A, B, C = iv[a], iv[b], iv[c]
        idx = {}
        pts = []
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                k = freq - i - j
                p = (k * A + i * B + j * C) / freq
                idx[(i, j)] = base + len(pts)
                pts.append(p)
        for i in range(freq):
            for j in range(freq - i):
                faces.append((idx[(i, j)], idx[(i + 1, j)], idx[(i, j + 1)]))
                if i + j + 1 < freq:
                    faces.append((idx[(i + 1, j)], idx[(i + 1, j + 1)], idx[(i, j + 1)]))
        verts.extend(pts)
        base += len(pts)
    v = np.array(verts, np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array(faces, np.int64)
    key = np.round(v * 1e7).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    v = v[first]
    f = inv.reshape(-1)[f]
    # outward winding check
    pv = v[f]
    nrm = np.cross(pv[:, 1] - pv[:, 0], pv[:, 2] - pv[:, 0])
    flip = np.einsum("ij,ij->i", nrm, pv.mean(axis=1)) < 0
    f[flip] = f[flip][:, ::-1]
    return v, f


def icosphere(subdiv):
    """Loop-style midpoint subdivision of the icosahedron (20 * 4^subdiv faces)."""
    v, f = _icosahedron()
    v = list(v)
    for _ in range(subdiv):
        mid = {}
        nf = []

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = (v[a] + v[b]) / 2.0
                v.append(p / np.linalg.norm(p))
                mid[key] = len(v) - 1
            return mid[key]
        for a, b, c in f:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = np.array(nf, np.int64)
    v = np.array(v, np.float64)
    pv = v[f]
    nrm = np.cross(pv[:, 1] - pv[:, 0], pv[:, 2] - pv[:, 0])
    flip = np.einsum("ij,ij->i", nrm, pv.mean(axis=1)) < 0
    f[flip] = f[flip][:, ::-1]
    return v, f


def uv_sphere(slices, bands):
    """UV sphere with `bands` latitude bands: pole fans + quads,
    2*slices*(bands-1) triangles, no degenerate triangles."""
    verts = [np.array([0.0, 1.0, 0.0])]
    for b in range(1, bands):
        th = np.pi * b / bands
        for s in range(slices):
            ph = 2 * np.pi * s / slices
            verts.append(np.array([np.sin(th) * np.cos(ph), np.cos(th), np.sin(th) * np.sin(ph)]))
    verts.append(np.array([0.0, -1.0, 0.0]))
    v = np.array(verts)
    south = len(v) - 1

    def ring(b, s):
        return 1 + (b - 1) * slices + (s % slices)
    f = []
    for s in range(slices):
        f.append((0, ring(1, s + 1), ring(1, s)))
    for b in range(1, bands - 1):
        for s in range(slices):
            a0, a1 = ring(b, s), ring(b, s + 1)
            b0, b1 = ring(b + 1, s), ring(b + 1, s + 1)
            f.append((a0, a1, b1))
            f.append((a0, b1, b0))
    for s in range(slices):
        f.append((south, ring(bands - 1, s), ring(bands - 1, s + 1)))
    f = np.array(f, np.int64)
    pv = v[f]
    nrm = np.cross(pv[:, 1] - pv[:, 0], pv[:, 2] - pv[:, 0])
    flip = np.einsum("ij,ij->i", nrm, pv.mean(axis=1)) < 0
    f[flip] = f[flip][:, ::-1]
    return v, f


def _hash_u32(x):
    x = np.asarray(x, np.uint32).copy()
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x7feb352d)
    x ^= x >> np.uint32(15)
    x *= np.uint32(0x846ca68b)
    x ^= x >> np.uint32(16)
    return x


def value_noise3(p, seed, octaves=3, base_freq=3.0):
    """Seeded 3D value noise in [-1, 1] (trilinear, smoothstep-faded)."""
    p = np.asarray(p, np.float64)
    total = np.zeros(p.shape[0])
    amp, norm = 1.0, 0.0
    freq = base_freq
    for o in range(octaves):
        q = p * freq
        i0 = np.floor(q).astype(np.int64)
        fr = q - i0
        w = fr * fr * (3 - 2 * fr)
        acc = np.zeros(p.shape[0])
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    ix = (i0[:, 0] + dx).astype(np.uint32)
                    iy = (i0[:, 1] + dy).astype(np.uint32)
                    iz = (i0[:, 2] + dz).astype(np.uint32)
                    h = _hash_u32(ix * np.uint32(73856093) ^ iy * np.uint32(19349663) ^ iz * np.uint32(83492791)
                                  ^ np.uint32((seed * 1000003 + o * 7919) & 0xffffffff))
                    val = h.astype(np.float64) / 4294967295.0 * 2.0 - 1.0
                    wx = w[:, 0] if dx else 1 - w[:, 0]
                    wy = w[:, 1] if dy else 1 - w[:, 1]
                    wz = w[:, 2] if dz else 1 - w[:, 2]
                    acc += wx * wy * wz * val
        total += amp * acc
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def vertex_normals(v, f):
    pv = v[f]
    fn = np.cross(pv[:, 1] - pv[:, 0], pv[:, 2] - pv[:, 0])   # area weighted
    vn = np.zeros_like(v)
    for k in range(3):
        np.add.at(vn, f[:, k], fn)
    vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-30)
    return vn


def mesh_arrays(v, f, vn):
    return v[f].astype(F32), vn[f].astype(F32)


# ---------------------------------------------------------------------------
# C1: Cornell-style room, icospheres
# ---------------------------------------------------------------------------
def cornell_c1(width=256, height=256, spp=16, max_depth=8):
    sb = SceneBuilder()
    white = sb.add_bsdf(pod.bsdf_diffuse((0.7, 0.7, 0.7)))
    mirror = sb.add_bsdf(pod.bsdf_mirror())
    glass = sb.add_bsdf(pod.bsdf_dielectric(1.5, 1.0))
    sb.add_mesh(room_quads(), None, white)
    sb.add_mesh(ceiling_light(0.0, 0.0, 0.5, 0.5, 1.999), None, white, emission=(15.0, 15.0, 15.0))
    v, f = icosphere(3)
    for center, bid in (((-0.45, 0.35, -0.3), mirror), ((0.45, 0.35, 0.25), glass)):
        pos = v * 0.35 + np.asarray(center)
        p, n = mesh_arrays(pos, f, v)
        sb.add_mesh(p, n, bid)
    cam = make_camera((0, 1, 3.6), (0, 1, 0), (0, 1, 0), 40.0, width, height)
    return sb.build("c1_cornell", width, height, spp, max_depth, cam)


# ---------------------------------------------------------------------------
# C2: room + 200k-triangle displaced geodesic terrain, 64 Disney slots, 4 lights
# ---------------------------------------------------------------------------
def random_disney_slots(rng, count):
    out = []
    for _ in range(count):
        bc = rng.uniform(0, 1, 3)
        metallic, roughness, specular, clearcoat, sheen = rng.uniform(0, 1, 5)
        out.append(pod.bsdf_disney(tuple(bc), roughness=roughness, metallic=metallic, specular=specular,
                                   clearcoat=clearcoat, sheen=sheen))
    return out


def terrain_c2(width=1280, height=720, spp=64, max_depth=10, freq=100, seed=None, camera=None):
    """seed None: the fixed benchmark scene; otherwise another draw of the
    same generator (materials, terrain noise, lights).  camera: optional
    (eye, target, fov) override."""
    s_mat, s_noise, s_light = (3, 2, 4) if seed is None else (1000 + 3 * seed, 1001 + 3 * seed, 1002 + 3 * seed)
    sb = SceneBuilder()
    white = sb.add_bsdf(pod.bsdf_diffuse((0.7, 0.7, 0.7)))
    sb.add_mesh(room_quads(), None, white)
    slots = [sb.add_bsdf(b) for b in random_disney_slots(np.random.default_rng(s_mat), 64)]
    v, f = geodesic_sphere(freq)
    disp = value_noise3(v, seed=s_noise)
    pos = v * (0.6 + 0.08 * disp)[:, None] + np.array([0.0, 0.6, 0.0])
    vn = vertex_normals(pos, f)
    p, n = mesh_arrays(pos, f, vn)
    tri_ids = np.arange(f.shape[0], dtype=np.uint32)
    slot_of = (_hash_u32(tri_ids * np.uint32(0x9E3779B9) + np.uint32(3)) % np.uint32(64)).astype(np.int32)
    sb.add_mesh(p, n, np.asarray(slots, np.int32)[slot_of])
    rng = np.random.default_rng(s_light)
    for _ in range(4):
        sx, sz = rng.uniform(0.2, 0.5, 2)
        cx, cz = rng.uniform(-0.7, 0.7, 2)
        le = rng.uniform(5, 20)
        sb.add_mesh(ceiling_light(cx, cz, sx, sz, 1.98), None, white, emission=(le, le, le))
    cam = camera if camera is not None else make_camera((0, 1, 3.6), (0, 1, 0), (0, 1, 0), 40.0, width, height)
    return sb.build("c2_terrain", width, height, spp, max_depth, cam)


# ---------------------------------------------------------------------------
# C3: env-lit field of 500 displaced UV spheres (1M triangles)
# ---------------------------------------------------------------------------
def sky_env(width=2048, height=1024, seed=5, sun_elev_deg=40.0, sun_azim_deg=60.0, sun_radius_deg=0.5,
            sun_radiance=1e4):
    """Procedural equirect HDR sky, row 0 = +Y, u = (phi + pi) / 2pi with
    phi = atan2(z, x) (the convention of rt_env_map.cuh:36-67)."""
    rng = np.random.default_rng(seed)
    u = (np.arange(width) + 0.5) / width
    vv = (np.arange(height) + 0.5) / height
    phi = u * 2 * np.pi - np.pi
    th = vv * np.pi
    st, ct = np.sin(th)[:, None], np.cos(th)[:, None]
    d = np.stack([st * np.cos(phi)[None, :], np.broadcast_to(ct, (height, width)), st * np.sin(phi)[None, :]], -1)
    horizon = np.array([1.0, 1.0, 1.0]) * 0.8
    zenith = np.array([0.25, 0.45, 1.0]) * 1.5
    ground = np.array([0.3, 0.28, 0.25]) * 0.3
    y = d[..., 1:2]
    t = np.clip(y, 0, 1) ** 0.5
    sky = horizon * (1 - t) + zenith * t
    col = np.where(y > 0, sky, ground)
    el, az = np.radians(sun_elev_deg), np.radians(sun_azim_deg)
    sun = np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)])
    cosang = np.einsum("hwc,c->hw", d, sun)
    disc = cosang >= np.cos(np.radians(sun_radius_deg))
    col = np.where(disc[..., None], sun_radiance, col)
    noise = rng.lognormal(0.0, 1.0, (height, width, 1))
    col = col * noise
    rgba = np.concatenate([col, np.ones((height, width, 1))], -1).astype(F32)
    return rgba


C3_EYE, C3_TARGET, C3_FOV = (0.0, 2.5, 8.0), (0.0, 0.0, 0.0), 45.0


def envlit_c3(width=1920, height=1080, spp=16, max_depth=12, n_spheres=500, slices=50, bands=21,
              env_size=(2048, 1024), seed=None, camera=None):
    """seed None: the fixed benchmark scene; otherwise another draw of the same
    generators (sphere layout, materials, sky noise and sun position)."""
    s_geo, s_mat, s_env = (6, 7, 5) if seed is None else (2000 + 3 * seed, 2001 + 3 * seed, 2002 + 3 * seed)
    sb = SceneBuilder()
    ground_b = sb.add_bsdf(pod.bsdf_diffuse((0.5, 0.5, 0.5)))
    G = 10.0
    sb.add_mesh(quad((-G, 0, -G), (-G, 0, G), (G, 0, G), (G, 0, -G)), None, ground_b)
    rng = np.random.default_rng(s_geo)
    mat_rng = np.random.default_rng(s_mat)
    v, f = uv_sphere(slices, bands)
    for i in range(n_spheres):
        r = rng.uniform(0.05, 0.2)
        cx, cz = rng.uniform(-4, 4, 2)
        disp = value_noise3(v + i * 7.31, seed=6, octaves=2, base_freq=2.0)
        pos = v * (r * (1.0 + 0.15 * disp))[:, None] + np.array([cx, r, cz])
        vn = vertex_normals(pos, f)
        if mat_rng.uniform() < 0.8:
            b = pod.bsdf_disney(tuple(mat_rng.uniform(0.5, 1.0, 3)), roughness=mat_rng.uniform(0.05, 1.0),
                                metallic=1.0, spec_trans=0.0)
        else:
            b = pod.bsdf_dielectric(1.5, 1.0)
        bid = sb.add_bsdf(b)
        p, n = mesh_arrays(pos, f, vn)
        sb.add_mesh(p, n, bid)
    cam = camera if camera is not None else make_camera(C3_EYE, C3_TARGET, (0, 1, 0), C3_FOV, width, height)
    if seed is None:
        env = sky_env(*env_size)
    else:
        srng = np.random.default_rng(s_env)
        env = sky_env(*env_size, seed=s_env, sun_elev_deg=srng.uniform(15, 75), sun_azim_deg=srng.uniform(0, 360))
    return sb.build("c3_envlit", width, height, spp, max_depth, cam, env=env, env_intensity=1.0)


def small_envlit(width=128, height=72, n_spheres=12, env_size=(256, 128)):
    """Reduced C3 (same generators) for CPU-sized parity tests."""
    sd = envlit_c3(width, height, spp=2, max_depth=12, n_spheres=n_spheres, slices=16, bands=9, env_size=env_size)
    sd.name = "c3_small"
    return sd


def serve_c5(n=64):
    """The batched-serving workload: n uneven scenes (seeds 2000..2000+n-1),
    square frames of 256..1024 px (multiples of 64), 10k..500k triangles
    (log-uniform), the terrain-room generator's Disney materials and quad
    lights plus a 512x256 procedural sky with its own sun, 8 spp, depth 10."""
    rs = np.random.default_rng(2000)
    out = []
    for k in range(n):
        side = int(rs.integers(4, 17)) * 64
        tris = float(np.exp(rs.uniform(np.log(1e4), np.log(5e5))))
        freq = max(4, int(round(np.sqrt(tris / 20.0))))
        seed = 2000 + k
        sd = terrain_c2(side, side, spp=8, max_depth=10, freq=freq, seed=seed)
        srng = np.random.default_rng(seed)
        sd.env_rgba = sky_env(512, 256, seed=seed, sun_elev_deg=srng.uniform(15, 75),
                              sun_azim_deg=srng.uniform(0, 360))
        sd.name = f"c5_serve_{k}"
        out.append(sd)
    return out


CONFIGS = {"c1": cornell_c1, "c2": terrain_c2, "c3": envlit_c3, "c5": serve_c5}


# ---------------------------------------------------------------------------
# scene.xml / OBJ / .hdr through the library's host layer (src/scene.cpp,
# src/main.cpp:524-930)
# ---------------------------------------------------------------------------
def load_xml(path, base_dir=None, spp=1, max_depth=8, name=None):
    """Parses a scene.xml with libbte's loader and returns its flattened scene
    as SceneData (the arrays the reference's main.cpp hands to the renderer).
    Relative mesh / map paths resolve against base_dir, default the working
    directory (as the reference's)."""
    import ctypes
    from . import renderer
    L = renderer.lib()
    handle = ctypes.c_void_p()
    base = "" if base_dir is None else base_dir     # the reference resolves against the working directory
    if L.bte_scene_load_xml(os.fsencode(path), os.fsencode(base), ctypes.byref(handle)) != 0:
        raise ValueError(L.bte_scene_last_error().decode(errors="replace"))
    try:
        info = renderer.BteSceneInfo()
        arr = renderer.BteSceneArrays()
        L.bte_scene_info(handle, ctypes.byref(info))
        L.bte_scene_arrays(handle, ctypes.byref(arr))

        def take(ptr, dtype, count, shape=None):
            if count == 0 or not ptr:
                return np.zeros(0 if shape is None else (0,) + tuple(shape[1:]), dtype)
            n = count * np.dtype(dtype).itemsize
            buf = (ctypes.c_char * n).from_address(ptr)
            a = np.frombuffer(bytes(buf), dtype=dtype).copy()
            return a.reshape(shape) if shape is not None else a

        T = info.triangleCount
        cam = np.frombuffer(bytes(info.camera), dtype=pod.CAMERA).reshape(()).copy()
        sd = SceneData(
            name=name or os.path.splitext(os.path.basename(path))[0], width=info.width, height=info.height,
            spp=spp, max_depth=max_depth,
            triangles=take(arr.triangles, pod.TRIANGLE, T),
            bsdfs=take(arr.bsdfs, pod.BSDF, info.bsdfCount),
            tri_bsdf_ids=take(arr.triangleBsdfIds, np.int32, T),
            tri_material_ids=take(arr.triangleMaterialIds, np.int32, T),
            tri_emission=take(arr.triangleEmission, np.float32, 3 * T, (T, 3)),
            tri_emitter_flags=take(arr.triangleEmitterFlags, np.uint8, T),
            camera=cam)
        if info.emitterCount:
            sd.emitters = take(arr.emitters, pod.EMITTER, info.emitterCount)
            sd.emitter_tri_indices = take(arr.emitterTriIndices, np.int32, info.emitterTriCount)
            sd.emitter_tri_cdf = take(arr.emitterTriCdf, np.float32, info.emitterCdfCount)
            sd.scene_emitter_cdf = take(arr.sceneEmitterCdf, np.float32, info.emitterCount + 1)
        if info.spotlightCount:
            sd.spotlights = take(arr.spotlights, pod.SPOTLIGHT, info.spotlightCount)
        if info.envWidth:
            sd.env_rgba = take(arr.envRGBA, np.float32, info.envWidth * info.envHeight * 4,
                               (info.envHeight, info.envWidth, 4))
            sd.env_intensity = float(info.envIntensity)
            sd.env_rotation_deg = float(info.envRotationDeg)
        sd.extra["materials"] = take(arr.materials, np.float32, 3 * info.materialCount, (info.materialCount, 3))
        M = info.materialCount

        def images(p_arr, p_w, p_h):
            if not (M and p_arr):
                return None
            ptrs = (ctypes.c_void_p * M).from_address(p_arr)
            ws, hs = take(p_w, np.int32, M), take(p_h, np.int32, M)
            tex = [take(ptrs[i], np.uint8, int(hs[i]) * int(ws[i]) * 4, (int(hs[i]), int(ws[i]), 4))
                   if ptrs[i] else None for i in range(M)]
            return tex if any(t is not None for t in tex) else None
        sd.textures = images(arr.albedoTextures, arr.albedoTextureWidths, arr.albedoTextureHeights)
        sd.mr_textures = images(arr.mrTextures, arr.mrTextureWidths, arr.mrTextureHeights)
        if sd.mr_textures is not None and sd.textures is None:
            # metallic-roughness maps apply through the per-material texture table
            # (rt_material.cuh:26-56), which the reference always uploads
            sd.textures = [None] * M
        return sd
    finally:
        L.bte_scene_free(handle)
