"""glTF 2.0 / GLB <model> entries of the scene host layer (bte_gltf.cpp,
following src/gltf_loader.cpp and src/main.cpp:760-818), on the CPU.

Files are written by tests/gltf_writer.py; expected values are recomputed
here in binary64 from the reference's rules: placement T Rz Ry Rx S after the
node graph, node matrices (column-major `matrix`, else T R S with the
quaternion's rotation matrix), normals by the inverse transpose, geometric
normals without NORMAL, normalized integer texture coordinates, list / strip
(odd triangles swap their first two corners) / fan modes, Disney mapping of
the pbrMetallicRoughness model and the KHR_materials_* extensions, the base
colour factor baked into the albedo map, one emitter per emissive material,
backdrop / outline / outlier pruning, and the default material appended only
when a primitive has none."""
import math
import os

import numpy as np
import pytest

from bluetruthengine_amd import pod, scenes
from gltf_writer import Gltf
from test_png_decoder import encode_png
from test_scene_loader import _synthetic

F32 = np.float32


def _quat_mat(x, y, z, w):
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y), 0],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x), 0],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y), 0], [0, 0, 0, 1]])


def _trs(t=(0, 0, 0), q=(0, 0, 0, 1), s=(1, 1, 1)):
    T = np.eye(4)
    T[:3, 3] = t
    return T @ _quat_mat(*q) @ np.diag(list(s) + [1])


def _axis_rot(axis, deg):
    a = math.radians(deg)
    c, s = math.cos(a), math.sin(a)
    R = np.eye(4)
    i, j = (axis + 1) % 3, (axis + 2) % 3
    R[i, i], R[i, j], R[j, i], R[j, j] = c, -s, s, c
    return R


def _placement(pos, rot, scale):
    T = np.eye(4)
    T[:3, 3] = pos
    return T @ _axis_rot(2, rot[2]) @ _axis_rot(1, rot[1]) @ _axis_rot(0, rot[0]) @ np.diag(list(scale) + [1])


def _xp(M, p):
    return (M @ np.append(p, 1.0))[:3]


def _xn(M, n):
    v = np.linalg.inv(M[:3, :3]).T @ n
    return v / np.linalg.norm(v)


def _write_scene(tmp_path, model_name, place_attrs, with_meshes=True):
    if with_meshes:
        _synthetic(tmp_path)
        xml = (tmp_path / "scene.xml").read_text()
    else:
        xml = ('<scene><window width="32" height="24" /><camera eyeX="0" eyeY="0" eyeZ="5" lookAtX="0" lookAtY="0" '
               'lookAtZ="0" upX="0" upY="1" upZ="0" fovY="40" /></scene>')
    xml = xml.replace("</scene>", f'<model filename="{model_name}" {place_attrs} />\n</scene>')
    (tmp_path / "m.xml").write_text(xml)
    return str(tmp_path / "m.xml")


def _geometry_model():
    g = Gltf()
    rs = np.random.default_rng(3)
    posA = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], F32)
    nrmA = rs.normal(size=(4, 3)).astype(F32)
    nrmA /= np.linalg.norm(nrmA, axis=1, keepdims=True)
    uvA = rs.uniform(-1, 2, (4, 2)).astype(F32)
    mA = g.mesh([{"attributes": {"POSITION": g.accessor(posA, minmax=True), "NORMAL": g.accessor(nrmA),
                                 "TEXCOORD_0": g.accessor(uvA)},
                  "indices": g.accessor(np.array([0, 1, 2, 0, 2, 3], np.uint16)), "material": 0},
                 # lines: skipped; an index past the vertices: that triangle skipped
                 {"attributes": {"POSITION": g.accessor(posA)}, "mode": 1},
                 {"attributes": {"POSITION": g.accessor(posA)}, "indices": g.accessor(np.array([0, 1, 9, 1, 2, 3], np.uint32)),
                  "material": 0}])
    posB = np.array([[0, 0, 0], [1, 0, 0.2], [0, 1, 0], [1, 1, 0.3], [0.5, 2, 0]], F32)
    uvB = np.array([[0, 0], [65535, 0], [0, 65535], [65535, 65535], [32768, 1]], np.uint16)
    mB = g.mesh([{"attributes": {"POSITION": g.accessor(posB), "TEXCOORD_0": g.accessor(uvB, normalized=True)},
                  "indices": g.accessor(np.array([0, 1, 2, 3, 4], np.uint8)), "mode": 5, "material": 1}])
    posC = np.array([[0, 0, 0], [1, 0, 0], [1.2, 0.8, 0], [0.5, 1.4, 0.1], [-0.3, 0.9, 0]], F32)
    nrmC = np.tile(np.array([[0, 0.6, 0.8]], F32), (5, 1))
    ids = g.interleaved([posC, nrmC])
    mC = g.mesh([{"attributes": {"POSITION": ids[0], "NORMAL": ids[1]}, "mode": 6}])
    q = np.array([0.2, -0.3, 0.1, 0.9])
    q = q / np.linalg.norm(q)
    n1 = g.node(mesh=mA, name="a")
    Mb = _trs((0.1, 0.2, 0.3), (0, 0.5 ** 0.5, 0, 0.5 ** 0.5), (1, 2, 1))
    n2 = g.node(mesh=mB, matrix=[float(v) for v in Mb.T.reshape(-1)])
    n0 = g.node(translation=[1, -2, 0.5], rotation=[float(v) for v in q], scale=[2, 1, 0.5], children=[n1, n2])
    n3 = g.node(mesh=mC, scale=[-1, 1, 1])        # mirrored: the normal matrix keeps the sign of 1/det
    g.doc["scenes"] = [{"nodes": [n0, n3]}]
    g.doc["scene"] = 0
    g.material(name="red", pbrMetallicRoughness={"baseColorFactor": [0.8, 0.2, 0.1, 1.0], "metallicFactor": 0.3,
                                                 "roughnessFactor": 0.6})
    g.material(name="plain")
    N0 = _trs((1, -2, 0.5), tuple(q), (2, 1, 0.5))
    return g, dict(posA=posA, nrmA=nrmA, uvA=uvA, posB=posB, uvB=uvB, posC=posC, nrmC=nrmC, N0=N0, Mb=Mb,
                   N3=np.diag([-1.0, 1, 1, 1]))


@pytest.mark.parametrize("form", ["glb", "gltf_data_uri", "gltf_external"])
def test_gltf_geometry_nodes_and_placement(tmp_path, form):
    g, e = _geometry_model()
    if form == "glb":
        name = "model.glb"
        (tmp_path / name).write_bytes(g.glb())
    elif form == "gltf_data_uri":
        name = "model.gltf"
        (tmp_path / name).write_bytes(g.gltf_data_uri())
    else:
        name = "model.gltf"
        js, b = g.gltf_external("my%20data.bin")
        (tmp_path / name).write_bytes(js)
        (tmp_path / "my data.bin").write_bytes(b)
    xml = _write_scene(tmp_path, name, 'posX="0.5" posY="-1" posZ="2" rotX="30" rotY="45" rotZ="-60" '
                                       'scaleX="1.5" scaleY="0.5" scaleZ="2"')
    sd = scenes.load_xml(xml, base_dir=str(tmp_path))
    P = _placement((0.5, -1, 2), (30, 45, -60), (1.5, 0.5, 2))
    tris = sd.triangles[10:]                      # after the OBJ meshes' 10 triangles
    assert tris.shape[0] == 2 + 1 + 3 + 3
    WA, WB, WC = P @ e["N0"], P @ e["N0"] @ e["Mb"], P @ e["N3"]
    want = []
    for a, b, c in ((0, 1, 2), (0, 2, 3)):
        want.append(("A", (a, b, c)))
    want.append(("A2", (1, 2, 3)))
    for i in range(3):
        want.append(("B", (i, i + 1, i + 2) if i % 2 == 0 else (i + 1, i, i + 2)))
    for i in range(1, 4):
        want.append(("C", (0, i, i + 1)))
    scale = 10.0
    for t, (kind, idx) in zip(tris, want):
        W = {"A": WA, "A2": WA, "B": WB, "C": WC}[kind]
        pos = e["posC"] if kind == "C" else e["posB"] if kind == "B" else e["posA"]
        for k, vi in enumerate(idx):
            assert np.allclose(t[f"v{k}"], _xp(W, pos[vi]), atol=2e-6 * scale), (kind, k)
        if kind == "A":
            for k, vi in enumerate(idx):
                assert np.allclose(t[f"n{k}"], _xn(W, e["nrmA"][vi]), atol=2e-6)
                assert np.array_equal(t[f"uv{k}"], e["uvA"][vi])
        elif kind == "C":
            for k in range(3):
                assert np.allclose(t[f"n{k}"], _xn(W, e["nrmC"][0]), atol=2e-6)
                assert np.array_equal(t[f"uv{k}"], [0, 0])
        else:
            v = [t["v0"], t["v1"], t["v2"]]
            gn = np.cross(v[1] - v[0], v[2] - v[0])
            gn /= np.linalg.norm(gn)
            assert np.allclose(t["n0"], gn, atol=2e-6) and np.array_equal(t["n0"], t["n2"])
            if kind == "B":
                for k, vi in enumerate(idx):
                    assert np.allclose(t[f"uv{k}"], e["uvB"][vi] / 65535.0, atol=1e-7)
    # materials: 5 from the XML, then red, plain and the default slot (mesh C has none)
    assert list(sd.tri_material_ids[10:]) == [5, 5, 5, 6, 6, 6, 7, 7, 7]
    assert list(sd.tri_bsdf_ids[10:]) == [5, 5, 5, 6, 6, 6, 7, 7, 7]
    b = sd.bsdfs[5:]
    assert list(b["type"]) == [pod.BSDF_DISNEY] * 3
    assert np.allclose(b["p0"][0], [0.8, 0.2, 0.1, 0.6]) and np.allclose(b["p1"][0], [0.3, 0.5, 0, 0])
    assert np.allclose(b["p0"][1], [1, 1, 1, 1]) and np.allclose(b["p1"][1], [1, 0.5, 0, 0])   # glTF defaults
    assert np.allclose(b["p2"][1], [0, 0.5, 0, 0]) and np.allclose(b["p3"][1], [0, 1, 1.5, 0])
    assert np.allclose(b["p0"][2], [0.8, 0.8, 0.8, 0.5]) and np.allclose(b["p2"][2], [0, 0.5, 0, 0])
    assert np.array_equal(sd.extra["materials"][5:], np.ones((3, 3), np.float32))
    assert list(sd.tri_emitter_flags[10:]) == [0] * 9 and sd.emitters.shape[0] == 2     # only the OBJ lights


def _tri_area(t):
    return 0.5 * np.linalg.norm(np.cross(t["v1"].astype(np.float64) - t["v0"], t["v2"].astype(np.float64) - t["v0"]))


def test_gltf_materials_textures_and_emitters(tmp_path):
    g = Gltf()
    rs = np.random.default_rng(9)
    albedo = rs.integers(0, 256, (3, 4, 4))
    mr = rs.integers(0, 256, (5, 2, 4))
    t_alb = g.image(encode_png(albedo, 8, 6))
    t_mr = g.image(encode_png(mr, 8, 6))
    quad = np.array([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0]], F32)
    idx = g.accessor(np.array([0, 1, 2, 0, 2, 3], np.uint16))
    pa = g.accessor(quad)
    g.material(name="tex", pbrMetallicRoughness={"baseColorFactor": [0.5, 1.0, 0.25, 0.8],
                                                 "baseColorTexture": {"index": t_alb},
                                                 "metallicRoughnessTexture": {"index": t_mr},
                                                 "metallicFactor": 0.9, "roughnessFactor": 0.4},
               extensions={"KHR_materials_ior": {"ior": 1.33}, "KHR_materials_transmission": {"transmissionFactor": 0.7},
                           "KHR_materials_clearcoat": {"clearcoatFactor": 0.4, "clearcoatRoughnessFactor": 0.25},
                           "KHR_materials_sheen": {"sheenColorFactor": [0.2, 0.4, 0.6]}})
    g.material(name="lamp", pbrMetallicRoughness={"baseColorFactor": [0.3, 0.3, 0.3, 1]},
               emissiveFactor=[1.0, 0.5, 0.25], extensions={"KHR_materials_emissive_strength": {"emissiveStrength": 4},
                                                            "KHR_materials_unlit": {}})
    g.material(name="unused lamp", emissiveFactor=[2, 2, 2])
    m = g.mesh([{"attributes": {"POSITION": pa}, "indices": idx, "material": 0},
                {"attributes": {"POSITION": g.accessor(quad * F32(0.5) + F32([0, 0, 1]))}, "indices": idx, "material": 1}])
    g.doc["scenes"] = [{"nodes": [g.node(mesh=m)]}]
    (tmp_path / "mat.glb").write_bytes(g.glb())
    sd = scenes.load_xml(_write_scene(tmp_path, "mat.glb", "", with_meshes=False), base_dir=str(tmp_path))
    # only glTF materials (no default slot: every primitive has one)
    assert sd.bsdfs.shape[0] == 3 and sd.triangles.shape[0] == 4
    b = sd.bsdfs
    assert np.allclose(b["p0"][0], [1, 1, 1, 0.4]) and np.allclose(b["p1"][0], [0.9, 0.5, 0.7, 0])
    lum = 0.2126 * 0.2 + 0.7152 * 0.4 + 0.0722 * 0.6
    assert np.allclose(b["p2"][0], [lum, 1.0, 0, 0]) and np.allclose(b["p3"][0], [0.4, 0.75, 1.33, 0])
    assert np.allclose(b["p0"][1], [0.3, 0.3, 0.3, 1.0]) and np.allclose(b["p1"][1], [0, 0.5, 0, 0])     # unlit
    # the factor is baked into the albedo map: uint8((p / 255) f 255 + 0.5) in binary32
    f = np.array([0.5, 1.0, 0.25, 0.8], F32)
    baked = np.clip((albedo.astype(F32) / F32(255)) * f * F32(255) + F32(0.5), 0, 255).astype(np.uint8)
    assert np.array_equal(sd.textures[0], baked) and sd.textures[1] is None and sd.textures[2] is None
    assert np.array_equal(sd.mr_textures[0], mr.astype(np.uint8)) and sd.mr_textures[1] is None
    # emitters: "lamp" over its two triangles; "unused lamp" has none and leaves a lone CDF start
    assert list(sd.tri_emitter_flags) == [0, 0, 1, 1]
    assert np.array_equal(sd.tri_emission[2], [4, 2, 1]) and np.array_equal(sd.tri_emission[0], [0, 0, 0])
    e = sd.emitters
    assert e.shape[0] == 1 and list(sd.emitter_tri_indices) == [2, 3]
    a0, a1 = _tri_area(sd.triangles[2]), _tri_area(sd.triangles[3])
    assert np.allclose(sd.emitter_tri_cdf, [0, a0, a0 + a1, 0]) and np.isclose(e["areaSum"][0], a0 + a1)
    lum_e = 0.2126 * 4 + 0.7152 * 2 + 0.0722 * 1
    assert np.isclose(e["powerWeight"][0], (a0 + a1) * lum_e, rtol=1e-6)
    assert np.allclose(sd.scene_emitter_cdf, [0, (a0 + a1) * lum_e], rtol=1e-6)


def _box(g, center, size, material=None):
    c = np.asarray(center, F32)
    p = np.array([[-1, -1, 0], [1, -1, 0], [0, 1, 0.5]], F32) * F32(size / 2) + c
    prim = {"attributes": {"POSITION": g.accessor(p, minmax=True)}}
    if material is not None:
        prim["material"] = material
    return g.mesh([prim])


def test_gltf_pruning_and_root_rules(tmp_path):
    g = Gltf()
    outline = g.material(name="Toon_OUTLINE")
    nodes = [g.node(mesh=_box(g, (i * 0.3, 0, 0), 1.0), name=f"part{i}") for i in range(5)]
    nodes.append(g.node(mesh=_box(g, (0, 0, 0), 100.0), name="cape"))          # extent > 12 x median
    nodes.append(g.node(mesh=_box(g, (50, 0, 0), 1.0), name="collider"))        # centre > 10 x median away
    nodes.append(g.node(mesh=_box(g, (0, 0, 0), 1.0), name="BackGround"))       # backdrop node
    nodes.append(g.node(mesh=_box(g, (0, 0, 0), 1.0, material=outline)))        # outline material
    g.doc["scenes"] = [{"nodes": nodes}]
    (tmp_path / "p.glb").write_bytes(g.glb())
    sd = scenes.load_xml(_write_scene(tmp_path, "p.glb", "", with_meshes=False), base_dir=str(tmp_path))
    assert sd.triangles.shape[0] == 5
    assert np.allclose(sorted(sd.triangles["v2"][:, 0]), [i * 0.3 for i in range(5)], atol=1e-6)
    # the outline material is unused, so the default slot was appended and kept
    assert sd.bsdfs.shape[0] == 2 and list(sd.tri_material_ids) == [1] * 5
    # without `scenes` every node is a root: a child is reached from its parent and again on its own
    g2 = Gltf()
    child = g2.node(mesh=_box(g2, (0, 0, 0), 1.0))
    g2.node(children=[child], translation=[0, 0, 5])
    (tmp_path / "r.glb").write_bytes(g2.glb())
    sd2 = scenes.load_xml(_write_scene(tmp_path, "r.glb", "", with_meshes=False), base_dir=str(tmp_path))
    assert sd2.triangles.shape[0] == 2
    assert sorted(float(v) for v in sd2.triangles["v0"][:, 2]) == [0.0, 5.0]


def test_gltf_errors(tmp_path):
    def load(name, data):
        (tmp_path / name).write_bytes(data)
        return scenes.load_xml(_write_scene(tmp_path, name, "", with_meshes=False), base_dir=str(tmp_path))
    with pytest.raises(ValueError, match="JSON"):
        load("bad.gltf", b'{"asset": {"version": "2.0"}, "nodes": [}')
    g = Gltf()
    g.image(b"\x89PNG\r\n\x1a\n")                  # an image that does not decode fails the load
    g.node(mesh=_box(g, (0, 0, 0), 1.0))
    with pytest.raises(ValueError, match="image 0"):
        load("img.glb", g.glb())
    g = Gltf()
    g.node(name="empty")
    with pytest.raises(ValueError, match="no triangles"):
        load("none.glb", g.glb())
    with pytest.raises(ValueError, match="cannot read"):
        scenes.load_xml(_write_scene(tmp_path, "missing.glb", "", with_meshes=False), base_dir=str(tmp_path))
    g = Gltf()
    acc = g.accessor(np.zeros((3, 3), F32))
    g.doc["accessors"][acc]["count"] = 1000       # past the end of its buffer: the attribute reads as absent
    g.node(mesh=g.mesh([{"attributes": {"POSITION": acc}}]))
    with pytest.raises(ValueError, match="no triangles"):
        load("short.glb", g.glb())
    # absurd numbers (counts and offsets past any buffer, negative, beyond int) read as absent
    for key, val in (("count", 1e18), ("count", -3), ("byteOffset", -8), ("byteOffset", 1e300)):
        g = Gltf()
        acc = g.accessor(np.zeros((3, 3), F32))
        g.doc["accessors"][acc][key] = val
        g.node(mesh=g.mesh([{"attributes": {"POSITION": acc}}]))
        with pytest.raises(ValueError, match="no triangles"):
            load("absurd.glb", g.glb())
    g = Gltf()
    g.node(mesh=1e20, children=[-1, 2.5e12])
    with pytest.raises(ValueError, match="no triangles"):
        load("index.glb", g.glb())
