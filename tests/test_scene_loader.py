"""The scene.xml / OBJ / .hdr host layer (libbte's bte_scene_*), on the CPU.

Expected values are recomputed here from the files' text with the reference's
rules (src/scene.cpp:83-523, src/main.cpp:267-401 and 524-930,
include/mesh_transform.h:9-80): face counts and fan triangulation, index forms
and negative indices, Scale->Rx->Ry->Rz->Translate transforms, normal and
planar-UV fallbacks, BSDF packing per type, emitter areas and CDFs, the camera
basis, spotlight cones, and RGBE decoding (flat and run-length scanlines),
plus the reference's own asset files when /root/reference is present.
"""
import math
import os
import re
import struct

import numpy as np
import pytest

from bluetruthengine_amd import pod, scenes

REF_ASSETS = "/root/reference/assets"


def _rgbe(rgb):
    """float RGB -> RGBE bytes (the encoder side, used to write test files)"""
    m = max(rgb)
    if m < 1e-32:
        return bytes([0, 0, 0, 0])
    f, e = math.frexp(m)
    s = f * 256.0 / m
    return bytes([int(c * s) for c in rgb] + [e + 128])


def _decode_rgbe(b):
    if b[3] == 0:
        return (0.0, 0.0, 0.0)
    f = math.ldexp(1.0, b[3] - 136)
    return tuple(np.float32(b[k]) * np.float32(f) for k in range(3))


def _write_hdr(path, pixels_rgbe, w, h, rle):
    out = bytearray(b"#?RADIANCE\n# test\nFORMAT=32-bit_rle_rgbe\n\n")
    out += f"-Y {h} +X {w}\n".encode()
    for y in range(h):
        row = pixels_rgbe[y * w:(y + 1) * w]
        if not rle:
            for px in row:
                out += px
            continue
        out += bytes([2, 2, w >> 8, w & 255])
        for k in range(4):
            comp = [px[k] for px in row]
            x = 0
            while x < w:
                # a run when the next 3+ bytes repeat, else literals
                run = 1
                while x + run < w and comp[x + run] == comp[x] and run < 127:
                    run += 1
                if run >= 3:
                    out += bytes([128 + run, comp[x]])
                    x += run
                else:
                    n = min(w - x, 2)
                    out += bytes([n] + comp[x:x + n])
                    x += n
    open(path, "wb").write(out)


def _synthetic(tmp_path, rle=True):
    (tmp_path / "quad.obj").write_text(
        "# quad as one polygon, no normals, one texcoord\n"
        "v -1 0 -1\nv 1 0 -1\nv 1 0 1\nv -1 0 1\nvt 0.5 0.5\nf 1/1 2/1 3/1 4/1\n")
    (tmp_path / "tri.obj").write_text(
        "v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 1\nvn 0 0.6 0.8\nvt 0 0\nvt 1 0\nvt 0 1\n"
        "f -3/-3/-2 -2/-2/-2 -1/-1/-1\nf 1//1 2//1 3//1\n")
    (tmp_path / "light.obj").write_text("v -0.2 1.9 -0.2\nv 0.2 1.9 -0.2\nv 0.2 1.9 0.2\nv -0.2 1.9 0.2\n"
                                        "f 1 3 2\nf 1 4 3\n")
    w, h = 12, 3
    rs = np.random.default_rng(0)
    px = [_rgbe(tuple(rs.uniform(0, 4, 3))) for _ in range(w * h)]
    px[5] = bytes([0, 0, 0, 0])
    _write_hdr(tmp_path / "sky.hdr", px, w, h, rle)
    xml = """<scene>
  <environment map="sky.hdr" intensity="2.5" rotation="30" />
  <bsdf name="white" type="diffuse" albedoR="0.7" albedoG="0.6" albedoB="0.5" />
  <bsdf name="mir" type="mirror" albedoR="0" albedoG="0.5" />
  <bsdf name="mf" type="microfacet" albedoR="0.2" albedoG="0.3" albedoB="0.1" alpha="0.25" intIOR="1.7" />
  <bsdf name="glass" type="dielectric" intIOR="1.45" />
  <bsdf name="dis" type="disney" base_colorR="0.9" roughness="0.3" metallic="1" clearcoat="0.5" eta="1.6" />
  <material name="mWhite" albedoTexture="missing.png" bsdf="white" />
  <material name="mMir" bsdf="mir" />
  <material name="mMf" bsdf="mf" />
  <material name="mDis" bsdf="dis" />
  <material name="mGlass" bsdf="glass" />
  <mesh filename="quad.obj" material="mWhite" />
  <mesh filename="tri.obj" material="mMf" scaleX="2" rotX="30" rotY="45" rotZ="-60" posX="0.5" posY="0.25" />
  <mesh filename="tri.obj" material="mDis" posZ="-0.5" />
  <mesh filename="light.obj" material="mWhite" isEmitter="true" />
  <mesh filename="light.obj" material="mMir" isEmitter="1" radianceR="5" radianceG="6" radianceB="7" posY="-0.1" />
  <spotlight posX="0" posY="1.5" posZ="0" dirX="0" dirY="-2" dirZ="0.5" radianceR="3" intensity="4"
             innerConeAngle="20" outerConeAngle="35" />
  <window width="64" height="48" />
  <camera eyeX="0" eyeY="1" eyeZ="3" lookAtX="0.1" lookAtY="0.9" lookAtZ="0" upX="0" upY="1" upZ="0" fovY="40" />
</scene>
"""
    (tmp_path / "scene.xml").write_text(xml)
    return px, w, h


def _xform(p, sx=1, sy=1, sz=1, rx=0, ry=0, rz=0, tx=0, ty=0, tz=0, normal=False):
    f = np.float32
    d2r = f(3.14159265358979323846) / f(180)
    x, y, z = (f(v) for v in p)
    if not normal:
        x, y, z = x * f(sx), y * f(sy), z * f(sz)
    for axis, ang in (("x", rx), ("y", ry), ("z", rz)):
        a = f(ang) * d2r
        c, s = f(math.cos(a)), f(math.sin(a))
        if axis == "x":
            y, z = y * c - z * s, y * s + z * c
        elif axis == "y":
            x, z = x * c + z * s, -x * s + z * c
        else:
            x, y = x * c - y * s, x * s + y * c
    if normal:
        n = math.sqrt(x * x + y * y + z * z)
        return np.array([x / n, y / n, z / n])
    return np.array([x + f(tx), y + f(ty), z + f(tz)])


@pytest.mark.parametrize("rle", [True, False])
def test_synthetic_scene_follows_the_reference_rules(tmp_path, rle):
    px, ew, eh = _synthetic(tmp_path, rle)
    sd = scenes.load_xml(str(tmp_path / "scene.xml"), base_dir=str(tmp_path))
    assert (sd.width, sd.height) == (64, 48)
    tris = sd.triangles
    # quad: one polygon fanned into 2; tri.obj twice (2 faces); light twice (2 faces)
    assert tris.shape[0] == 2 + 2 + 2 + 2 + 2
    q = tris[:2]
    assert np.array_equal(q["v0"][0], [-1, 0, -1]) and np.array_equal(q["v1"][0], [1, 0, -1])
    assert np.array_equal(q["v2"][0], [1, 0, 1]) and np.array_equal(q["v2"][1], [-1, 0, 1])
    # no normals: geometric normal e1 x e2 normalised, for all three vertices
    assert np.allclose(q["n0"], [[0, -1, 0], [0, -1, 0]]) and np.array_equal(q["n0"], q["n2"])
    # one texcoord only: planar UV from the dominant normal axis (y -> (x, z))
    assert np.array_equal(q["uv1"][0], [1, -1])
    # transformed triangle with negative indices and v/vt/vn
    t0 = tris[2]
    kw = dict(sx=2, rx=30, ry=45, rz=-60, tx=0.5, ty=0.25)
    for k, p in (("v0", (0, 0, 0)), ("v1", (1, 0, 0)), ("v2", (0, 1, 0))):
        assert np.allclose(t0[k], _xform(p, **kw), atol=2e-7), k
    # vn index -2 is the first normal, -1 the second
    assert np.allclose(t0["n0"], _xform((0, 0, 1), normal=True, **kw), atol=2e-7)
    assert np.allclose(t0["n2"], _xform((0, 0.6, 0.8), normal=True, **kw), atol=2e-7)
    assert np.array_equal(t0["uv1"], [1, 0])
    # BSDF packing
    b = sd.bsdfs
    assert list(b["type"]) == [pod.BSDF_DIFFUSE, pod.BSDF_MIRROR, pod.BSDF_MICROFACET, pod.BSDF_DIELECTRIC,
                               pod.BSDF_DISNEY]
    assert np.allclose(b["p0"][0], [0.7, 0.6, 0.5, 0])
    assert np.array_equal(b["p0"][1], [1, 0.5, 1, 0])                   # non-positive -> 1
    assert np.allclose(b["p1"][2], [0.25, 1.7, 1.0, 1 - 0.3])           # ks = 1 - max(albedo)
    assert np.allclose(b["p1"][3], [1.45, 1.0, 0, 0])
    assert np.allclose(b["p0"][4], [0.9, 0.5, 0.5, 0.3]) and np.allclose(b["p1"][4], [1, 0.5, 0, 0])
    assert np.allclose(b["p2"][4], [0, 0.5, 0, 0]) and np.allclose(b["p3"][4], [0.5, 1, 1.6, 0])
    assert list(sd.tri_bsdf_ids) == [0, 0, 2, 2, 4, 4, 0, 0, 1, 1]
    assert list(sd.tri_material_ids) == [0, 0, 2, 2, 3, 3, 0, 0, 1, 1]
    assert np.array_equal(sd.extra["materials"], np.ones((5, 3), np.float32))   # texture missing -> white
    # emitters: default radiance 10, areas, CDFs, power weights
    assert list(sd.tri_emitter_flags) == [0] * 6 + [1] * 4
    assert np.array_equal(sd.tri_emission[6], [10, 10, 10]) and np.array_equal(sd.tri_emission[9], [5, 6, 7])
    e = sd.emitters
    assert list(e["triIndexOffset"]) == [0, 2] and list(e["triCount"]) == [2, 2] and list(e["cdfOffset"]) == [0, 3]
    assert list(sd.emitter_tri_indices) == [6, 7, 8, 9]
    area = np.float32(0.5) * np.float32(0.16)    # each light triangle: 0.4 x 0.4 / 2
    assert np.allclose(sd.emitter_tri_cdf, [0, area, 2 * area, 0, area, 2 * area])
    lum2 = 0.2126 * 5 + 0.7152 * 6 + 0.0722 * 7
    assert np.allclose(e["powerWeight"], [2 * area * 10, 2 * area * lum2], rtol=1e-6)
    assert np.allclose(sd.scene_emitter_cdf, [0, 2 * area * 10, 2 * area * (10 + lum2)], rtol=1e-6)
    # spotlight: normalised direction, cone cosines
    s = sd.spotlights[0]
    d = np.array([0, -2, 0.5]) / math.sqrt(4.25)
    assert np.allclose(s["direction"], d, atol=1e-7)
    assert np.isclose(s["innerConeCosine"], math.cos(math.radians(20)), atol=1e-7)
    assert np.isclose(s["outerConeCosine"], math.cos(math.radians(35)), atol=1e-7)
    assert s["intensity"] == 4 and np.array_equal(s["radiance"], [3, 1, 1])
    # camera basis: forward, right = f x up, up' = right x f; aspect w/h
    c = sd.camera
    f = np.array([0.1, -0.1, -3.0])
    f /= np.linalg.norm(f)
    r = np.cross(f, [0, 1, 0])
    r /= np.linalg.norm(r)
    assert np.allclose(c["forward"], f, atol=1e-7) and np.allclose(c["right"], r, atol=1e-7)
    assert np.allclose(c["up"], np.cross(r, f), atol=1e-7)
    assert np.isclose(c["fovYRadians"], math.radians(40)) and np.isclose(c["aspect"], 64 / 48)
    # environment: RGBE decode, alpha 1, intensity / rotation attributes
    env = sd.env_rgba
    assert env.shape == (eh, ew, 4) and sd.env_intensity == 2.5 and sd.env_rotation_deg == 30
    want = np.array([_decode_rgbe(p) for p in px], np.float32).reshape(eh, ew, 3)
    assert np.array_equal(env[..., :3], want) and np.all(env[..., 3] == 1)


def test_loader_errors_are_reported(tmp_path):
    _synthetic(tmp_path)
    bad = (tmp_path / "scene.xml").read_text().replace('material="mMf"', 'material="nope"')
    (tmp_path / "bad.xml").write_text(bad)
    with pytest.raises(ValueError, match="nope"):
        scenes.load_xml(str(tmp_path / "bad.xml"), base_dir=str(tmp_path))
    # a texture file that exists but cannot be decoded is an error (a missing
    # one is no texture, as in the reference)
    (tmp_path / "tex.png").write_bytes(b"\x89PNG\r\n\x1a\n")
    tex = (tmp_path / "scene.xml").read_text().replace("missing.png", "tex.png")
    (tmp_path / "tex.xml").write_text(tex)
    with pytest.raises(ValueError, match="tex.png.*PNG without IHDR"):
        scenes.load_xml(str(tmp_path / "tex.xml"), base_dir=str(tmp_path))
    (tmp_path / "tex.bmp").write_bytes(b"BM" + bytes(60))
    (tmp_path / "bmp.xml").write_text(tex.replace("tex.png", "tex.bmp"))
    with pytest.raises(ValueError, match="unsupported image format"):
        scenes.load_xml(str(tmp_path / "bmp.xml"), base_dir=str(tmp_path))
    with pytest.raises(ValueError, match="unreadable"):
        scenes.load_xml(str(tmp_path / "none.xml"))


def test_png_albedo_texture(tmp_path):
    from test_png_decoder import encode_png
    _synthetic(tmp_path)
    px = np.random.default_rng(6).integers(0, 256, (5, 7, 4))
    (tmp_path / "wall.png").write_bytes(encode_png(px, 8, 6))
    xml = (tmp_path / "scene.xml").read_text().replace("missing.png", "wall.png")
    (tmp_path / "p.xml").write_text(xml)
    sd = scenes.load_xml(str(tmp_path / "p.xml"), base_dir=str(tmp_path))
    assert np.array_equal(sd.textures[0], px.astype(np.uint8)) and all(t is None for t in sd.textures[1:])
    avg = (px[..., :3] / 255.0).reshape(-1, 3).mean(0)
    assert np.allclose(sd.extra["materials"][0], avg, atol=1e-7)


def test_jpeg_albedo_texture_is_decoded_and_averaged(tmp_path):
    from test_jpeg_decoder import QUANT, _image, _to_rgb, _ycc, encode
    _synthetic(tmp_path)
    H, W = 24, 40
    y, cb, cr = _ycc(_image(H, W, 5))
    data, recon = encode([y, cb, cr], [(1, 1)] * 3, QUANT)
    (tmp_path / "wall.jpg").write_bytes(data)
    xml = (tmp_path / "scene.xml").read_text().replace('name="mMf" bsdf', 'name="mMf" albedoTexture="wall.jpg" bsdf')
    (tmp_path / "t.xml").write_text(xml)
    sd = scenes.load_xml(str(tmp_path / "t.xml"), base_dir=str(tmp_path))
    # one entry per material, only mMf (material 2) has an image
    assert len(sd.textures) == 5 and [t is None for t in sd.textures] == [True, True, False, True, True]
    img = sd.textures[2]
    want = _to_rgb(recon, [(1, 1)] * 3, H, W)
    assert img.shape == (H, W, 4) and np.all(img[..., 3] == 255)
    assert np.abs(img[..., :3].astype(np.int64) - want.astype(np.int64)).max() <= 1
    # loadTextureAverage (src/main.cpp:407-440): mean of byte/255 per channel
    mats = sd.extra["materials"]
    avg = (img[..., :3].astype(np.float64) / 255.0).reshape(-1, 3).mean(0)
    assert np.allclose(mats[2], avg.astype(np.float32), rtol=0, atol=2e-7)
    assert np.array_equal(np.delete(mats, 2, 0), np.ones((4, 3), np.float32))


@pytest.mark.skipif(not os.path.isfile(os.path.join(REF_ASSETS, "checkerboard.jpg")),
                    reason="reference assets not present")
def test_reference_scene_xml_loads_its_floor_texture(tmp_path):
    # assets/scene.xml without its <model> entries: the .glb files they name
    # are not shipped (the reference exits on them, src/gltf_loader.cpp:784-791)
    xml = re.sub(r"<model\b[^>]*/>", "", open(os.path.join(REF_ASSETS, "scene.xml")).read())
    (tmp_path / "scene.xml").write_text(xml)
    sd = scenes.load_xml(str(tmp_path / "scene.xml"), base_dir=os.path.dirname(REF_ASSETS))
    mats = re.findall(r'<material name="([^"]+)"([^>]*)>', xml)
    textured = [i for i, (_, attrs) in enumerate(mats) if "albedoTexture" in attrs]
    assert textured and sd.textures is not None
    for i, t in enumerate(sd.textures):
        assert (t is not None) == (i in textured)
    img = sd.textures[textured[0]]
    assert img.shape == (2048, 2048, 4)
    avg = (img[..., :3].astype(np.float64) / 255.0).reshape(-1, 3).mean(0)
    assert np.allclose(sd.extra["materials"][textured[0]], avg, atol=1e-6)


def _obj_faces(path):
    n = 0
    for line in open(path):
        if line.startswith("f "):
            n += len(line.split()) - 3
    return n


@pytest.mark.skipif(not os.path.isdir(REF_ASSETS), reason="reference assets not present")
def test_reference_scene_files_load(monkeypatch):
    monkeypatch.chdir(os.path.dirname(REF_ASSETS))      # the reference resolves paths from its repo root
    xml_path = "assets/scene_anil.xml"
    sd = scenes.load_xml(xml_path)
    xml = open(xml_path).read()
    meshes = re.findall(r'<mesh filename="([^"]+)"', xml)
    assert sd.triangles.shape[0] == sum(_obj_faces(m) for m in meshes)
    bsdf_names = re.findall(r'<bsdf name="([^"]+)"', xml)
    assert sd.bsdfs.shape[0] == len(bsdf_names)
    assert (sd.width, sd.height) == tuple(int(v) for v in re.search(r'<window width="(\d+)" height="(\d+)"', xml).groups())
    # the one emitter is light.obj with radiance 18: area from its vertices
    v = np.array([[float(x) for x in l.split()[1:4]] for l in open("assets/light.obj") if l.startswith("v ")],
                 np.float32)
    e = sd.emitters
    assert e.shape[0] == 1 and np.array_equal(e["radiance"][0], [18, 18, 18])
    faces = [[int(t.split("/")[0]) - 1 for t in l.split()[1:]] for l in open("assets/light.obj") if l.startswith("f ")]
    area = sum(0.5 * np.linalg.norm(np.cross(v[f[1]] - v[f[0]], v[f[2]] - v[f[0]])) for f in faces)
    assert np.isclose(e["areaSum"][0], area, rtol=1e-6)
    assert np.all(np.isfinite(sd.triangles["n0"])) and np.allclose(np.linalg.norm(sd.triangles["n0"], axis=1), 1,
                                                                     atol=1e-5)


@pytest.mark.skipif(not os.path.isfile(os.path.join(REF_ASSETS, "sky.hdr")), reason="reference sky.hdr not present")
def test_reference_sky_hdr_decodes(tmp_path):
    # a scene pointing at the reference's 4096x2048 sky.hdr (run-length encoded)
    (tmp_path / "t.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    (tmp_path / "s.xml").write_text(
        f'<scene><environment map="{REF_ASSETS}/sky.hdr" /><bsdf name="w" type="diffuse" />'
        '<material name="m" bsdf="w" /><mesh filename="t.obj" material="m" /><window width="8" height="8" />'
        '<camera eyeX="0" eyeY="0" eyeZ="3" lookAtX="0" lookAtY="0" lookAtZ="0" upX="0" upY="1" upZ="0" fovY="40" />'
        "</scene>")
    sd = scenes.load_xml(str(tmp_path / "s.xml"), base_dir=str(tmp_path))
    data = open(os.path.join(REF_ASSETS, "sky.hdr"), "rb").read()
    hdr_end = data.index(b"\n\n") + 2
    res_end = data.index(b"\n", hdr_end)
    _, h, _, w = data[hdr_end:res_end].split()
    h, w = int(h), int(w)
    assert sd.env_rgba.shape == (h, w, 4)
    # independent decode of the first two scanlines
    p = res_end + 1
    for y in range(2):
        assert data[p] == 2 and data[p + 1] == 2
        p += 4
        comps = []
        for k in range(4):
            out = []
            while len(out) < w:
                c = data[p]
                p += 1
                if c > 128:
                    out += [data[p]] * (c - 128)
                    p += 1
                else:
                    out += list(data[p:p + c])
                    p += c
            comps.append(out)
        row = np.array([_decode_rgbe((comps[0][x], comps[1][x], comps[2][x], comps[3][x])) for x in range(w)],
                       np.float32)
        assert np.array_equal(sd.env_rgba[y, :, :3], row)
