"""A scene.xml / OBJ / .hdr scene loaded by libbte's host layer renders on the
GPU bit-identically to the oracle, and the library's own upload
(bte_scene_upload, the reference's init order) equals the Python upload.
The scene files are written by the test (no reference files on the GPU box).
"""
import ctypes

import numpy as np
import pytest

import oracle
from bluetruthengine_amd import renderer, scenes
from test_scene_loader import _synthetic

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _add_jpeg_textures(tmp_path):
    """Baseline JPEG albedo maps on the floor quad (4:4:4) and on the
    microfacet triangles (4:2:0), written with the test encoder."""
    from test_jpeg_decoder import QUANT, _image, _ycc, encode
    y, cb, cr = _ycc(_image(32, 48, 2))
    (tmp_path / "floor.jpg").write_bytes(encode([y, cb, cr], [(1, 1)] * 3, QUANT)[0])
    y, cb, cr = _ycc(_image(16, 16, 3))
    half = lambda p: p.reshape(8, 2, 8, 2).mean((1, 3))
    (tmp_path / "mf.jpg").write_bytes(encode([y, half(cb), half(cr)], [(2, 2), (1, 1), (1, 1)], QUANT)[0])
    xml = (tmp_path / "scene.xml").read_text().replace("missing.png", "floor.jpg")
    xml = xml.replace('name="mMf" bsdf', 'name="mMf" albedoTexture="mf.jpg" bsdf')
    (tmp_path / "scene.xml").write_text(xml)


def _add_gltf_model(tmp_path):
    """A GLB model in view: a quad with PNG albedo and metallic-roughness maps
    (texture coordinates past [0, 1]), a triangle strip with only a
    metallic-roughness map, and an emissive pair of triangles."""
    from gltf_writer import Gltf
    from test_png_decoder import encode_png
    rs = np.random.default_rng(21)
    g = Gltf()
    t_alb = g.image(encode_png(rs.integers(0, 256, (8, 8, 4)), 8, 6))
    t_mr = g.image(encode_png(rs.integers(0, 256, (4, 4, 4)), 8, 6))
    t_mr2 = g.image(encode_png(rs.integers(0, 256, (2, 3, 4)), 8, 6))
    g.material(name="tex", pbrMetallicRoughness={"baseColorFactor": [0.9, 0.8, 0.7, 1.0], "metallicFactor": 0.8,
                                                 "roughnessFactor": 0.7, "baseColorTexture": {"index": t_alb},
                                                 "metallicRoughnessTexture": {"index": t_mr}})
    g.material(name="mronly", pbrMetallicRoughness={"baseColorFactor": [0.2, 0.6, 0.9, 1.0],
                                                    "metallicRoughnessTexture": {"index": t_mr2}})
    g.material(name="glow", emissiveFactor=[3.0, 3.0, 2.0])
    f32 = np.float32
    quad = np.array([[-0.6, 0.2, -0.4], [0.2, 0.2, -0.6], [0.2, 1.0, -0.6], [-0.6, 1.0, -0.4]], f32)
    uv = np.array([[-0.5, -0.5], [1.5, -0.5], [1.5, 1.5], [-0.5, 1.5]], f32)
    idx = g.accessor(np.array([0, 1, 2, 0, 2, 3], np.uint16))
    strip = np.array([[0.3, 0.1, -0.2], [0.7, 0.1, -0.3], [0.3, 0.6, -0.2], [0.7, 0.6, -0.4], [0.4, 1.1, -0.3],
                      [0.8, 1.0, -0.5]], f32)
    suv = np.array([[0, 0], [1, 0], [0, 1], [1, 1], [0, 2], [1, 2]], f32)
    lamp = np.array([[-0.2, 1.6, -0.2], [0.2, 1.6, -0.2], [0.2, 1.6, 0.2], [-0.2, 1.6, 0.2]], f32)
    m = g.mesh([{"attributes": {"POSITION": g.accessor(quad, minmax=True), "TEXCOORD_0": g.accessor(uv)},
                 "indices": idx, "material": 0},
                {"attributes": {"POSITION": g.accessor(strip, minmax=True), "TEXCOORD_0": g.accessor(suv)},
                 "mode": 5, "material": 1},
                {"attributes": {"POSITION": g.accessor(lamp, minmax=True)}, "indices": idx, "material": 2}])
    g.doc["scenes"] = [{"nodes": [g.node(mesh=m, translation=[0.0, 0.05, 0.0])]}]
    (tmp_path / "model.glb").write_bytes(g.glb())
    xml = (tmp_path / "scene.xml").read_text().replace("</scene>", '<model filename="model.glb" rotY="10" />\n</scene>')
    (tmp_path / "scene.xml").write_text(xml)


@pytest.mark.parametrize("variant", ["plain", "jpeg", "gltf"])
def test_loaded_scene_matches_oracle_and_library_upload(tmp_path, variant):
    extra = {"plain": "", "jpeg": ", JPEG albedo maps", "gltf": ", GLB model with PNG albedo/MR maps and an emitter"}
    _say(f"case: synthetic scene.xml (OBJ quads/tris, emitters, spotlight, RLE .hdr sky{extra[variant]}), "
         f"64x48, 3 frames, depth 6")
    _synthetic(tmp_path, rle=True)
    if variant == "jpeg":
        _add_jpeg_textures(tmp_path)
    elif variant == "gltf":
        _add_gltf_model(tmp_path)
    xml = str(tmp_path / "scene.xml")
    sd = scenes.load_xml(xml, base_dir=str(tmp_path), max_depth=6)
    assert (sd.textures is not None) == (variant != "plain")
    if variant == "gltf":
        assert sd.emitters.shape[0] == 3 and sd.mr_textures is not None and sd.triangles.shape[0] == 10 + 2 + 4 + 2
    r = renderer.Renderer(sd)
    r.set_hit_capture(3)
    r.render(3)
    gpu = r.sample_radiance(3)
    acc_py = r.accum().copy()
    env_cdf = r.env_cdf()
    ref = oracle.OracleScene(sd, env_cdf=env_cdf)
    pix = np.arange(sd.width * sd.height, dtype=np.int32)
    for f in range(3):
        cpu, _ = ref.trace_samples(pix, sd.width, sd.height, f, sd.max_depth)
        exact = np.mean(np.all(gpu[f].reshape(-1, 3) == cpu, axis=1))
        _say(f"frame {f}: bit-exact samples {exact:.6f}")
        assert exact == 1.0
    # the same scene through bte_scene_upload
    L = renderer.lib()
    renderer.cudaCleanup()
    h = ctypes.c_void_p()
    assert L.bte_scene_load_xml(xml.encode(), str(tmp_path).encode(), ctypes.byref(h)) == 0
    try:
        assert L.bte_scene_upload(h) == 0
    finally:
        L.bte_scene_free(h)
    cfg = renderer.BteConfig(6, 5, 0, 0, 16, 1, 4, 0, 2)
    renderer._call("bte_set_config", ctypes.byref(cfg))
    renderer._call("bte_reset_accumulation", sd.width, sd.height)
    renderer._call("bte_render_frames", sd.width, sd.height, 3)
    acc = np.empty_like(acc_py)
    renderer._call("bte_read_accum", acc.ctypes.data_as(ctypes.c_void_p), acc.size)
    assert np.array_equal(acc, acc_py)
