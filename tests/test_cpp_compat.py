"""The reference-named C++ host layer (include/bte_render_kernel.hpp) compiles
and links against libbte.so; on a GPU it renders an emissive triangle whose
directly seen pixels carry exactly its emission, and a scene.xml loaded through
loadSceneXml renders to the same pixels as the Python path."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "compat_triangle.cpp")
LIBDIR = os.path.join(ROOT, "bluetruthengine_amd", "_lib")


def _build(tmp_path, src=SRC):
    exe = str(tmp_path / os.path.splitext(os.path.basename(src))[0])
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), src, "-L", LIBDIR, "-lbte",
                    f"-Wl,-rpath,{LIBDIR}", "-o", exe], check=True)
    return exe


SCENE_SRC = os.path.join(ROOT, "tests", "cpp", "scene_render.cpp")


def test_compat_header_compiles_and_links(tmp_path):
    assert os.path.exists(_build(tmp_path))
    assert os.path.exists(_build(tmp_path, SCENE_SRC))


@pytest.mark.gpu
def test_compat_program_renders(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0 and "lit=" in out.stdout


@pytest.mark.gpu
def test_scene_xml_program_matches_python_path(tmp_path):
    import numpy as np
    from bluetruthengine_amd import renderer, scenes
    from test_scene_loader import _synthetic
    _synthetic(tmp_path)
    exe = _build(tmp_path, SCENE_SRC)
    out = subprocess.run([exe, "scene.xml", "2", "out.ppm"], capture_output=True, text=True, timeout=120,
                         cwd=str(tmp_path))
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    data = open(tmp_path / "out.ppm", "rb").read()
    hdr = b"P6\n64 48\n255\n"
    assert data.startswith(hdr)
    img = np.frombuffer(data[len(hdr):], np.uint8).reshape(48, 64, 3)[::-1]
    sd = scenes.load_xml(str(tmp_path / "scene.xml"), base_dir=str(tmp_path), max_depth=200)
    r = renderer.Renderer(sd)
    r.render(2)
    fb = r.framebuffer()
    rgb = np.stack([fb & 255, (fb >> 8) & 255, (fb >> 16) & 255], -1).astype(np.uint8)
    assert np.array_equal(img, rgb)
