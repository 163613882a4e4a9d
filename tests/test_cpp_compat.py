"""The reference-named C++ host layer (include/bte_render_kernel.hpp) compiles
and links against libbte.so; on a GPU it renders an emissive triangle whose
directly seen pixels carry exactly its emission."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "compat_triangle.cpp")
LIBDIR = os.path.join(ROOT, "bluetruthengine_amd", "_lib")


def _build(tmp_path):
    exe = str(tmp_path / "compat_triangle")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), SRC, "-L", LIBDIR, "-lbte",
                    f"-Wl,-rpath,{LIBDIR}", "-o", exe], check=True)
    return exe


def test_compat_header_compiles_and_links(tmp_path):
    assert os.path.exists(_build(tmp_path))


@pytest.mark.gpu
def test_compat_program_renders(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0 and "lit=" in out.stdout
