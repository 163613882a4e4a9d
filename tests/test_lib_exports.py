"""The C-ABI library loads without a GPU and exports every entry point that
include/bte_api.h declares; the Python mirror binds each of them; calls fail
with a status (never exit) when no device or scene is present."""
import os
import re
import subprocess

from bluetruthengine_amd import renderer

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "bte_api.h")


def declared():
    src = open(HDR).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|long long|const char\*)\s+(bte_\w+)\s*\(", src, re.M)))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", renderer.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(bte_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert set(names) == set(renderer.API)


def test_library_loads_and_reports_errors_without_device():
    L = renderer.lib()
    assert L.bte_get_render_mode() in (0, 1)
    import ctypes
    cfg = renderer.BteConfig(0, 5, 0, 0, 0)
    assert L.bte_set_config(ctypes.byref(cfg)) == -1        # invalid maxDepth
    assert b"maxDepth" in L.bte_last_error()


def test_config_validation_without_device():
    import ctypes
    L = renderer.lib()
    saved = renderer.BteConfig()
    assert L.bte_get_config(ctypes.byref(saved)) == 0
    assert (saved.graphs, saved.workChunk, saved.traversalBlocksPerCU, saved.bvhWidth) == (1, 4, 0, 4)
    assert renderer.default_config()["tailBounce"] == 6
    for fields, word in [(dict(graphs=2), b"graphs"), (dict(workChunk=3), b"workChunk"),
                         (dict(traversalBlocksPerCU=-1), b"traversalBlocksPerCU"),
                         (dict(traversalStack=12), b"traversalStack"), (dict(bvhWidth=3), b"bvhWidth"),
                         (dict(tailBounce=-1), b"tailBounce")]:
        cfg = renderer.BteConfig(12, 5, 0, 0, 16, 1, 4, 0, 2)
        for k, v in fields.items():
            setattr(cfg, k, v)
        assert L.bte_set_config(ctypes.byref(cfg)) == -1
        assert word in L.bte_last_error()
    good = renderer.BteConfig(12, 5, 0, 0, 8, 0, 2, 3, 4, 5)
    assert L.bte_set_config(ctypes.byref(good)) == 0
    back = renderer.BteConfig()
    L.bte_get_config(ctypes.byref(back))
    assert (back.traversalStack, back.graphs, back.workChunk, back.traversalBlocksPerCU, back.bvhWidth,
            back.tailBounce) == (8, 0, 2, 3, 4, 5)
    assert L.bte_set_config(ctypes.byref(saved)) == 0


def test_weights_file_reader_rejects_bad_files(tmp_path):
    L = renderer.lib()
    bad = tmp_path / "bad.npy"
    bad.write_bytes(b"not a numpy file")
    assert L.bte_denoiser_load_file(str(bad).encode()) == -1
    assert b"not a .npy" in L.bte_last_error()
    import numpy as np
    wrong = tmp_path / "wrong.npy"
    np.save(wrong, np.zeros(10, np.float64))
    assert L.bte_denoiser_load_file(str(wrong).encode()) == -1
    assert b"float32" in L.bte_last_error()
    short = tmp_path / "short.npy"
    np.save(short, np.zeros(10, np.float32))
    assert L.bte_denoiser_load_file(str(short).encode()) == -1
    assert b"exactly" in L.bte_last_error()
    assert L.bte_denoiser_param_count() == 1211907
