"""The BVH4 and BVH8 the kernels traverse (the BVH2 collapsed by surface-area cost) are
structurally sound: tests/cpp/bvh4_check.cpp, compiled here with g++ against
the library's own bte_bvh.cpp, checks that every triangle is referenced once,
that every box contains its subtree, and that the depth stays within the
bound the traversal stack is sized for.  The GPU parity tests then show that
traversing them gives the same hits as the BVH2 and the oracle."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "bluetruthengine_amd", "csrc")


def test_bvh4_structure(tmp_path):
    exe = str(tmp_path / "bvh4_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-I", CSRC,
                    os.path.join(ROOT, "tests", "cpp", "bvh4_check.cpp"), os.path.join(CSRC, "bte_bvh.cpp"),
                    "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(out.stdout, out.stderr)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.count("BVH4") == 6 and out.stdout.count("BVH8") == 6
