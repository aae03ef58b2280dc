"""Host-side allocator shared by the live server's output arena and the
builder's workspace (sparse-voxel-octrees_amd/csrc/range_alloc.hpp),
compiled with g++ and checked against a brute-force model (CPU only)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_range_allocator_is_first_fit_and_merges(tmp_path):
    exe = str(tmp_path / "range_alloc_test")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "sparse-voxel-octrees_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "range_alloc_test.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe, "20000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok 20000")
