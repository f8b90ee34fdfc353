"""The serving arena's sub-allocator (gpu_recip_amd/csrc/pme_arena.hpp),
host code only: a randomised C++ check of first fit, coalescing, the block
returned by a release and allocation from it (the serving loop's eviction
path), compiled with g++ and run on the CPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_arena_invariants(tmp_path):
    exe = tmp_path / "arena_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "gpu_recip_amd", "csrc"),
                    os.path.join(HERE, "cpp", "arena_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
