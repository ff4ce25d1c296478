"""The device arena's allocator (exatensor_amd/csrc/arena.cpp) on the host:
tests/cxx/arena_test.cpp is compiled with g++ against the engine's own
source and run (TRY_LATER / DEVICE_UNABLE leave it unchanged, best fit,
fragmentation, coalescing, alignment, and 20000 random allocate/release
steps against a shadow model). The allocator only does bookkeeping over the
range it is given, so a host buffer stands in for device memory."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_arena_allocator_contract(tmp_path):
    exe = str(tmp_path / "arena_test")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "exatensor_amd", "csrc"), os.path.join(ROOT, "tests", "cxx", "arena_test.cpp"),
                    os.path.join(ROOT, "exatensor_amd", "csrc", "arena.cpp"), "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "ARENA_OK" in r.stdout
