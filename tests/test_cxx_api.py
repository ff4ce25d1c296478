"""The C++ API (include/talshxx.hpp, include/talsh_task.hpp): the headers
compile as C++11 and C++14 without warnings and link against libtalsh.so
(CPU), and tests/cxx/talshxx_test -- the closed-form checks of the
reference's C++ test (TALSH/test.cpp:166-293) plus asynchronous polling,
complex data, permuted copy, accumulate, scale, norm1 and the XL form --
passes on the GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cxx", "talshxx_test.cpp")
BIN = os.path.join(ROOT, "tests", "cxx", "talshxx_test")
LIBDIR = os.path.join(ROOT, "exatensor_amd")

needs_gxx = pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")


@needs_gxx
@pytest.mark.parametrize("std", ["c++11", "c++14", "c++17"])
def test_headers_compile(std):
    r = subprocess.run(["g++", f"-std={std}", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), SRC], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@needs_gxx
def test_links_against_library(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libtalsh.so")):
        pytest.skip("libtalsh.so not built")
    out = tmp_path / "t"
    r = subprocess.run(["g++", "-std=c++14", "-I" + os.path.join(ROOT, "include"), SRC, "-L" + LIBDIR, "-ltalsh",
                        "-Wl,--no-undefined", "-o", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_cxx_api_on_gpu():
    if not os.path.exists(BIN):
        pytest.fail("tests/cxx/talshxx_test is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=180)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
