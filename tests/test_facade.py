"""The C++ drop-in facade (include/SparseGridGpu.hpp): the reference's own tests restated in
C++ (tests/cpp/test_facade.cpp) compile with plain g++ against the C ABI, and pass on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "openfpm_data_amd")


def _build(tmp_path):
    exe = str(tmp_path / "test_facade")
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror",
                           "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "test_facade.cpp"),
                           "-L", LIBDIR, "-lsg_hip", "-Wl,-rpath," + LIBDIR, "-o", exe])
    return exe


def test_facade_compiles(tmp_path):
    assert os.path.exists(_build(tmp_path))


@pytest.mark.gpu
def test_facade_reference_tests_on_gpu(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "facade tests passed" in out.stdout
