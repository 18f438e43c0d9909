"""The facade's conv family with user __device__ lambdas (include/SparseGridGpu_device.hpp):
tests/cpp/test_conv.hip compiles with hipcc against the C ABI, and its checks pass on the
GPU (each lambda form against the library's named stencils or host-computed values)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "openfpm_data_amd")
HIPCC = "/opt/rocm/bin/hipcc"


def _build(tmp_path):
    exe = str(tmp_path / "test_conv")
    subprocess.check_call([HIPCC, "-std=c++17", "-O2", "--offload-arch=gfx950", "-ffp-contract=off",
                           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wextra",
                           "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "test_conv.hip"),
                           "-L", LIBDIR, "-lsg_hip", "-Wl,-rpath," + LIBDIR, "-o", exe])
    return exe


def test_conv_lambdas_compile(tmp_path):
    assert os.path.exists(_build(tmp_path))


@pytest.mark.gpu
def test_conv_lambdas_on_gpu(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "conv tests passed" in out.stdout
