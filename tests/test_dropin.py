"""Drop-in check of the C++ facade compiled by hipcc (tests/cpp/test_dropin.hip): the
reference's testFlushInsert and testStencil_lap_simplified with the reference's call forms
(grid_smb constructor, host insertFlush + hostToDevice, get<p> on toKernel() in a kernel,
applyStencils with a user boundary functor), and user stencil functors (HeatStencil,
Conv3x3x3, Conv3x3x3_noshared) bitwise against the library's named stencils."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "openfpm_data_amd")
HIPCC = "/opt/rocm/bin/hipcc"


def _build(tmp_path):
    exe = str(tmp_path / "test_dropin")
    subprocess.check_call([HIPCC, "-std=c++17", "-O2", "--offload-arch=gfx950", "-ffp-contract=off",
                           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wextra",
                           "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "test_dropin.hip"),
                           "-L", LIBDIR, "-lsg_hip", "-Wl,-rpath," + LIBDIR, "-o", exe])
    return exe


def test_dropin_compiles(tmp_path):
    assert os.path.exists(_build(tmp_path))


@pytest.mark.gpu
def test_dropin_reference_tests_on_gpu(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "dropin tests passed" in out.stdout
