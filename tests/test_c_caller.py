"""A C program written against include/ocfd_hip.h alone (tests/c_abi/):
it compiles with gcc, links with the library, and runs -- without a GPU it
must stop with the library's message; on a GPU a uniform flow must come
back unchanged after 20 whole steps, to the rounding of the RK3 stage
combinations and the primitive recovery (relative 1e-13)."""
import os
import shutil
import subprocess

import pytest

from util import gpu_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "opencfd-sc_amd")


def _build(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    if not os.path.exists(os.path.join(LIBDIR, "libocfd_hip.so")):
        pytest.skip("library not built")
    exe = str(tmp_path / "uniform_flow")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "c_abi", "uniform_flow.c"), "-lm", "-o", exe,
                        "-L" + LIBDIR, "-locfd_hip", "-Wl,-rpath," + LIBDIR], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_caller_builds_and_fails_cleanly_without_gpu(tmp_path):
    exe = _build(tmp_path)
    if gpu_available():
        pytest.skip("a GPU is present (the GPU test runs the program)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, env=dict(os.environ, HIP_VISIBLE_DEVICES=""))
    assert r.returncode == 1, r.stdout + r.stderr
    assert r.stdout.startswith("ocfd_create:") and "device" in r.stdout


@pytest.mark.gpu
def test_c_caller_uniform_flow_on_gpu(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "steps 20" in r.stdout and "values changed 0" in r.stdout
