"""The C++ drop-in header (cuda-sort_amd/cusort.hpp): compiles and links
against libcusort_hip.so with the reference's entry-point names (CPU), and
sorts correctly when called like the reference's own callers (GPU)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "cuda-sort_amd")


def build(out):
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           os.path.join(HERE, "cpp", "dropin_test.cpp"), "-o", str(out), f"-L{PKG}", "-lcusort_hip",
           "-L/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)


def test_dropin_compiles_and_links(tmp_path):
    build(tmp_path / "dropin_test")
    assert (tmp_path / "dropin_test").exists()


@pytest.mark.gpu
def test_dropin_sorts(tmp_path):
    exe = tmp_path / "dropin_test"
    build(exe)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "ALL PASSED" in r.stdout, r.stdout[-3000:] + r.stderr[-2000:]
