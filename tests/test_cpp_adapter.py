"""The C++ adapter (include/orqg_adapter.hpp) compiles with g++ against
include/orqg.h, links liborqg.so, and runs: its argument and error paths on
the CPU, and a radix sort through it on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "adapter_test.cpp")
LIBDIR = os.path.join(ROOT, "orq_amd")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    import orq_amd

    assert os.path.exists(orq_amd.LIB_PATH), "build liborqg.so first"
    out = str(tmp_path_factory.mktemp("cpp") / "adapter_test")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           SRC, "-o", out, "-L", LIBDIR, "-lorqg", "-Wl,-rpath," + LIBDIR,
           "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_adapter_cpu_paths(exe):
    r = subprocess.run([exe, "cpu"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "cpu ok" in r.stdout


@pytest.mark.gpu
def test_adapter_radix_sort_gpu(exe):
    r = subprocess.run([exe, "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "gpu ok" in r.stdout
