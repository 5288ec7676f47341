"""The C++ testbed-lifecycle adapter (sparse-register-tiling_amd/cpp) compiles
against include/srt.h, and its driver refuses a missing mapping directory
argument without touching a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "sparse-register-tiling_amd", "cpp")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC) and not shutil.which("hipcc"), reason="hipcc not installed")
def test_adapter_compiles():
    r = subprocess.run([HIPCC, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-Wno-unused-command-line-argument",
                        "--offload-arch=gfx950",
                        "-I" + os.path.join(ROOT, "include"), "-I" + CPP, os.path.join(CPP, "testbed_demo.cpp")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr


def test_demo_usage_without_arguments():
    demo = os.path.join(CPP, "bin", "testbed_demo")
    if not os.path.exists(demo):
        pytest.skip("demo not built")
    r = subprocess.run([demo], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
