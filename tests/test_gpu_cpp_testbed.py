"""The C++ testbed-style caller (sparse-register-tiling_amd/cpp: the
reference's SpMMFunctor lifecycle over the C ABI) run on the GPU: set_config
with the testbed's keys re-packs, every configuration's C passes the
reference's 64-epsilon tolerance rule against a double-precision host
product (cpp_testbed/demo/SPMM_demo.cpp:99-137), and an unknown key is
refused."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sparse-register-tiling_amd")
DEMO = os.path.join(PKG, "cpp", "bin", "testbed_demo")


@pytest.mark.parametrize("args", [["da01e"], ["400fa", "37", "70", "24", "0.3", "3"],
                                  ["61fee", "1030", "300", "200", "0.05", "5"]])
def test_cpp_testbed_functor(args):
    if not os.path.exists(DEMO):
        raise AssertionError("%s is not built (build_lib.build builds it)" % DEMO)
    print("testbed demo", args, flush=True)
    r = subprocess.run([DEMO, os.path.join(PKG, "mappings")] + args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and r.stdout.strip().endswith("PASS"), r.stdout + r.stderr
