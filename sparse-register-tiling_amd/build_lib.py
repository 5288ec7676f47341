"""Build libsrt.so for gfx950: generate mappings and executors, compile, link.

Run as ``python sparse-register-tiling_amd/build_lib.py [--incremental]`` or
through ``__graft_entry__.build()``.  Objects go to ``_build/`` next to this
file, the library to ``libsrt.so`` next to this file.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import time
from typing import List, Sequence, Tuple

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
GEN = os.path.join(CSRC, "generated")
OBJ = os.path.join(PKG_DIR, "_build")
LIB = os.path.join(PKG_DIR, "libsrt.so")
INCLUDE = os.path.join(REPO_DIR, "include")
ARCH = os.environ.get("SRT_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    # Packed fp32 (v_pk_*_f32) stays off: the executors issue scalar v_fmac_f32.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Keep every multiply-add explicit: results must reproduce the reference's
    # fmaf chains bit for bit.
    "-ffp-contract=off",
    # Leave uniform (wave-invariant) branches unstructurized: the executors
    # dispatch on wave-uniform kernel ids, and structurizing those switches
    # chains the cases through flag blocks (register copies, extra waits).
    "-mllvm", "-structurizecfg-skip-uniform-regions=true",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + INCLUDE, "-I" + CSRC,
]
HAND_SOURCES = ["capi.hip", "packer.hip", "csr_spmm.hip", "mfma_panel.hip", "mfma_capi.hip", "probe.hip", "dense_panels.hip",
                "mfma_exec.hip"]


def _load_sibling(name):
    import importlib.util
    spec = importlib.util.spec_from_file_location("_srt_build_" + name, os.path.join(PKG_DIR, name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def pattern_sets() -> List[Tuple[int, List[int]]]:
    """Id-ordered pattern lists that get a prebuilt executor: every shipped
    strategy (its own kernel list) and every shipped mapping file read as the
    reference's load_mapping reads it (the file's set)."""
    mapping = _load_sibling("mapping")
    mapping.generate_shipped_mappings()
    sets = {}
    for name in mapping.STRATEGIES:
        m = mapping.strategy_mapping(name)
        sets[(m.M_r, tuple(m.patterns))] = True
    for f in sorted(os.listdir(mapping.MAPPINGS_DIR)):
        if f.startswith("mapping_") and f.endswith(".txt"):
            m = mapping.read_mapping_file(os.path.join(mapping.MAPPINGS_DIR, f))
            sets[(m.M_r, tuple(m.patterns))] = True
    return [(M_r, list(p)) for (M_r, p) in sets]


def _compile(src: str, obj: str) -> Tuple[str, float, str]:
    t0 = time.time()
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(obj + ".tmp", obj)
    return src, time.time() - t0, r.stderr


def _stale(src: str, obj: str, deps: Sequence[str]) -> bool:
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in [src] + list(deps))


def build(incremental: bool = False, jobs: int = 0, verbose: bool = True) -> str:
    codegen = _load_sibling("codegen")
    sets = pattern_sets()
    gen_files = codegen.write_sources(sets, GEN)
    os.makedirs(OBJ, exist_ok=True)
    headers = [os.path.join(CSRC, h) for h in ("srt_internal.h", "exec_kernel.h")] + \
        [os.path.join(INCLUDE, "srt.h")]
    sources = [os.path.join(CSRC, s) for s in HAND_SOURCES] + gen_files
    objs = []
    todo = []
    for src in sources:
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        objs.append(obj)
        if not incremental or _stale(src, obj, headers):
            todo.append((src, obj))
    for f in os.listdir(OBJ):
        p = os.path.join(OBJ, f)
        if f.endswith(".o") and p not in objs:
            os.remove(p)
    jobs = jobs or min(len(todo) or 1, os.cpu_count() or 4, 16)
    t0 = time.time()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, dt, err in ex.map(lambda a: _compile(*a), todo):
            if verbose:
                print("  compiled %-48s %5.1fs" % (os.path.basename(src), dt), flush=True)
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", LIB + ".tmp"] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print("built %s (%d executors, %.1fs)" % (LIB, len(gen_files), time.time() - t0), flush=True)
    build_testbed_demo(verbose)
    return LIB


CPP_DIR = os.path.join(PKG_DIR, "cpp")
DEMO = os.path.join(CPP_DIR, "bin", "testbed_demo")


def build_testbed_demo(verbose: bool = True) -> str:
    """The C++ testbed-style caller of the C ABI (cpp/testbed_demo.cpp),
    linked against libsrt.so."""
    os.makedirs(os.path.dirname(DEMO), exist_ok=True)
    cmd = [HIPCC, "-O2", "-std=c++17", "--offload-arch=" + ARCH, "-Wall", "-I" + INCLUDE, "-I" + CPP_DIR,
           os.path.join(CPP_DIR, "testbed_demo.cpp"), "-o", DEMO + ".tmp", "-L" + PKG_DIR, "-lsrt",
           "-Wl,-rpath," + PKG_DIR]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("testbed demo build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(DEMO + ".tmp", DEMO)
    if verbose:
        print("built %s" % DEMO, flush=True)
    return DEMO


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=0)
    a = ap.parse_args()
    build(incremental=a.incremental, jobs=a.jobs)
