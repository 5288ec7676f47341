"""Build an executor for a pattern set that libsrt was not built with.

The reference JIT-compiles one torch extension per pattern set
(sbench/SOP/sop_driver.py:100-127).  Here the generator writes the CDNA4
translation units, hipcc builds them into ``_jit/libsrt_exec_<id>.so`` (linked
against libsrt.so), and loading that library registers the executor with
libsrt.  Built libraries are reused across processes.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

from . import _lib
from .codegen import executor_id, gen_executor_sources

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
JIT_DIR = os.path.join(PKG_DIR, "_jit")
_loaded = {}


def _compile(eid: str, sources) -> str:
    """Compile the translation units in parallel, link against libsrt."""
    import concurrent.futures as cf
    from importlib import util
    spec = util.spec_from_file_location("_srt_build_lib", os.path.join(PKG_DIR, "build_lib.py"))
    bl = util.module_from_spec(spec)
    spec.loader.exec_module(bl)
    os.makedirs(JIT_DIR, exist_ok=True)
    objs = []
    for name, text in sorted(sources.items()):
        path = os.path.join(JIT_DIR, name)
        with open(path, "w") as f:
            f.write(text)
        objs.append((path, path + ".o"))
    jobs = min(len(objs), os.cpu_count() or 4, 16)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda a: bl._compile(*a), objs))
    out = os.path.join(JIT_DIR, "libsrt_exec_%s.so" % eid)
    cmd = [bl.HIPCC, "-shared", "-fPIC", "--offload-arch=" + bl.ARCH] + [o for _, o in objs] + \
        ["-o", out + ".tmp", "-L" + PKG_DIR, "-lsrt", "-Wl,-rpath," + PKG_DIR]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("executor link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(out + ".tmp", out)
    return out


def ensure_executor(native, regen: bool = False, lean: bool = False) -> str:
    """Executor id for ``native``'s pattern set, building it if needed
    (`lean`: only the default launch variants, see codegen)."""
    try:
        if not regen:
            return native.executor_id()
    except _lib.SrtError as e:
        if e.code != _lib.SRT_ERR_UNSUPPORTED:
            raise
    eid = executor_id(native.M_r, native.patterns)
    path = os.path.join(JIT_DIR, "libsrt_exec_%s.so" % eid)
    # A built executor is reused unless libsrt or the sources it was generated
    # from are newer (the executor table layout lives in srt_internal.h).
    deps = [_lib.LIB_PATH, os.path.join(PKG_DIR, "codegen.py"),
            os.path.join(PKG_DIR, "csrc", "srt_internal.h"), os.path.join(PKG_DIR, "csrc", "exec_kernel.h")]
    stale = not os.path.exists(path) or any(
        os.path.exists(d) and os.path.getmtime(d) > os.path.getmtime(path) for d in deps)
    if regen or stale:
        path = _compile(eid, gen_executor_sources(native.M_r, native.patterns, lean=lean))
    if path not in _loaded:
        _lib.lib()
        _loaded[path] = ct.CDLL(path)
    return native.executor_id()
