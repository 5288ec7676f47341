"""Executor code generator: one CDNA4 HIP translation unit per pattern set.

The reference generates one AVX-512 micro-kernel header per pattern set
(``UKernelCodegenBase(...).gen_header``, used at sbench/SOP/sop_driver.py:93-107;
legacy generator tools/codegen/gen_vector_specialized.py:112-349).  Here the
unrolled per-pattern body is the ``srt::exec_kernel`` template
(csrc/exec_kernel.h); the generator emits, for a pattern list in id order
(the mapping's kernel list stable-sorted by popcount, as the legacy generator
sorts, :113; the same set in another order is another executor),

* the pattern table,
* the kernel instantiations for both B paths (rows read from global memory,
  or the workgroup's B tile staged in LDS), every lane width V (1, 2, 4 fp32
  columns per lane) and every unroll U (packed columns in flight: 1, 2, 4, 8),
* a static registrar that adds the executor to libsrt's registry under the id
  ``cdna4_M<M_r>_<hash>``.
"""
from __future__ import annotations

import hashlib
import os
from typing import Dict, Iterable, List, Sequence, Tuple

VEC_WIDTHS = (1, 2, 4)
UNROLLS = (1, 2, 4, 8)


def popcount(x: int) -> int:
    return bin(int(x)).count("1")


def id_ordered_patterns(patterns: Iterable[int]) -> List[int]:
    """The kernel patterns of an executor in id order, as given (a mapping's
    ``patterns``: its kernel list stable-sorted by popcount).  The order is
    part of the executor: it fixes the packed column order and so the
    accumulation order of every output."""
    out = [int(p) for p in patterns]
    if any(p <= 0 for p in out) or len(set(out)) != len(out):
        raise ValueError("kernel patterns must be distinct and nonzero")
    if any(popcount(a) > popcount(b) for a, b in zip(out, out[1:])):
        raise ValueError("kernel patterns are not in id order (stable popcount order)")
    return out


def nanokernel_hash(M_r: int, patterns: Sequence[int]) -> str:
    key = "%d:%s" % (M_r, ",".join(str(p) for p in id_ordered_patterns(patterns)))
    return hashlib.md5(key.encode()).hexdigest()[:10]


def executor_id(M_r: int, patterns: Sequence[int]) -> str:
    return "cdna4_M%d_%s" % (M_r, nanokernel_hash(M_r, patterns))


def variants_for(num_patterns: int, M_r: int = 8) -> List[Tuple[int, int]]:
    """(V, U) pairs to instantiate (for each B path).  Large pattern sets get
    one unroll to bound code size and compile time; V=4 with U=8 needs more
    than the 128 VGPRs a 1024-thread workgroup allows once M_r >= 8 (it would
    spill), so it is left out there."""
    unrolls = (4,) if num_patterns > 64 else (2, 4, 8)
    return [(v, u) for v in VEC_WIDTHS for u in unrolls if not (M_r >= 8 and v == 4 and u == 8)]


def _dispatch_text(M_r: int, pats: Sequence[int], U: int) -> str:
    cases = "\n".join(
        "            case %d: srt::batch_fmas<%d, MR, V, U, %d>(acc, rows, vals, vp); return;"
        % (i * U + n - 1, p, n) for i, p in enumerate(pats) for n in range(1, U + 1))
    return """#include "exec_kernel.h"

namespace {{

constexpr int MR = {M_r};

// (kernel id, batch size) -> unrolled FMAs of one batch.
struct Dispatch {{
    template <int V, int U>
    static __device__ __forceinline__ void fmas(int key, float (&acc)[MR][V],
                                                const typename srt::VecOf<V>::type (&rows)[U],
                                                const float (&vals)[srt::kValWindow],
                                                const float *__restrict__ vp) {{
        static_assert(U == {U}, "one unroll per translation unit");
        switch (key) {{
{cases}
            default: __builtin_unreachable();
        }}
    }}
}};

}}  // namespace
""".format(M_r=M_r, U=U, cases=cases)


# The grouped executor (several matrices in one launch) is instantiated for
# the global-memory B path at this unroll, every lane width.
GROUPED_UNROLL = 4


def _grouped_name(eid: str, v: int) -> str:
    return "srt_grouped_%s_v%d" % (eid, v)


def _persistent_name(eid: str, v: int) -> str:
    return "srt_persistent_%s_v%d" % (eid, v)


def _launcher_name(eid: str, stage: int, v: int, u: int) -> str:
    return "srt_launch_%s_s%d_v%d_u%d" % (eid, stage, v, u)


def gen_executor_sources(M_r: int, patterns: Sequence[int], lean: bool = False) -> Dict[str, str]:
    """Translation units of one executor: ``exec_<id>.hip`` (pattern table,
    launch table, registrar) and one ``exec_<id>__s<S>v<V>u<U>.hip`` per B
    path, lane width and unroll holding that kernel instantiation, so the
    variants of a large pattern set compile in parallel.  `lean` keeps only
    the global-memory B path at U = GROUPED_UNROLL (every lane width): the
    default launch, for executors built at run time."""
    if not patterns:
        raise ValueError("empty pattern set")
    pats = id_ordered_patterns(patterns)
    if not pats:
        raise ValueError("empty pattern set")
    if max(pats) >= (1 << M_r):
        raise ValueError("pattern wider than M_r")
    eid = executor_id(M_r, pats)
    variants = set(variants_for(len(pats), M_r))
    if lean:
        variants = {(v, u) for (v, u) in variants if u == GROUPED_UNROLL}
    head = "// Generated by sparse-register-tiling_amd/codegen.py; do not edit.\n" \
           "// Executor %s: M_r = %d, %d kernel patterns (id order):\n//   %s\n" \
           % (eid, M_r, len(pats), ", ".join(str(p) for p in pats))
    files = {}
    decls = []
    rows = []
    grouped = {}
    persistent = {}
    for stage in ((0,) if lean else (0, 1)):
        vrows = []
        for v in VEC_WIDTHS:
            cells = []
            for u in UNROLLS:
                if (v, u) in variants:
                    name = _launcher_name(eid, stage, v, u)
                    cells.append("&" + name)
                    decls.append("hipError_t %s(const srt::ExecLaunch &, hipStream_t);" % name)
                    body = ("__attribute__((visibility(\"hidden\"))) hipError_t %s(const srt::ExecLaunch &a, "
                            "hipStream_t s) {\n    return srt::launch_exec<MR, %d, %d, %s, Dispatch>(a, s);\n}"
                            % (name, v, u, "true" if stage else "false"))
                    if stage == 0 and u == GROUPED_UNROLL:
                        gname = _grouped_name(eid, v)
                        pname = _persistent_name(eid, v)
                        decls.append("hipError_t %s(const srt::ExecProblem *, int, int, int, size_t, int, "
                                     "hipStream_t);" % gname)
                        decls.append("hipError_t %s(const srt::ExecProblem *, const int *, int, int, size_t, "
                                     "hipStream_t);" % pname)
                        grouped[v] = "&" + gname
                        persistent[v] = "&" + pname
                        body += ("\n\n__attribute__((visibility(\"hidden\"))) hipError_t %s(const srt::ExecProblem *p, "
                                 "int np, int nb, int w, size_t lds, int x, hipStream_t s) {\n    return "
                                 "srt::launch_exec_grouped<MR, %d, %d, false, Dispatch>(p, np, nb, w, lds, x, s);\n}"
                                 % (gname, v, u))
                        body += ("\n\n__attribute__((visibility(\"hidden\"))) hipError_t %s(const srt::ExecProblem *p, "
                                 "const int *h, int g, int w, size_t lds, hipStream_t s) {\n    return "
                                 "srt::launch_exec_persistent<MR, %d, %d, false, Dispatch>(p, h, g, w, lds, s);\n}"
                                 % (pname, v, u))
                    files["exec_%s__s%dv%du%d.hip" % (eid, stage, v, u)] = (
                        head + _dispatch_text(M_r, pats, u) + "\n" + body + "\n")
                else:
                    cells.append("nullptr")
            vrows.append("         {" + ", ".join(cells) + "}")
        rows.append("        {\n" + ",\n".join(vrows) + "}")
    files["exec_%s.hip" % eid] = head + """#include "srt_internal.h"

{decls}

namespace {{

const int32_t k_patterns[] = {{{pat_list}}};

const srt::ExecutorDesc k_desc = {{
    "{eid}", {M_r}, {n}, k_patterns,
    {{
{rows}
    }},
    {{{grouped}}},
    {{{persistent}}}}};

struct Registrar {{
    Registrar() {{ srt::register_executor(&k_desc); }}
}} k_registrar;

}}  // namespace
""".format(decls="\n".join('__attribute__((visibility("hidden"))) ' + d for d in decls), eid=eid, M_r=M_r,
           n=len(pats), pat_list=", ".join(str(p) for p in pats), rows=",\n".join(rows),
           grouped=", ".join(grouped.get(v, "nullptr") for v in VEC_WIDTHS),
           persistent=", ".join(persistent.get(v, "nullptr") for v in VEC_WIDTHS))
    return files


def write_sources(pattern_sets: Sequence[Tuple[int, Sequence[int]]], outdir: str) -> List[str]:
    """Write the translation units of every distinct pattern set; remove
    stale generated files."""
    os.makedirs(outdir, exist_ok=True)
    wanted = {}
    for M_r, pats in pattern_sets:
        wanted.update(gen_executor_sources(M_r, pats))
    for f in os.listdir(outdir):
        if f.startswith("exec_") and f.endswith(".hip") and f not in wanted:
            os.remove(os.path.join(outdir, f))
    paths = []
    for name, text in sorted(wanted.items()):
        path = os.path.join(outdir, name)
        if not os.path.exists(path) or open(path).read() != text:
            with open(path, "w") as f:
                f.write(text)
        paths.append(path)
    return paths
