"""Host-side checks that need no GPU: the pattern-set generator and mapping IO
of the package, the executor code generator, and the C ABI of libsrt.so
(loads, exports every declared symbol, host-only calls behave)."""
import ctypes as ct
import glob
import os
import re

import pytest

from conftest import GOLDEN_MAPPINGS, ROOT


# ----------------------------------------------------------- mapping module
def test_shipped_strategies_reproduce_reference_ids(srt, tmp_path):
    ids = srt.mapping.generate_shipped_mappings(str(tmp_path))
    assert ids["M4_identity"] == "61fee"
    assert ids["M4_orig"] == "da01e"
    assert ids["M8_orig"] == "400fa"
    assert ids["M8_alt"] == "747f9"
    assert ids["M8_identity"] == "161fb"
    assert ids["M2_identity"] == "4af51"
    assert ids["M6_identity"] == "001e3"
    for name, mid in ids.items():
        golden = os.path.join(GOLDEN_MAPPINGS, "mapping_%s.txt" % mid)
        if os.path.exists(golden):
            assert open(tmp_path / ("mapping_%s.txt" % mid)).read() == open(golden).read()


def test_named_mappings(srt):
    assert srt.named_mapping(4, "identity").id == "61fee"
    assert srt.named_mapping(4, "orig").id == "da01e"
    assert srt.named_mapping(8, "orig").id == "400fa"
    assert srt.named_mapping(8, "alt").id == "747f9"
    m = srt.named_mapping(8, "orig")
    assert m(0) == [] and m(3) == [195]


# Kernel id tables, written out: the reference generator's stable popcount
# sort (tools/codegen/gen_vector_specialized.py:113) of the list its benchmark
# passes, config.all_patterns() = merge then split patterns
# (sbench/SOP/sop_hand_configs.py:16-17,99-103,160-169,209-219).
KAT_400FA = [1, 2, 4, 8, 16, 32, 64, 128, 85, 170, 195, 60, 15, 240, 252, 243, 207, 63, 255]
KAT_DA01E = [1, 2, 4, 8, 14, 13, 11, 7, 15]
KAT_61FEE = [1, 2, 4, 8, 3, 5, 6, 9, 10, 12, 7, 11, 13, 14, 15]
# the same 400fa / 747f9 pattern set read from its file: list(set) of the
# file's patterns (load_mapping, sbench/SOP/sop_bench.py:72-95) -- 128 sits in
# slot 0 of the interpreter's 64-slot table -- then the same sort
KAT_400FA_FILE = [128, 1, 2, 4, 8, 16, 32, 64, 15, 170, 60, 195, 85, 240, 63, 207, 243, 252, 255]


def test_kernel_id_kats(srt):
    assert srt.named_mapping(8, "orig").patterns == KAT_400FA
    assert srt.named_mapping(8, "alt").patterns == KAT_400FA
    assert srt.named_mapping(4, "orig").patterns == KAT_DA01E
    assert srt.named_mapping(4, "identity").patterns == KAT_61FEE
    assert srt.named_mapping(8, "400fa").patterns == KAT_400FA   # by id: the strategy's list
    for M_r, name, kat in [(8, "orig", KAT_400FA), (8, "alt", KAT_400FA), (4, "orig", KAT_DA01E),
                           (4, "identity", KAT_61FEE)]:
        assert srt.resolve_mapping(name, M_r=M_r).patterns == kat    # through srt_mapping_create_ordered
    assert srt.resolve_mapping("400fa").patterns == KAT_400FA
    # the file path, in Python and through the C reader
    for mid in ("400fa", "747f9"):
        assert srt.read_mapping_file(os.path.join(GOLDEN_MAPPINGS, "mapping_%s.txt" % mid)).patterns == KAT_400FA_FILE
        assert srt.NativeMapping.from_id(mid, [GOLDEN_MAPPINGS]).patterns == KAT_400FA_FILE


def test_make_sop_module_ids_follow_the_pattern_list(srt):
    """make_sop_module(acc, patterns, mapping): the ids come from the
    ``patterns`` argument (sop_driver.py:93), unused patterns included."""
    strat = srt.SOPOrig([14, 13, 11, 7, 15], [1, 2, 4, 8], split_cost=10000)
    pats = strat.all_patterns()
    m = srt.Mapping(4, srt.mapping.mapping_table(4, strat.map_pattern), kernel_list=pats + [3])
    assert m.patterns == [1, 2, 4, 8, 3, 14, 13, 11, 7, 15]
    with pytest.raises(ValueError):
        srt.Mapping(4, srt.mapping.mapping_table(4, strat.map_pattern), kernel_list=[1, 2, 4, 8])


def _shuffled_mapping_file(path, M_r, rng, keep_frac):
    """A mapping file whose lines come in random order: a kept code maps to
    itself, the others to the full pattern.  Returns the kernel patterns in
    the order the lines add them."""
    full = (1 << M_r) - 1
    codes = list(range(1, full + 1))
    rng.shuffle(codes)
    keep = set(c for c in codes if rng.random() < keep_frac)
    seq = []
    with open(path, "w") as f:
        f.write("%d\n" % M_r)
        for c in codes:
            k = c if c in keep else full
            f.write("%d: [%d]\n" % (c, k))
            seq.append(k)
    return seq


def test_file_id_order_is_the_interpreters_set_order(srt, tmp_path):
    """The C reader's ids equal the stable popcount sort of ``list(set)``
    built by this interpreter in file-line order, for random line orders and
    pattern subsets (the construct of the reference's load_mapping)."""
    import random
    rng = random.Random(5)
    for trial in range(24):
        M_r = [3, 5, 7, 9, 11][trial % 5]
        d = tmp_path / ("t%d" % trial)
        d.mkdir()
        seq = _shuffled_mapping_file(str(d / "mapping_zzzzz.txt"), M_r, rng, rng.choice([0.02, 0.2, 0.7, 1.0]))
        pyset = set()
        for k in seq:
            pyset.add(k)
        want = sorted(list(pyset), key=lambda p: bin(p).count("1"))
        assert srt.NativeMapping.from_id("zzzzz", [str(d)]).patterns == want, trial
        assert srt.read_mapping_file(str(d / "mapping_zzzzz.txt")).patterns == want


def test_load_mapping_like_reference(srt):
    acc, patterns, m = srt.load_mapping("400fa", [GOLDEN_MAPPINGS])
    assert acc["M_r"] == 8 and acc["N_r"]["avx512"] == 2
    assert patterns == set(m.patterns)


def test_sop_orig_matches_generator_rules(srt):
    s = srt.SOPOrig([0b1111], [1, 2, 4, 8], split_cost=10000)
    assert s.map_pattern(0) == [0]
    assert s.map_pattern(0b0001) == [0b0001]          # below min_merge
    assert s.map_pattern(0b0101) == [0b1111]          # merge with padding
    r = srt.SOPOrig([0b0011, 0b1100], [1, 2, 4, 8], split_cost=2.1, recursive=True)
    assert r.map_pattern(0b1111) == [0b0011, 0b1100]
    assert srt.gen_all_combos(4, 3) == [7, 11, 13, 14]


def test_code_arrays(srt):
    m = srt.named_mapping(4, "orig")
    ptr, flat = m.code_arrays()
    assert len(ptr) == 17 and ptr[0] == 0 and ptr[1] == 0
    for code in range(1, 16):
        assert flat[ptr[code]:ptr[code + 1]] == m.table[code]


# ---------------------------------------------------------------- codegen
def test_codegen_sources_are_current(srt):
    cg = srt.mapping  # noqa: F841
    import importlib
    codegen = importlib.import_module("sparse-register-tiling_amd.codegen")
    gen_dir = os.path.join(ROOT, "sparse-register-tiling_amd", "csrc", "generated")
    files = sorted(f for f in glob.glob(os.path.join(gen_dir, "exec_*.hip")) if "__" not in f)
    assert files, "build() has not generated executors"
    for f in files:
        text = open(f).read()
        M_r = int(re.search(r"M_r = (\d+)", text).group(1))
        pats = [int(x) for x in re.search(r"k_patterns\[\] = \{([^}]*)\}", text).group(1).split(",")]
        units = codegen.gen_executor_sources(M_r, pats)
        assert os.path.basename(f) == "exec_%s.hip" % codegen.executor_id(M_r, pats)
        for name, want in units.items():
            with open(os.path.join(gen_dir, name)) as g:
                assert g.read() == want, name


def test_codegen_rejects_bad_sets():
    import importlib
    codegen = importlib.import_module("sparse-register-tiling_amd.codegen")
    with pytest.raises(ValueError):
        codegen.gen_executor_sources(4, [])
    with pytest.raises(ValueError):
        codegen.gen_executor_sources(4, [16])
    assert codegen.id_ordered_patterns([2, 1, 15]) == [2, 1, 15]
    with pytest.raises(ValueError):
        codegen.id_ordered_patterns([1, 3, 2])     # not in stable popcount order
    with pytest.raises(ValueError):
        codegen.id_ordered_patterns([1, 1])
    # the same set in another id order is another executor
    assert codegen.executor_id(4, [1, 2, 4, 8, 14, 13, 11, 7, 15]) != codegen.executor_id(4, [1, 2, 4, 8, 7, 11, 13, 14, 15])


# ------------------------------------------------------------------ C ABI
def _declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "srt.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(srt_[a-z0-9_]+)\s*\(", hdr)))


def test_every_declared_symbol_is_exported(srt):
    L = srt._lib.lib()
    syms = _declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) <= set(srt._lib.SIGNATURES), set(syms) - set(srt._lib.SIGNATURES)


def test_c_mapping_reader_on_golden_files(srt):
    for f in sorted(glob.glob(os.path.join(GOLDEN_MAPPINGS, "mapping_*.txt"))):
        mid = os.path.basename(f)[len("mapping_"):-4]
        nm = srt.NativeMapping.from_id(mid, [GOLDEN_MAPPINGS])
        py = srt.read_mapping_file(f)
        assert nm.M_r == py.M_r
        assert nm.patterns == py.patterns


def test_shipped_executors_cover_named_mappings(srt):
    ids = {e[0] for e in srt.executors()}
    for M_r, name in [(4, "identity"), (4, "orig"), (8, "orig"), (8, "alt")]:
        nm = srt.resolve_mapping(name, M_r=M_r)
        assert nm.executor_id() in ids
    for mid in ("61fee", "da01e", "400fa", "747f9"):      # the file path (load_mapping's ids)
        nm = srt.NativeMapping.from_id(mid)
        assert nm.executor_id() in ids


def test_c_errors(srt):
    L = srt._lib.lib()
    h = ct.c_void_p()
    rc = L.srt_mapping_read(b"nonexistent", (ct.c_char_p * 1)(GOLDEN_MAPPINGS.encode()), 1, ct.byref(h))
    assert rc == srt._lib.SRT_ERR_IO
    assert "not found" in srt._lib.last_error()
    # overlapping kernel patterns would add a value twice
    ptr = [0, 0, 1, 2, 4] + [4] * 13
    pats = [1, 2, 3, 1]
    rc = L.srt_mapping_create(4, srt._lib.int32_array(ptr[:17]), srt._lib.int32_array(pats), ct.byref(h))
    assert rc == srt._lib.SRT_ERR_INVALID
    # a pattern set without a generated executor
    m = srt.Mapping(3, {c: [c] for c in range(1, 8)})
    nm = srt.NativeMapping.from_mapping(m)
    with pytest.raises(srt.SrtError) as e:
        nm.executor_id()
    assert e.value.code == srt._lib.SRT_ERR_UNSUPPORTED
    # plan creation argument checks happen before any device work
    plan = ct.c_void_p()
    rc = L.srt_plan_create(4, 4, 4, None, None, None, 3, None, 1, None, nm.handle, None, ct.byref(plan))
    assert rc == srt._lib.SRT_ERR_INVALID
    rc = L.srt_csr_spmm(4, 4, 4, 0, None, None, 4, None, None, 4, None, 4, 48, None)
    assert rc == srt._lib.SRT_ERR_INVALID


def test_default_tile_config(srt):
    c = srt._lib.TileConfigC()
    srt._lib.lib().srt_tile_config_default(ct.byref(c))
    t = srt.TileConfig.from_c(c)
    # the reference's SOPConfig defaults, except K_c: 0 = one K tile (256 would tile K like the reference)
    assert (t.M_c, t.K_c, t.N_c, t.beta, t.max_tlb_entries, t.tlb_page_size) == (16, 0, 64, 1.0, 64, 4096)
    # GPU fields default to "choose" (the balanced workgroup layout, among others)
    assert (t.vec_width, t.waves_per_block, t.unroll, t.nsplit, t.chunk_cols, t.b_staging, t.layout) == (0,) * 7
    assert srt.TileConfig.from_c(t.to_c()) == t
    # the ctypes mirror has every field of srt_tile_config, in order
    assert [f[0] for f in srt._lib.TileConfigC._fields_][-4:] == ["chunk_cols", "b_staging", "layout", "matrix_cores"]
    assert t.matrix_cores == 0
    assert ct.sizeof(srt._lib.TileConfigC) == 18 * 4


def test_group_abi_rejects_bad_arguments(srt):
    """Grouped-execution entry points fail with a status, not a crash."""
    L = srt._lib.lib()
    out = ct.c_void_p()
    assert L.srt_group_create(None, 1, None, None, None, None, None, None, ct.byref(out)) == srt._lib.SRT_ERR_INVALID
    plans = (ct.c_void_p * 1)()
    assert L.srt_group_create(plans, 0, None, None, None, None, None, None, ct.byref(out)) == srt._lib.SRT_ERR_INVALID
    assert L.srt_group_create(plans, 1, None, None, None, None, None, None, ct.byref(out)) == srt._lib.SRT_ERR_INVALID
    assert L.srt_group_execute(None, None) == srt._lib.SRT_ERR_INVALID
    assert L.srt_group_info(None, None, None, None) == srt._lib.SRT_ERR_INVALID
    assert L.srt_group_rebind(None, None, None, None, None, None, None) == srt._lib.SRT_ERR_INVALID
    L.srt_group_destroy(None)


def test_sparsity_bucket_rounds_pruning_counts(srt):
    """A matrix pruned to a nominal sparsity keeps floor or round of the kept
    count; both land in the nominal sparsity's bucket of the tuning table."""
    at = srt.autotune
    assert at.sparsity_bucket(1 - 78643 / 262144.0) == "(0,0.7]"        # 512 x 512 at 70 %
    assert at.sparsity_bucket(1 - 104858 / 1048576.0) == "(0.8,0.9]"    # 2048 x 512 at 90 %
    assert at.sparsity_bucket(0.95) == "(0.9,1]" and at.sparsity_bucket(0.75) == "(0.7,0.8]"


def test_rank_with_evidence_prefers_less_traffic_within_slack(srt):
    """Cold time orders the records; inside 3 % of the fastest, cold time x
    sqrt(HBM bytes / fewest) picks the leader; records without counters
    never jump."""
    at = srt.autotune
    recs = [dict(seconds=1.0, seconds_cold=10.0e-6, hbm_bytes=9_000_000, id="a"),
            dict(seconds=1.0, seconds_cold=10.2e-6, hbm_bytes=6_000_000, id="b"),
            dict(seconds=1.0, seconds_cold=10.25e-6, id="c"),
            dict(seconds=1.0, seconds_cold=12.0e-6, hbm_bytes=1_000_000, id="d")]
    # b: 2 % slower, a third less traffic -> leads
    assert [r["id"] for r in at.rank_with_evidence(recs)] == ["b", "a", "c", "d"]
    # b: 1 % slower but only 1 % less traffic -> the fastest keeps the lead
    recs[1].update(seconds_cold=10.1e-6, hbm_bytes=8_910_000)
    assert [r["id"] for r in at.rank_with_evidence(recs)] == ["a", "b", "c", "d"]
    # b beyond the slack: time alone
    recs[1].update(seconds_cold=10.4e-6, hbm_bytes=1_000_000)
    assert [r["id"] for r in at.rank_with_evidence(recs)] == ["a", "c", "b", "d"]


def test_counter_means_reads_rocprof_csv(srt, tmp_path):
    """Per-kernel, per-counter means over dispatches from rocprofv3's
    counter_collection CSV (the layout rocprofv3 --output-format csv writes)."""
    at = srt.autotune
    d = tmp_path / "host" / "1234"
    d.mkdir(parents=True)
    (d / "run_counter_collection.csv").write_text(
        "Correlation_Id,Dispatch_Id,Agent_Id,Queue_Id,Process_Id,Thread_Id,Grid_Size,Kernel_Id,Kernel_Name,"
        "Workgroup_Size,LDS_Block_Size,Scratch_Size,VGPR_Count,Accum_VGPR_Count,SGPR_Count,Counter_Name,Counter_Value,"
        "Start_Timestamp,End_Timestamp\n"
        "1,1,1,1,1,1,262144,3,exec_kernel<4>,1024,0,0,64,0,32,FETCH_SIZE,4000,0,1\n"
        "2,2,1,1,1,1,262144,3,exec_kernel<4>,1024,0,0,64,0,32,FETCH_SIZE,6000,0,1\n"
        "2,2,1,1,1,1,262144,3,exec_kernel<4>,1024,0,0,64,0,32,WRITE_SIZE,2048,0,1\n"
        "3,3,1,1,1,1,1024,4,elementwise_kernel,256,0,0,16,0,16,FETCH_SIZE,262144,0,1\n")
    m = at._counter_means(str(tmp_path))
    assert m["exec_kernel<4>"] == {"FETCH_SIZE": 5000.0, "WRITE_SIZE": 2048.0}
    assert m["elementwise_kernel"] == {"FETCH_SIZE": 262144.0}
    assert at._counter_means(str(tmp_path), last=1)["exec_kernel<4>"]["FETCH_SIZE"] == 6000.0
