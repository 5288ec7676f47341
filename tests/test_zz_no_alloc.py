"""The allocation-free entry point srt_mfma_plan_execute_noalloc, called
through the C ABI, against srt_mfma_plan_execute (same kernels, partial
tiles in a reserved workspace): bit-identical results on two calls in a row,
the second with a smaller input, for BF16 and FP8 with K splits; sizes past
the scratch maximum fail with a clear error."""
import ctypes as ct
import importlib

import numpy as np
import pytest

from oracle import precision_oracle as po

pytestmark = pytest.mark.gpu


def _sparse(M, K, density, seed):
    rng = np.random.default_rng(seed)
    return ((rng.random((M, K)) < density) * rng.standard_normal((M, K))).astype(np.float32)


def _call(srt, name, plan, C, Bq, t, n, shape, ks):
    from importlib import import_module
    native = import_module("sparse-register-tiling_amd.native")
    return getattr(srt._lib.lib(), name)(plan._h, ct.c_void_p(C.data_ptr()), C.stride(0), ct.c_void_p(Bq.data_ptr()),
                                Bq.stride(0), native.ptr(t), n, shape, ks, native.stream_handle(None))


@pytest.mark.parametrize("dtype", ["bf16", "fp8"])
def test_noalloc_matches_reserved_execute(dtype):
    import torch
    srt = importlib.import_module("sparse-register-tiling_amd")
    L = srt._lib.lib()
    cases = [(1024, 4096, 512, 4, 6), (300, 1100, 77, 3, 2)]   # second call: smaller input
    for M, K, N, ks, shape in cases:
        print("noalloc %s %dx%d n=%d ks=%d" % (dtype, M, K, N, ks), flush=True)
        A = _sparse(M, K, 0.3, seed=M)
        B = np.random.default_rng(K).uniform(-1, 1, (K, N)).astype(np.float32)
        plan = srt.precision.PanelPlan.from_dense(torch.from_numpy(A), dtype)
        Bq, t = plan.prepare_b(torch.from_numpy(B).cuda())
        plan.reserve(N, ks)
        C_ref = torch.full((M, N), float("nan"), device="cuda")
        C_new = torch.full((M, N), float("nan"), device="cuda")
        assert _call(srt, "srt_mfma_plan_execute", plan, C_ref, Bq, t, N, shape, ks) == 0
        assert _call(srt, "srt_mfma_plan_execute_noalloc", plan, C_new, Bq, t, N, shape, ks) == 0
        torch.cuda.synchronize()
        a, b = C_ref.cpu().numpy(), C_new.cpu().numpy()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        ref, bound = po.reference_f64(A, B, dtype)
        assert np.all(np.abs(b - ref) <= bound + 1e-30)
        plan.close()
    # past the maximum: a clear error, nothing launched
    plan = srt.precision.PanelPlan.from_dense(torch.from_numpy(_sparse(4096, 4096, 0.05, 1)), dtype)
    Bq, t = plan.prepare_b(torch.rand((4096, 1024), device="cuda"))
    C = torch.zeros((4096, 1024), device="cuda")
    rc = _call(srt, "srt_mfma_plan_execute_noalloc", plan, C, Bq, t, 1024, 6, 8)
    assert rc == srt._lib.SRT_ERR_INVALID and "scratch maximum" in srt._lib.last_error()
    plan.close()
