"""Plans are immutable after creation: one plan executed concurrently on
several streams (different B and C) gives each stream the same bits as a
lone launch, and a captured graph of plan launches replays them exactly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_concurrent_streams_and_graph(srt):
    import torch
    A, _ = srt.workloads.case_inputs(srt.workloads.SpMMCase(1024, 512, 0.9, 256), seed=2)
    plan = srt.MatMulSpecialized.from_dense(torch.from_numpy(A), 256, mapping_id=srt.named_mapping(4, "orig"),
                                            tile_config=srt.TileConfig(K_c=128))
    Bs = [torch.rand((512, 256), device="cuda") * 2 - 1 for _ in range(4)]
    ref = []
    for B in Bs:
        C = torch.empty((1024, 256), device="cuda")
        plan(C, B)
        ref.append(C)
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream() for _ in range(4)]
    outs = [torch.full((1024, 256), float("nan"), device="cuda") for _ in range(4)]
    for rep in range(3):
        for s, B, C in zip(streams, Bs, outs):
            s.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(s):
                for _ in range(5):
                    plan(C, B, stream=s)
        torch.cuda.synchronize()
        for a, b in zip(ref, outs):
            assert torch.equal(a.view(torch.int32), b.view(torch.int32))
    # graph capture of several launches, replayed with eager work in between
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    gouts = [torch.zeros_like(o) for o in outs]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        for B, C in zip(Bs, gouts):
            plan(C, B, stream=s)
    for rep in range(3):
        for o in gouts:
            o.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(ref, gouts):
            assert torch.equal(a.view(torch.int32), b.view(torch.int32)), rep
    del g
    plan.close()
