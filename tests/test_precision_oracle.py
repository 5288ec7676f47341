"""Pins of the precision-form oracle (CPU): number formats against torch's
own conversions, the MFMA block layout against a dense reconstruction, the
scale rule, and the error bound against planted mistakes."""
import numpy as np
import pytest
import torch

from oracle import precision_oracle as po


def _samples(n=50000, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n).astype(np.float32) * rng.choice([1e-4, 1e-2, 1, 100], n).astype(np.float32)
    edge = np.array([0, -0.0, 1, 448, 447.9, -448, 2 ** -9, 2 ** -10, 3 * 2 ** -10, 2 ** -6, 1.0625,
                     1.125, 1.1875, 240, 464, -464], dtype=np.float32)
    return np.concatenate([x, edge])


def test_bf16_matches_torch():
    x = _samples()
    ref = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(po.bf16_bits(x), ref)
    assert np.array_equal(po.bf16_value(ref), torch.from_numpy(x).to(torch.bfloat16).float().numpy())


def test_e4m3_matches_torch_in_range():
    x = _samples(seed=1)
    x = x[np.abs(x) <= 448]
    ref = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    assert np.array_equal(po.e4m3_bits(x), ref)
    codes = np.arange(256, dtype=np.uint8)
    tv = torch.from_numpy(codes).view(torch.float8_e4m3fn).float().numpy()
    assert np.array_equal(np.nan_to_num(po.e4m3_value(codes), nan=7.0), np.nan_to_num(tv, nan=7.0))


def test_e4m3_saturates_and_rounds_even():
    assert po.e4m3_bits(np.float32([1000.0, -1000.0, 464.0, 460.0]))[:2].tolist() == [0x7E, 0xFE]
    # 1 + 1/16 is halfway between 1 and 1.125: even (1.0) wins; 1 + 3/16 -> 1.25
    assert po.e4m3_value(po.e4m3_bits(np.float32([1.0625, 1.1875]))).tolist() == [1.0, 1.25]


def test_scales_rule():
    A = np.array([[0, 0, 0], [1, -896, 2], [0.5, 0, 0]], dtype=np.float32)
    s = po.row_scales(A)
    assert s.tolist() == [1.0, 2.0, np.float32(0.5 / 448)]
    qa, s2 = po.quantize_a(A, "fp8")
    assert np.array_equal(s, s2)
    assert po.e4m3_value(qa[1, 1]) == -448.0


@pytest.mark.parametrize("dtype", ["bf16", "fp8"])
def test_block_layout_reconstructs_dense(dtype):
    rng = np.random.default_rng(3)
    A = ((rng.random((37, 70)) < 0.1) * rng.standard_normal((37, 70))).astype(np.float32)
    A[16:32, 32:64] = 0.0  # one empty block
    slot, kb, blocks = po.pack_blocks(A, dtype)
    assert slot[1, 1] == -1  # the empty block is not stored
    codes, _ = po.quantize_a(A, dtype)
    rec = np.zeros((48, 96), dtype=codes.dtype)
    lane = np.arange(64)
    for p in range(slot.shape[0]):
        for b in range(slot.shape[1]):
            if slot[p, b] < 0:
                continue
            blk = blocks[slot[p, b]].reshape(64, 8)
            for e in range(8):
                rec[p * 16 + lane % 16, b * 32 + 8 * (lane // 16) + e] = blk[:, e]
    assert np.array_equal(rec[:37, :70], np.where(A != 0, codes, 0))
    # stored blocks in (panel, K block) order: slots ascend, blk_kb names each
    # block's K block, and exactly the blocks holding a nonzero are stored
    stored = slot.ravel()[slot.ravel() >= 0]
    assert np.array_equal(stored, np.arange(len(blocks)))
    assert np.array_equal(kb, np.nonzero(slot >= 0)[1])
    nz = np.zeros((48, 96), dtype=bool)
    nz[:37, :70] = A != 0
    assert np.array_equal(slot >= 0, nz.reshape(3, 16, 3, 32).any(axis=(1, 3)))
    assert len(blocks) == 8 and all(blocks[i].any() for i in range(len(blocks)))


@pytest.mark.parametrize("dtype", ["bf16", "fp8"])
def test_reference_and_bound_catch_mistakes(dtype):
    rng = np.random.default_rng(4)
    A = ((rng.random((64, 96)) < 0.4) * rng.standard_normal((64, 96))).astype(np.float32)
    B = rng.uniform(-1, 1, (96, 40)).astype(np.float32)
    C, bound = po.reference_f64(A, B, dtype)
    assert np.all(np.abs(C.astype(np.float32) - C) <= bound)
    # a dropped term, a transposed tile and a missing scale all break the bound
    bad = C.copy()
    bad[3, 5] -= C[3, 5] if C[3, 5] != 0 else 1.0
    assert not np.all(np.abs(bad - C) <= bound)
    assert not np.all(np.abs(C[:40, :40].T - C[:40, :40]) <= bound[:40, :40])
    if dtype == "fp8":
        qa, s = po.quantize_a(A, dtype)
        qb, t = po.quantize_b(B, dtype)
        noscale = (po.dequant(qa, dtype).astype(np.float64) * (A != 0)) @ po.dequant(qb, dtype)
        assert not np.all(np.abs(noscale - C) <= bound)
    # quantisation error vs FP64 of the unquantised inputs is what the format allows
    err = po.rel_error_vs_fp64(C, A, B)
    assert err < (1e-2 if dtype == "bf16" else 8e-2)
