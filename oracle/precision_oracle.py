"""CPU restatement of the precision form (BF16 / FP8-e4m3 16-row MFMA panels).

TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py cpu_baseline).

The precision form has no counterpart in the reference (its executors are
FP32 only, tools/codegen/gen_vector_specialized.py:266-349); what it must
reproduce is fixed by public definitions, each pinned in tests:
  * bfloat16: upper 16 bits of IEEE binary32, round to nearest even
    (checked against torch.bfloat16 conversion);
  * FP8 e4m3 (OCP "fn"): bias 7, 3 mantissa bits, no infinities, max 448,
    subnormals in steps of 2^-9, round to nearest even (checked against
    torch.float8_e4m3fn conversion);
  * per-row / per-column scales max|x| / 448 in FP64 (1.0 for an all-zero
    row or column), stored as float32;
  * the MFMA operand layout of a 16 x 32 block: lane l holds row l % 16,
    columns 8 * (l // 16) .. + 8 (ISA layout of v_mfma_f32_16x16x32_*).
"""
from __future__ import annotations

import numpy as np

PANEL = 16
BLOCK_K = 32


def bf16_bits(x) -> np.ndarray:
    """float32 -> bfloat16 bit patterns (uint16), round to nearest even."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return r.astype(np.uint16)


def bf16_value(bits) -> np.ndarray:
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32)


def e4m3_bits(x) -> np.ndarray:
    """float32 -> e4m3fn codes (uint8), round to nearest even, saturating at 448."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32)
    sign = ((u >> 24) & 0x80).astype(np.uint8)
    a = u & 0x7FFFFFFF
    out = np.zeros(f.shape, dtype=np.uint8)
    e = (a >> 23).astype(np.int64) - 127
    # subnormal target
    sub = e < -6
    q = np.rint(a.view(np.float32) * np.float32(512.0)).astype(np.int64)
    out[sub] = q[sub].astype(np.uint8)
    # normal target
    m = (a & 0x7FFFFF).astype(np.int64)
    r = m >> 20
    rem = m & 0xFFFFF
    ex = e + 7
    up = (rem > 0x80000) | ((rem == 0x80000) & ((r & 1) == 1))
    r = r + up
    carry = r == 8
    r = np.where(carry, 0, r)
    ex = np.where(carry, ex + 1, ex)
    sat = (ex > 15) | ((ex == 15) & (r == 7))
    normal = ~sub
    code = np.where(sat, 0x7E, (ex << 3) | r)
    out[normal] = code[normal].astype(np.uint8)
    out[a >= 0x43E00000] = 0x7E
    out = out | sign
    out[a > 0x7F800000] = 0x7F
    return out


def e4m3_value(code) -> np.ndarray:
    c = np.asarray(code, dtype=np.uint8).astype(np.int64)
    s = np.where(c & 0x80, -1.0, 1.0)
    e = (c >> 3) & 0xF
    m = c & 7
    v = np.where(e == 0, m * 2.0 ** -9, (1.0 + m / 8.0) * 2.0 ** (e - 7))
    v = np.where((c & 0x7F) == 0x7F, np.nan, v)
    return (s * v).astype(np.float32)


def row_scales(A: np.ndarray) -> np.ndarray:
    mx = np.abs(A.astype(np.float64)).max(axis=1) if A.shape[1] else np.zeros(A.shape[0])
    return np.where(mx > 0, mx / 448.0, 1.0).astype(np.float32)


def col_scales(B: np.ndarray) -> np.ndarray:
    mx = np.abs(B.astype(np.float64)).max(axis=0) if B.shape[0] else np.zeros(B.shape[1])
    return np.where(mx > 0, mx / 448.0, 1.0).astype(np.float32)


def quantize_a(A: np.ndarray, dtype: str):
    """(codes of every entry of A, row scales or None); zeros stay zero."""
    A = np.asarray(A, dtype=np.float32)
    if dtype == "bf16":
        return bf16_bits(A), None
    s = row_scales(A)
    q = (A.astype(np.float64) / s.astype(np.float64)[:, None]).astype(np.float32)
    return e4m3_bits(q), s


def quantize_b(B: np.ndarray, dtype: str):
    B = np.asarray(B, dtype=np.float32)
    if dtype == "bf16":
        return bf16_bits(B), None
    t = col_scales(B)
    q = (B.astype(np.float64) / t.astype(np.float64)[None, :]).astype(np.float32)
    return e4m3_bits(q), t


def dequant(codes, dtype: str) -> np.ndarray:
    return bf16_value(codes) if dtype == "bf16" else e4m3_value(codes)


def pack_blocks(A: np.ndarray, dtype: str):
    """Reference packing: (slot [P, KB], blk_kb, blocks [n, 512]) with
    blocks in MFMA operand order; only the n blocks holding a nonzero are
    stored, in (panel, K block) order, and slot[p, kb] is the stored index of
    block (p, kb) or -1 (the executor reads zeros for it)."""
    A = np.asarray(A, dtype=np.float32)
    M, K = A.shape
    P = -(-M // PANEL)
    KB = -(-K // BLOCK_K)
    codes, _ = quantize_a(A, dtype)
    pad = np.zeros((P * PANEL, KB * BLOCK_K), dtype=codes.dtype)
    pad[:M, :K] = codes
    nzpad = np.zeros((P * PANEL, KB * BLOCK_K), dtype=bool)
    nzpad[:M, :K] = A != 0
    slot = -np.ones((P, KB), dtype=np.int32)
    kb_list, blocks = [], []
    lane = np.arange(64)
    for p in range(P):
        for b in range(KB):
            if not nzpad[p * PANEL:(p + 1) * PANEL, b * BLOCK_K:(b + 1) * BLOCK_K].any():
                continue
            slot[p, b] = len(blocks)
            kb_list.append(b)
            tile = pad[p * PANEL:(p + 1) * PANEL, b * BLOCK_K:(b + 1) * BLOCK_K]
            tile = np.where(nzpad[p * PANEL:(p + 1) * PANEL, b * BLOCK_K:(b + 1) * BLOCK_K], tile, 0)
            # lane l, element e  <-  row l % 16, column 8 * (l // 16) + e
            blk = tile[(lane % 16)[:, None], (8 * (lane // 16))[:, None] + np.arange(8)[None, :]]
            blocks.append(blk.reshape(-1))
    dt = np.uint16 if dtype == "bf16" else np.uint8
    return slot, np.array(kb_list, dtype=np.int32), (np.array(blocks, dtype=dt) if blocks
                                                      else np.zeros((0, 512), dtype=dt))


# Accumulation precision of the FP8 matrix-core dot products, measured on
# gfx950 (v_mfma_f32_16x16x32_fp8_fp8): about 2^-14 relative to sum |a b| of a
# block even though the products are exact; the bound allows 2^-12.
FP8_MFMA_REL = 2.0 ** -12


def reference_f64(A: np.ndarray, B: np.ndarray, dtype: str):
    """C in FP64 from the quantised operands (what FP32 accumulation
    approximates), and the bound on |C_gpu - C| for FP32 accumulation of
    exact products over K terms: 2 * K * 2^-24 * (|qA| |qB|) scaled (FP8 adds
    FP8_MFMA_REL * (|qA| |qB|) for the matrix core's own dot-product rounding)."""
    A = np.asarray(A, dtype=np.float32)
    B = np.asarray(B, dtype=np.float32)
    qa, s = quantize_a(A, dtype)
    qb, t = quantize_b(B, dtype)
    a = dequant(qa, dtype).astype(np.float64) * (A != 0)
    b = dequant(qb, dtype).astype(np.float64)
    C = a @ b
    bound = 2.0 * max(1, A.shape[1]) * 2.0 ** -24 * (np.abs(a) @ np.abs(b))
    if dtype == "fp8":
        scale = s.astype(np.float64)[:, None] * t.astype(np.float64)[None, :]
        bound = (bound + FP8_MFMA_REL * (np.abs(a) @ np.abs(b))) * scale
        C = C * scale
        bound = bound + np.abs(C) * 2.0 ** -22
    return C, bound


def rel_error_vs_fp64(C: np.ndarray, A: np.ndarray, B: np.ndarray) -> float:
    """Normwise relative error of C against the FP64 product of the
    unquantised FP32 inputs: ||C - A B||_F / ||A B||_F."""
    ref = np.asarray(A, dtype=np.float64) @ np.asarray(B, dtype=np.float64)
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(np.asarray(C, dtype=np.float64) - ref) / (den if den > 0 else 1.0))
