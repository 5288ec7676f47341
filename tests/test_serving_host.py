"""Host logic of the register-tiled serving form, on the CPU: the slab
budget, merging the requests of one weight into one SpMM, the split of a
batch's SpMMs between the wide table (4 floats per lane), the narrow one
(1 float per lane) and the middle one (2 floats per lane), and what happens
when the budget is full.  The backend's
per-weight records are stand-ins (no plans are built); the tables and the
slab arithmetic are the ones the GPU launches read."""
import numpy as np
import pytest
import torch


def _backend(sv, shapes, wblocks, mid_cols=0):
    """A RegTileBackend over stand-in weights: (M, K) per weight and its
    workgroups per N tile."""
    be = object.__new__(sv.RegTileBackend)
    be.merge, be.narrow_cols, be.mid_cols, be.grid = True, 64, mid_cols, 64
    be.wM = np.array([m for m, _ in shapes], dtype=np.int64)
    be.wK = np.array([k for _, k in shapes], dtype=np.int64)
    be.wblocks = np.array(wblocks, dtype=np.int64)
    be.rec = np.zeros(len(shapes), dtype=sv.EXEC_PROBLEM_DTYPE)
    be.rec["M"] = be.wM
    be.rec["K"] = be.wK
    be.rec["n_wblocks"] = be.wblocks
    return be


class _Slot:
    def __init__(self, sv, n):
        self.prob = np.zeros(n, dtype=sv.EXEC_PROBLEM_DTYPE)
        self.prob2 = np.zeros(n, dtype=sv.EXEC_PROBLEM_DTYPE)
        self.prob3 = np.zeros(n, dtype=sv.EXEC_PROBLEM_DTYPE)
        self.prep = np.zeros(n, dtype=sv.BPREP_DTYPE)
        self.hdr = np.zeros(8, dtype=np.int32)


def _server(sv, be, budget):
    srv = object.__new__(sv.Server)
    srv.backend = be
    srv.pool = sv.SlabPool(budget, torch.device("cpu"))
    return srv


def _req(sv, rid, w, K, N):
    return sv.Request(rid=rid, weight=w, B=torch.zeros((K, N), dtype=torch.bfloat16))


@pytest.fixture(scope="module")
def sv(srt):
    return srt.serving


def test_merge_and_narrow_split(sv):
    shapes = [(48, 32), (16, 64), (40, 24)]
    be = _backend(sv, shapes, [3, 5, 2])
    srv = _server(sv, be, 4 << 20)
    reqs = [_req(sv, 0, 0, 32, 10), _req(sv, 1, 1, 64, 100), _req(sv, 2, 0, 32, 30), _req(sv, 3, 1, 64, 5),
            _req(sv, 4, 2, 24, 300)]
    for r in reqs:
        assert srv.admit(r)
    slot = _Slot(sv, 8)
    be.fill_batch(srv, slot, reqs)
    m, tiles, nw, bw, nn, bn, nm, bm = (int(x) for x in slot.hdr)
    assert (nm, bm) == (0, 0)   # no middle table here
    # weight 0: 10 + 30 columns, padded to 12 + 32 = 44 <= 64: one narrow SpMM;
    # weight 1: 100 + 5 -> 100 + 8 = 108 > 64: one wide SpMM; weight 2 alone (300, wide)
    assert (nw, nn, m) == (2, 1, 5)
    wide = {int(w): r for w, r in zip(slot.prob["M"][:nw], slot.prob[:nw])}
    assert int(wide[16]["N"]) == 108 and int(wide[16]["ldb"]) == 108 and int(wide[16]["ldc"]) == 108
    assert int(wide[40]["N"]) == 300 and int(wide[40]["ldb"]) == 300
    assert int(slot.prob2[0]["M"]) == 48 and int(slot.prob2[0]["N"]) == 44
    # workgroups: per weight n_wblocks x N tiles (256 columns wide, 64 narrow)
    assert bw == 5 * 1 + 2 * 2 and bn == 3 * 1
    assert list(slot.prob["block_begin"][:nw]) == [0, 5]   # weight 1 (5 x 1), then weight 2
    # widening: one row per request, each written into its merged block at its
    # column offset with pad_to = its padded width
    pp = slot.prep[:m]
    by_src = {int(p["src"]): p for p in pp}
    r0, r2 = reqs[0], reqs[2]
    p0, p2 = by_src[r0.B.data_ptr()], by_src[r2.B.data_ptr()]
    assert (int(p0["pad_to"]), int(p2["pad_to"])) == (12, 32) and int(p0["ldk"]) == int(p2["ldk"]) == 44
    assert int(p2["dst"]) - int(p0["dst"]) == 4 * 12
    assert r0._group is r2._group and r0._group.col[id(r2)] == 12
    assert reqs[4]._group is None
    # two merged blocks plus the lone request's own; members gave theirs back
    assert srv.pool.live == 3
    for r in reqs:
        srv.release(r)
    assert srv.pool.used == 0 and srv.pool.live == 0


def test_full_budget_merges_in_the_members_blocks(sv):
    """The slab is full with the two members (16 KiB each); their merged block
    (64 x 48 floats of B and of C: a 32 KiB class) fits once they give their
    own blocks back first."""
    be = _backend(sv, [(64, 64)], [1])
    srv = _server(sv, be, 32 << 10)
    a, b = _req(sv, 0, 0, 64, 24), _req(sv, 1, 0, 64, 24)
    for r in (a, b):
        assert srv.admit(r)
    assert srv.pool.used == 32 << 10 and srv.pool.alloc(1) is None
    slot = _Slot(sv, 4)
    be.fill_batch(srv, slot, [a, b])
    assert int(slot.hdr[2]) + int(slot.hdr[4]) + int(slot.hdr[6]) == 1 and a._group is b._group
    assert a._group is not None
    assert a._off == b._off == -1 and srv.pool.live == 1 and srv.pool.used == 32 << 10
    srv.release(a)
    assert srv.pool.live == 1          # the merged block lives until its last member is released
    srv.release(b)
    assert srv.pool.used == 0


def test_members_run_alone_when_the_merged_block_cannot_fit(sv):
    """Another request holds a third 16 KiB block: freeing the members leaves
    two 16 KiB blocks, not the 32 KiB class the merged block needs, so the
    members take blocks of their class back and run alone."""
    be = _backend(sv, [(64, 64), (64, 64)], [1, 1])
    srv = _server(sv, be, 48 << 10)
    other = _req(sv, 0, 1, 64, 24)
    a, b = _req(sv, 1, 0, 64, 24), _req(sv, 2, 0, 64, 24)
    for r in (other, a, b):
        assert srv.admit(r)                # three 16 KiB blocks: the slab is full
    slot = _Slot(sv, 4)
    be.fill_batch(srv, slot, [other, a, b])
    assert int(slot.hdr[2]) + int(slot.hdr[4]) + int(slot.hdr[6]) == 3 and a._group is None and b._group is None
    assert a._off >= 0 and b._off >= 0 and srv.pool.live == 3
    rows = slot.prob2[:int(slot.hdr[4])]
    assert sorted(int(x) for x in rows["B"]) == sorted(srv.pool.base + r._bt_off for r in (other, a, b))
    for r in (other, a, b):
        srv.release(r)
    assert srv.pool.used == 0


def test_middle_widths_get_their_own_table(sv):
    """With mid_cols = 128: SpMMs of 65..128 columns go to the middle table
    (2 floats per lane, 128-column tiles), narrower ones to the narrow table,
    wider ones to the wide table."""
    shapes = [(48, 32), (16, 64), (40, 24)]
    be = _backend(sv, shapes, [3, 5, 2], mid_cols=128)
    srv = _server(sv, be, 4 << 20)
    reqs = [_req(sv, 0, 0, 32, 10), _req(sv, 1, 1, 64, 100), _req(sv, 2, 0, 32, 30), _req(sv, 3, 1, 64, 5),
            _req(sv, 4, 2, 24, 300), _req(sv, 5, 2, 24, 130)]
    for r in reqs:
        assert srv.admit(r)
    slot = _Slot(sv, 8)
    be.fill_batch(srv, slot, reqs)
    m, tiles, nw, bw, nn, bn, nm, bm = (int(x) for x in slot.hdr)
    # weight 0: 12 + 32 = 44 columns (narrow); weight 1: 100 + 8 = 108 (middle);
    # weight 2: 300 + 132 = 432 (wide)
    assert (nw, nn, nm, m) == (1, 1, 1, 6)
    assert int(slot.prob[0]["M"]) == 40 and int(slot.prob[0]["N"]) == 432
    assert int(slot.prob2[0]["M"]) == 48 and int(slot.prob2[0]["N"]) == 44
    assert int(slot.prob3[0]["M"]) == 16 and int(slot.prob3[0]["N"]) == 108
    # workgroups: n_wblocks x N tiles of 256, 64 and 128 columns
    assert (bw, bn, bm) == (2 * 2, 3 * 1, 5 * 1)
    assert int(slot.prob3[0]["block_begin"]) == 0
    for r in reqs:
        srv.release(r)
    assert srv.pool.used == 0 and srv.pool.live == 0
