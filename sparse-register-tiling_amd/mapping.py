"""Pattern mappings: nonzero code of a panel column -> kernel patterns.

A mapping for panel height ``M_r`` sends every code ``1 .. 2^M_r - 1`` (bit r
set iff row r of the panel column is nonzero) to a list of kernel patterns
whose union covers the code.  A pattern wider than the code pads (stores
zeros); several patterns split the column into several packed columns.

This module holds the pattern-set generator and the mapping-file IO:

* ``SOPOrig`` - the greedy merge/split cover of sbench/SOP/sop_hand_configs.py:7-64
* ``save_mapping`` / ``mapping_id`` - file writer and 5-hex id
  (sbench/SOP/mapping.py:5-18)
* ``load_mapping`` - reader returning (acc, patterns, mapping) like
  sbench/SOP/sop_bench.py:72-95
* ``named_mapping`` - the names used in the reference's configs
  (tools/paper/configs.py:84-87): M_r=4 identity/orig, M_r=8 orig/alt.
"""
from __future__ import annotations

import hashlib
import json
import math
import os
from collections import OrderedDict
from typing import Callable, Dict, Iterable, List, Optional, Sequence

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
MAPPINGS_DIR = os.path.join(PKG_DIR, "mappings")


def popcount(x: int) -> int:
    return bin(int(x)).count("1")


class SOPOrig:
    """Greedy cover of a column pattern by kernel patterns.

    For a code ``p`` with at least ``min_merge`` set bits, the merge pattern
    ``q`` minimising ``split_cost*|p\\q| + merge_cost*|q\\p|`` is taken (the
    first one on ties); the bits of ``p`` it leaves out are split into single
    rows, or covered recursively by merge patterns disjoint from those already
    chosen.  Codes with fewer bits map to themselves.
    """

    def __init__(self, merge_patterns: Sequence[int], split_patterns: Sequence[int],
                 split_cost: float = 1.1, merge_cost: float = 1, min_merge: int = 2,
                 recursive: bool = False):
        self.merge_patterns = list(merge_patterns)
        self.split_patterns = list(split_patterns)
        self.split_cost = split_cost
        self.merge_cost = merge_cost
        self.min_merge = min_merge
        self.recursive = recursive

    def all_patterns(self) -> List[int]:
        return self.merge_patterns + self.split_patterns

    def map_pattern(self, pat_code: int, cant_overlap: int = 0) -> List[int]:
        pat_code = int(pat_code)
        if pat_code == 0:
            return [0]
        if popcount(pat_code) < self.min_merge:
            return [pat_code]
        best, best_cost = None, math.inf
        for q in self.merge_patterns:
            if q & cant_overlap:
                continue
            cost = self.split_cost * popcount(~q & pat_code) + \
                self.merge_cost * (popcount(q) - popcount(q & pat_code))
            if cost < best_cost:
                best, best_cost = q, cost
        if best is None:
            raise ValueError("no merge pattern is disjoint from %d" % cant_overlap)
        out = [best]
        rest = ~best & pat_code
        if rest and self.recursive:
            out += self.map_pattern(rest, cant_overlap=best | cant_overlap)
        else:
            bit = 0
            while rest:
                if rest & 1:
                    out.append(1 << bit)
                bit += 1
                rest >>= 1
        return out

    __call__ = map_pattern


def gen_all_combos(vec_height: int, nnz: int) -> List[int]:
    """All ``vec_height``-bit patterns with ``nnz`` set bits, ascending."""
    return [p for p in range(1, 1 << vec_height) if popcount(p) == nnz]


def mapping_table(M_r: int, fn: Callable[[int], Iterable[int]]) -> "OrderedDict[int, List[int]]":
    table: "OrderedDict[int, List[int]]" = OrderedDict()
    for code in range(1, 1 << M_r):
        table[code] = [int(x) for x in fn(code)]
    return table


def mapping_id(table: Dict[int, List[int]]) -> str:
    """Five-hex id: last 5 digits of md5 of the dict's repr."""
    return hashlib.md5(str(dict(table)).encode("utf-8")).hexdigest()[-5:]


def mapping_text(M_r: int, table: Dict[int, List[int]]) -> str:
    lines = ["%d\n" % M_r]
    for k, v in table.items():
        lines.append("%d: %s\n" % (k, list(v)))
    return "".join(lines)


def save_mapping(M_r: int, mapping: Callable[[int], Iterable[int]], directory: str = MAPPINGS_DIR) -> str:
    """Write ``mapping_<id>.txt`` and return the id."""
    table = mapping_table(M_r, mapping)
    mid = mapping_id(table)
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, "mapping_%s.txt" % mid)
    text = mapping_text(M_r, table)
    if not os.path.exists(path) or open(path).read() != text:
        with open(path + ".tmp", "w") as f:
            f.write(text)
        os.replace(path + ".tmp", path)
    return mid


def kernel_id_order(kernel_list: Iterable[int]) -> List[int]:
    """Kernel patterns in id order: the list the reference hands its executor
    generator (``nanokernels=list(patterns)``, sbench/SOP/sop_driver.py:93),
    stable-sorted by popcount (tools/codegen/gen_vector_specialized.py:113;
    id = index in the sorted list, :130-137).  Repeated patterns keep their
    first position."""
    seen, order = set(), []
    for p in kernel_list:
        p = int(p)
        if p and p not in seen:
            seen.add(p)
            order.append(p)
    return sorted(order, key=popcount)


def file_kernel_list(table: Dict[int, Sequence[int]]) -> List[int]:
    """The kernel list the reference builds for a mapping read from a file:
    ``load_mapping`` adds every kernel pattern to a Python ``set`` line by
    line (sbench/SOP/sop_bench.py:72-95) and ``make_sop_module`` passes
    ``list(patterns)`` on (sop_driver.py:93).  The same construct is used
    here, so the order is the interpreter's iteration order of that set."""
    pats = set()
    for v in table.values():
        for p in v:
            if int(p):
                pats.add(int(p))
    return list(pats)


class Mapping:
    """A loaded mapping: ``M_r``, the code table, its id and kernel patterns.

    ``kernel_list`` is the pattern list the reference's generator receives:
    a strategy's ``all_patterns()`` (merge then split patterns,
    sbench/SOP/sop_hand_configs.py:16-17, passed at sop_bench.py:105,154), or,
    when None, the ``list(set)`` of a mapping file (``file_kernel_list``).
    ``patterns`` is the id order derived from it (``kernel_id_order``); every
    pattern the table uses must be in the list, and listed patterns the table
    never uses still get an id, as in the reference.
    """

    def __init__(self, M_r: int, table: Dict[int, List[int]], mid: Optional[str] = None,
                 kernel_list: Optional[Sequence[int]] = None):
        self.M_r = int(M_r)
        self.table = OrderedDict((int(k), [int(x) for x in v]) for k, v in table.items())
        self.id = mid or mapping_id(self.table)
        self.kernel_list = [int(p) for p in (file_kernel_list(self.table) if kernel_list is None
                                             else kernel_list)]
        self.patterns = kernel_id_order(self.kernel_list)
        listed = set(self.patterns)
        for code, v in self.table.items():
            for p in v:
                if p and p not in listed:
                    raise ValueError("code %d maps to pattern %d, which the kernel list lacks" % (code, p))
                if p >= (1 << self.M_r):
                    raise ValueError("pattern %d does not fit M_r=%d" % (p, self.M_r))

    @classmethod
    def from_strategy(cls, M_r: int, strategy: "SOPOrig") -> "Mapping":
        """The mapping a strategy generates, with the strategy's pattern list
        (what sop_bench.py:105,154 hands the generator)."""
        return cls(M_r, mapping_table(M_r, strategy.map_pattern), kernel_list=strategy.all_patterns())

    def __call__(self, code: int) -> List[int]:
        code = int(code)
        return list(self.table.get(code, [])) if code else []

    def code_arrays(self):
        """(code_ptr[2^M_r+1], code_patterns[]) for srt_mapping_create."""
        ptr = [0]
        flat: List[int] = []
        for code in range(1 << self.M_r):
            if code:
                flat.extend(self.table.get(code, []))
            ptr.append(len(flat))
        return ptr, flat

    def __repr__(self):
        return "Mapping(M_r=%d, id=%s, %d kernel patterns)" % (self.M_r, self.id, len(self.patterns))


def read_mapping_file(path: str) -> Mapping:
    with open(path) as f:
        M_r = int(f.readline().strip())
        table: "OrderedDict[int, List[int]]" = OrderedDict()
        for line in f:
            line = line.strip()
            if not line:
                continue
            k, v = line.split(":", 1)
            table[int(k)] = [int(x) for x in json.loads(v)]
    base = os.path.basename(path)
    mid = base[len("mapping_"):-len(".txt")] if base.startswith("mapping_") else None
    return Mapping(M_r, table, mid)


def find_mapping(mid: str, search_dirs: Optional[Sequence[str]] = None) -> str:
    dirs = list(search_dirs or []) + [MAPPINGS_DIR]
    for d in dirs:
        path = os.path.join(d, "mapping_%s.txt" % mid)
        if os.path.exists(path):
            return path
    raise FileNotFoundError("mapping %s not found in %s" % (mid, dirs))


def load_mapping(mid: str, search_dirs: Optional[Sequence[str]] = None):
    """(acc, patterns, mapping) as sbench/SOP/sop_bench.py:72-95 returns them;
    ``acc`` carries M_r and the AVX-512 N_r the reference derives (16 // M_r)
    plus the GPU lane width."""
    m = read_mapping_file(find_mapping(mid, search_dirs))
    acc = {"M_r": m.M_r, "N_r": {"avx512": max(1, 16 // m.M_r), "avx2": max(1, 8 // m.M_r)}}
    return acc, set(m.patterns), m


# ----------------------------------------------------------- named strategies
_M8_MERGE = [0b01010101, 0b10101010, 0b11000011, 0b00111100, 0b00001111, 0b11110000,
             0b11111100, 0b11110011, 0b11001111, 0b00111111, 0b11111111]


def identity_strategy(M_r: int) -> SOPOrig:
    return SOPOrig([], list(range(1, 1 << M_r)), split_cost=10000, min_merge=M_r + 1)


STRATEGIES: Dict[str, Callable[[], SOPOrig]] = {
    # name -> strategy; the id each one produces is checked by the tests.
    "M2_identity": lambda: identity_strategy(2),
    "M4_identity": lambda: identity_strategy(4),
    "M6_identity": lambda: identity_strategy(6),
    "M8_identity": lambda: identity_strategy(8),
    # M_r=4 "orig": 3-of-4 patterns + 0b1111, singles split
    # (sop_hand_configs.py:99-103, merge order that yields id da01e).
    "M4_orig": lambda: SOPOrig([14, 13, 11, 7, 0b1111], [1 << i for i in range(4)], split_cost=10000),
    # M_r=4 "full": every non-empty pattern padded to the full 4-row column
    # (merge pattern 0b1111 with min_merge 1, as sop_hand_configs.py:88-93
    # builds its merge-only strategy); the packing the matrix-core executor runs.
    "M4_full": lambda: SOPOrig([0b1111], [], split_cost=10000, min_merge=1),
    # M_r=8 "orig" (sop_hand_configs.py:209-219) and "alt" (:160-169).
    "M8_orig": lambda: SOPOrig(_M8_MERGE, [1 << i for i in range(8)], split_cost=1.1, merge_cost=1,
                               recursive=True),
    "M8_alt": lambda: SOPOrig(_M8_MERGE, [1 << i for i in range(8)], split_cost=10000),
}

STRATEGY_M_R = {"M2_identity": 2, "M4_identity": 4, "M6_identity": 6, "M8_identity": 8,
                "M4_orig": 4, "M4_full": 4, "M8_orig": 8, "M8_alt": 8}

# tools/paper/configs.py:84-87, plus the identity mappings of the other
# panel heights (sbench/SOP/identity_mappings/{2,6,8}).
NAMED_IDS = {2: {"identity": "4af51"}, 4: {"identity": "61fee", "orig": "da01e", "full": "7da09"},
             6: {"identity": "001e3"}, 8: {"orig": "400fa", "alt": "747f9", "identity": "161fb"}}


def generate_shipped_mappings(directory: str = MAPPINGS_DIR) -> Dict[str, str]:
    """Write every named strategy's mapping file; returns name -> id."""
    out = {}
    for name, make in STRATEGIES.items():
        out[name] = save_mapping(STRATEGY_M_R[name], make().map_pattern, directory)
    return out


def strategy_for(M_r: int, name_or_id: str) -> Optional[str]:
    """The shipped strategy behind a config name or a mapping id, if any."""
    mid = NAMED_IDS.get(M_r, {}).get(name_or_id, name_or_id)
    for sname, make in STRATEGIES.items():
        if STRATEGY_M_R[sname] == M_r and _strategy_id(sname) == mid:
            return sname
    return None


_STRATEGY_IDS: Dict[str, str] = {}


def _strategy_id(name: str) -> str:
    if name not in _STRATEGY_IDS:
        _STRATEGY_IDS[name] = mapping_id(mapping_table(STRATEGY_M_R[name], STRATEGIES[name]().map_pattern))
    return _STRATEGY_IDS[name]


def strategy_mapping(name: str) -> Mapping:
    """Mapping of a shipped strategy, ids from the strategy's pattern list."""
    m = Mapping.from_strategy(STRATEGY_M_R[name], STRATEGIES[name]())
    return m


def named_mapping(M_r: int, name: str, search_dirs: Optional[Sequence[str]] = None) -> Mapping:
    """Mapping by the reference's config name ("identity", "orig", "alt") or by id.

    A name or id that a shipped strategy generates gets that strategy's
    kernel list (the path of sbench/SOP/sop_bench.py:105,154, which wrote the
    mapping files); any other id is read from its file and gets the file's
    kernel list (the load_mapping path, sop_bench.py:72-95)."""
    sname = strategy_for(M_r, name)
    if sname is not None:
        return strategy_mapping(sname)
    mid = NAMED_IDS.get(M_r, {}).get(name, name)
    return read_mapping_file(find_mapping(mid, search_dirs))
