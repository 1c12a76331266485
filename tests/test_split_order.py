"""The walk order of split long rows, checked on the CPU.

The analysis (analysis.hip, build_segments) cuts every row of the 32-lane
class with more than 256 off-diagonal entries: the row keeps its first 128
entries and each further run of 256 becomes a segment item, placed in the
multirow walk right before the wave that holds the row.  The kernels rely on
one property: every item waits only on LOWER items (segments on the waves of
the rows their entries depend on, a wave on those waves and on its own
segments).  That is what lets the ticketed walk and the block-claimed
persistent walk keep their deadlock-freedom arguments.

Here the walk order is restated from the oracle's 64-lane wave layout, the
property is checked on random matrices with long rows and chains of them, and
the block-claimed walk is simulated on the resulting item DAG with arbitrary
resident subsets (it must finish).  A planted mistake -- segments placed
after their wave -- must break the property.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_py as O
from conftest import lower_from_rows
from test_claim_order import run_claimed

SEG_MIN, SEG_OWN, SEG_ENTRIES = 256, 128, 256


def seg_count(off: int) -> int:
    return -(-(off - SEG_OWN) // SEG_ENTRIES) if off > SEG_MIN else 0


def random_rows(n, long_rows, seed):
    rng = np.random.default_rng(seed)
    rows, prev = [], None
    for i in range(n):
        if i in long_rows:
            m = min(long_rows[i], i)
            c = rng.choice(i, size=m, replace=False)
            if prev is not None and prev not in c:
                c[0] = prev
            prev = i
        else:
            lo = max(0, i - 60)
            m = min(int(rng.poisson(4)), i - lo)
            c = rng.choice(np.arange(lo, i), size=m, replace=False) if m else np.empty(0, np.int64)
        rows.append(sorted(int(v) for v in c))
    return rows


def walk_items(ia, ja, segments_first=True):
    """Items of the split multirow walk and each item's dependencies (item indices)."""
    n = len(ia) - 1
    an = O.analysis(ia, ja, 64)
    io, base, vs = an["iorder"], an["ibase_row"], an["ivect_size_warp"]
    nw = an["n_warps"]
    items = []          # ("wave", w) or ("seg", row, lo, hi)
    wave_rows = []
    for w in range(nw):
        rows = [int(r) for r in io[base[w]:base[w + 1]]]
        wave_rows.append(rows)
        segs = []
        if vs[w] == 32:
            for r in rows:
                rs, dpos = int(ia[r]), int(ia[r + 1]) - 1
                if seg_count(dpos - rs):
                    for lo in range(rs + SEG_OWN, dpos, SEG_ENTRIES):
                        segs.append(("seg", r, lo, min(lo + SEG_ENTRIES, dpos)))
        order = segs + [("wave", w)] if segments_first else [("wave", w)] + segs
        items += order
    pos = {}
    for t, it in enumerate(items):
        if it[0] == "wave":
            for r in wave_rows[it[1]]:
                pos[("row", r)] = t
        else:
            pos[("seg", it[1], it[2])] = t
    deps = []
    for it in items:
        d = set()
        if it[0] == "seg":
            d.update(pos[("row", int(c))] for c in ja[it[2]:it[3]])
        else:
            for r in wave_rows[it[1]]:
                rs, dpos = int(ia[r]), int(ia[r + 1]) - 1
                split = seg_count(dpos - rs) and vs[it[1]] == 32
                own_end = rs + SEG_OWN if split else dpos
                d.update(pos[("row", int(c))] for c in ja[rs:own_end])
                if split:
                    d.update(pos[("seg", r, lo)] for lo in range(rs + SEG_OWN, dpos, SEG_ENTRIES))
        deps.append(sorted(d))
    return items, deps


LONG = {400: 256, 401: 257, 402: 384, 403: 385, 900: 700, 1500: 1400, 1501: 300, 2500: 2000}


@pytest.mark.parametrize("seed", range(3))
def test_every_item_waits_only_on_lower_items(seed):
    ia, ja, _ = lower_from_rows(random_rows(3000, LONG, seed))
    items, deps = walk_items(ia, ja)
    nseg = sum(1 for it in items if it[0] == "seg")
    off = np.diff(ia) - 1
    assert nseg == sum(seg_count(int(o)) for o in off) > 0
    for t, d in enumerate(deps):
        assert all(u < t for u in d), (t, items[t])


@pytest.mark.parametrize("seed", range(2))
@pytest.mark.parametrize("nw,r", [(4, 1), (16, 1), (4, 2)])
def test_block_claimed_walk_finishes_on_split_items(seed, nw, r):
    ia, ja, _ = lower_from_rows(random_rows(2000, LONG, seed))
    _, deps = walk_items(ia, ja)
    rng = np.random.default_rng(seed)
    for k in (1, 3, 8):
        resident = sorted(rng.choice(8, size=k, replace=False).tolist())
        assert run_claimed(deps, resident, nw, r, rng, max_steps=400 * len(deps)), (seed, nw, r, k)


def test_segments_after_their_wave_break_the_order():
    ia, ja, _ = lower_from_rows(random_rows(3000, LONG, 0))
    _, deps = walk_items(ia, ja, segments_first=False)
    assert any(any(u > t for u in d) for t, d in enumerate(deps))
