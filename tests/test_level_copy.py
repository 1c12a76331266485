"""The level-ordered copy, checked on the CPU with the oracle.

The library's reorder = 1 solves walk a copy of L whose row p is row
iorder[p] of L (entries in storage order, every column c renamed
inv_iorder[c]) with the copy's own analysis, and claim results bit-identical
to solving L itself.  That rests on three facts, each checked here on random
matrices (short rows, long rows, chains) for both lane widths the analysis
builds:
  * the copy is lower triangular with the diagonal last (every dependency is
    at an earlier level, hence earlier in iorder);
  * its analysis is the original's renamed: iorder the identity, levels and
    keys permuted, buckets and warp arrays equal (the level sort is stable);
  * the oracle's forward substitution on the copy gives the original's x,
    permuted, bit for bit (same entries folded in the same order).
Planted mistakes must break them: entries re-sorted by their new column
change the fold order, and rows copied in level order alone (classes mixed)
do not come back as the identity iorder.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_py as O
from conftest import lower_from_rows


def random_rows(n, seed, long_every=0):
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        if long_every and i > 300 and i % long_every == 0:
            m = min(i, int(rng.integers(40, 300)))
            c = rng.choice(i, size=m, replace=False)
        else:
            lo = max(0, i - 80)
            m = min(int(rng.poisson(5)), i - lo)
            c = rng.choice(np.arange(lo, i), size=m, replace=False) if m else np.empty(0, np.int64)
        rng.shuffle(c)                          # storage order is not column order
        rows.append([int(v) for v in c])
    return rows


def level_copy(ia, ja, a, perm, sort_entries=False):
    """Row p = row perm[p], columns renamed through the inverse permutation."""
    n = len(ia) - 1
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    lens = np.diff(ia)[perm]
    nia = np.zeros(n + 1, np.int64)
    nia[1:] = np.cumsum(lens)
    nja = np.empty(len(ja), np.int32)
    na = np.empty(len(a), np.float64)
    for p, r in enumerate(perm):
        s, e, d = ia[r], ia[r + 1], nia[p]
        cols, vals = inv[ja[s:e]], a[s:e]
        if sort_entries:
            order = np.argsort(cols, kind="stable")
            cols, vals = cols[order], vals[order]
        nja[d:d + e - s] = cols
        na[d:d + e - s] = vals
    return nia.astype(np.int32), nja, na


CASES = [(2000, 0, 0), (3000, 1, 97), (1500, 2, 13)]


@pytest.mark.parametrize("n,seed,long_every", CASES)
@pytest.mark.parametrize("lanes", [32, 64])
def test_copy_analysis_is_the_original_renamed(n, seed, long_every, lanes):
    ia, ja, a = lower_from_rows(random_rows(n, seed, long_every))
    an = O.analysis(ia, ja, lanes)
    perm = an["iorder"]
    nia, nja, na = level_copy(ia, ja, a, perm)
    # lower triangular, diagonal last
    for p in range(n):
        s, e = nia[p], nia[p + 1]
        assert nja[e - 1] == p
        assert np.all(nja[s:e - 1] < p)
    ac = O.analysis(nia, nja, lanes)
    assert np.array_equal(ac["iorder"], np.arange(n))
    assert np.array_equal(ac["level"], an["level"][perm])
    assert np.array_equal(ac["key"], an["key"][perm])
    assert ac["nlevs"] == an["nlevs"] and ac["n_warps"] == an["n_warps"]
    for f in ("bucket_off", "warp_off", "ibase_row", "ivect_size_warp"):
        assert np.array_equal(ac[f], an[f]), f


@pytest.mark.parametrize("n,seed,long_every", CASES)
def test_copy_solves_bit_identically(n, seed, long_every):
    ia, ja, a = lower_from_rows(random_rows(n, seed, long_every))
    perm = O.analysis(ia, ja, 64)["iorder"]
    b = np.random.default_rng(seed + 10).standard_normal(n)
    x = O.fwd_f64(ia, ja, a, b)
    nia, nja, na = level_copy(ia, ja, a, perm)
    xc = O.fwd_f64(nia, nja, na, b[perm])
    assert np.array_equal(xc.view(np.uint64), x[perm].view(np.uint64))


def test_planted_mistakes_break_the_copy():
    ia, ja, a = lower_from_rows(random_rows(3000, 1, 97))
    an = O.analysis(ia, ja, 64)
    perm = an["iorder"]
    b = np.random.default_rng(11).standard_normal(len(ia) - 1)
    x = O.fwd_f64(ia, ja, a, b)
    # entries re-sorted by their new column: another fold order
    sia, sja, sa = level_copy(ia, ja, a, perm, sort_entries=True)
    xs = O.fwd_f64(sia, sja, sa, b[perm])
    assert not np.array_equal(xs.view(np.uint64), x[perm].view(np.uint64))
    # rows in level order only (classes mixed within a level): still lower
    # triangular, but its own analysis reorders it, so row p would not be
    # walked at position p
    lev_only = np.argsort(an["level"], kind="stable").astype(np.int32)
    ria, rja, _ = level_copy(ia, ja, a, lev_only)
    assert not np.array_equal(O.analysis(ria, rja, 64)["iorder"], np.arange(len(ia) - 1))
