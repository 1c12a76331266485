"""The progress argument of the persistent kernels (k_persist, k_persist_rows
in syncfree.hpp), checked on the CPU by simulation.

A workgroup of NW waves claims a block of NW * R consecutive work items from
one counter (blocks are handed out in increasing order); wave w of the
workgroup takes the block's items w, w + NW, ... in that order; a barrier
closes the block before the next claim.  An item completes once every lower
item it depends on has completed.  Only resident workgroups ever take a
step: the others stand for workgroups that other kernels keep off the
device.  The claimed walk must finish whatever subset of workgroups is
resident.  The static walk it replaced (wave g owns items g, g + W,
g + 2W, ...) must deadlock as soon as one wave is missing -- the planted
counter-example showing that the simulation can fail.
"""
from __future__ import annotations

import numpy as np
import pytest


def random_dag(n: int, window: int, deps: int, rng) -> list[list[int]]:
    out = []
    for t in range(n):
        lo = max(0, t - window)
        k = min(t - lo, deps)
        out.append(sorted(rng.choice(np.arange(lo, t), size=k, replace=False).tolist()) if k else [])
    return out


def run_claimed(dag, resident: list[int], nw: int, r: int, rng, max_steps: int) -> bool:
    """Block-claimed walk; resident = indices of the workgroups that run."""
    n = len(dag)
    done = np.zeros(n, bool)
    counter = 0
    block = nw * r
    # per workgroup: current block base (None: must claim) and each wave's position in it
    base = {g: None for g in resident}
    pos = {g: [0] * nw for g in resident}
    finished = set()
    for _ in range(max_steps):
        if len(finished) == len(resident):
            return bool(done.all())
        g = resident[int(rng.integers(len(resident)))]
        if g in finished:
            continue
        if base[g] is None:
            base[g], counter = counter * block, counter + 1
            if base[g] >= n:
                finished.add(g)
            pos[g] = [0] * nw
            continue
        w = int(rng.integers(nw))
        j = w + pos[g][w] * nw            # the wave's next item of the block
        if j < block and base[g] + j < n:
            t = base[g] + j
            if all(done[d] for d in dag[t]):
                done[t] = True
                pos[g][w] += 1
        # the barrier: every wave past its share of the block
        if all(w2 + pos[g][w2] * nw >= block or base[g] + w2 + pos[g][w2] * nw >= n for w2 in range(nw)):
            base[g] = None
    return False


def run_static(dag, waves: int, resident: list[int], rng, max_steps: int) -> bool:
    n = len(dag)
    done = np.zeros(n, bool)
    pos = {w: w for w in resident}
    for _ in range(max_steps):
        if all(pos[w] >= n for w in resident):
            return bool(done.all())
        w = resident[int(rng.integers(len(resident)))]
        t = pos[w]
        if t < n and all(done[d] for d in dag[t]):
            done[t] = True
            pos[w] = t + waves
    return False


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("nw,r", [(1, 1), (4, 1), (4, 3), (16, 1)])
def test_claimed_walk_finishes_with_any_resident_subset(seed, nw, r):
    rng = np.random.default_rng(seed)
    n, groups = 500, 8
    dag = random_dag(n, window=40, deps=4, rng=rng)
    for k in (1, 3, groups):
        resident = sorted(rng.choice(groups, size=k, replace=False).tolist())
        assert run_claimed(dag, resident, nw, r, rng, max_steps=2000 * n), (seed, nw, r, k)


def test_static_walk_deadlocks_without_a_resident_wave():
    rng = np.random.default_rng(7)
    n, waves = 600, 32
    dag = random_dag(n, window=40, deps=4, rng=rng)
    assert run_static(dag, waves, list(range(waves)), rng, max_steps=400 * n)
    missing = [w for w in range(waves) if w != 3]
    assert not run_static(dag, waves, missing, rng, max_steps=400 * n)
    # the claimed walk on the same inputs, one workgroup per wave, finishes
    assert run_claimed(dag, missing, 1, 1, rng, max_steps=2000 * n)
