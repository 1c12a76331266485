"""The progress argument of the persistent kernels (k_persist, k_persist_rows
in syncfree.hpp), checked on the CPU by simulation.

Waves claim work items from one counter in increasing order, each wave holds
its current item and the next one it claimed, and works them in claim order;
an item completes once every lower item it depends on has.  Only the waves
that are resident ever take a step: the others stand for workgroups that
other kernels keep off the device.  The claimed walk must finish whatever
subset of waves is resident.  The static walk it replaced (wave g owns items
g, g + W, g + 2W, ...) must deadlock as soon as one wave is missing -- the
planted counter-example showing that the simulation can fail.
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


def run_claimed(dag, waves: int, resident: list[int], rng, max_steps: int) -> bool:
    n = len(dag)
    done = np.zeros(n, bool)
    counter = 0
    cur = {w: None for w in resident}
    nxt = {w: None for w in resident}
    finished = set()
    for _ in range(max_steps):
        if len(finished) == len(resident):
            return bool(done.all())
        w = resident[int(rng.integers(len(resident)))]
        if w in finished:
            continue
        if cur[w] is None:              # first claim, then the next one as it starts
            cur[w], counter = counter, counter + 1
            if cur[w] >= n:
                finished.add(w)
                continue
            nxt[w], counter = counter, counter + 1
            continue
        t = cur[w]
        if all(done[d] for d in dag[t]):
            done[t] = True
            cur[w] = nxt[w]
            if cur[w] >= n:
                finished.add(w)
                continue
            nxt[w], counter = counter, counter + 1
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


@pytest.mark.parametrize("seed", range(6))
def test_claimed_walk_finishes_with_any_resident_subset(seed):
    rng = np.random.default_rng(seed)
    n, waves = 600, 32
    dag = random_dag(n, window=40, deps=4, rng=rng)
    for r in (1, 2, 7, waves):
        resident = sorted(rng.choice(waves, size=r, replace=False).tolist())
        assert run_claimed(dag, waves, resident, rng, max_steps=400 * n), (seed, r)


def test_static_walk_deadlocks_without_a_resident_wave():
    rng = np.random.default_rng(7)
    n, waves = 600, 32
    dag = random_dag(n, window=40, deps=4, rng=rng)
    assert run_static(dag, waves, list(range(waves)), rng, max_steps=400 * n)
    missing = [w for w in range(waves) if w != 3]
    assert not run_static(dag, waves, missing, rng, max_steps=400 * n)
    # the claimed walk on the same inputs finishes
    assert run_claimed(dag, waves, missing, rng, max_steps=400 * n)
