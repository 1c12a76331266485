"""Writes tests/golden/random_small.npz: small random lower-triangular systems
with expected outputs computed by methods independent of the oracle:

  x      scipy.sparse.linalg.spsolve_triangular (LAPACK-style library solve)
  level  Kahn peeling: level k = rows whose dependencies were all removed in
         peels 1..k-1 (a different algorithm from the row recurrence)
  iorder Python's stable sorted() on (7*(level-1)+class, row), with the class
         table of analysis_csr.cu:12-31 written out as a dict lookup

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve_triangular

OUT = Path(__file__).with_name("random_small.npz")

CLASS_OF = {0: 6, 1: 0, 2: 1, 3: 2, 4: 2, 5: 3, 6: 3, 7: 3, 8: 3}


def klass(nz: int) -> int:
    if nz in CLASS_OF:
        return CLASS_OF[nz]
    return 4 if nz <= 16 else 5


def random_lower(n: int, dens: int, rng: np.random.Generator, maxrow: int = 40):
    ia = [0]
    ja, a = [], []
    for i in range(n):
        k = min(i, int(rng.integers(0, dens + 1)) if rng.random() > 0.05 else int(rng.integers(0, maxrow)))
        cols = sorted(rng.choice(i, size=k, replace=False).tolist()) if k else []
        vals = rng.uniform(-1, 1, size=len(cols))
        ja += cols + [i]
        a += vals.tolist() + [1.0 + float(np.abs(vals).sum())]
        ia.append(len(ja))
    return np.array(ia, np.int32), np.array(ja, np.int32), np.array(a, np.float64)


def kahn_levels(n, ia, ja):
    deps = [set(ja[ia[i]:ia[i + 1] - 1].tolist()) for i in range(n)]
    users = [[] for _ in range(n)]
    for i in range(n):
        for j in deps[i]:
            users[j].append(i)
    remaining = [len(d) for d in deps]
    level = [0] * n
    frontier = [i for i in range(n) if remaining[i] == 0]
    k = 0
    while frontier:
        k += 1
        nxt = []
        for i in frontier:
            level[i] = k
        for i in frontier:
            for u in users[i]:
                remaining[u] -= 1
                if remaining[u] == 0:
                    nxt.append(u)
        frontier = nxt
    return np.array(level, np.int32)


def main() -> None:
    rng = np.random.default_rng(20261202)
    data = {}
    for case, (n, dens) in enumerate([(37, 3), (300, 6), (1500, 12)]):
        ia, ja, a = random_lower(n, dens, rng)
        b = rng.uniform(-1, 1, size=n)
        Lm = sp.csr_matrix((a, ja, ia), shape=(n, n))
        x = spsolve_triangular(Lm, b, lower=True)
        lev = kahn_levels(n, ia, ja)
        key = [7 * (lev[i] - 1) + klass(int(ia[i + 1] - ia[i] - 1)) for i in range(n)]
        iorder = np.array(sorted(range(n), key=lambda i: (key[i], i)), np.int32)
        for name, v in dict(ia=ia, ja=ja, a=a, b=b, x=x, level=lev, key=np.array(key, np.int32),
                            iorder=iorder).items():
            data[f"c{case}_{name}"] = v
    np.savez(OUT, **data)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
