import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library on the device)")
    config.addinivalue_line("markers", "slow: large inputs (full benchmark sizes)")


def lower_from_rows(rows, rng=None, dominant=True):
    """CSR (ia, ja, a) from a list of off-diagonal column lists."""
    import numpy as np
    rng = rng or np.random.default_rng(0)
    ia, ja, a = [0], [], []
    for i, cols in enumerate(rows):
        vals = rng.uniform(-1, 1, size=len(cols))
        ja += list(cols) + [i]
        a += vals.tolist() + [1.0 + float(np.abs(vals).sum()) if dominant else 1.0]
        ia.append(len(ja))
    return (np.array(ia, np.int32), np.array(ja, np.int32), np.array(a, np.float64))


@pytest.fixture(scope="session")
def golden():
    import json

    import numpy as np
    g = ROOT / "tests" / "golden"
    return {"hand": json.loads((g / "hand_expected.json").read_text()),
            "random": dict(np.load(g / "random_small.npz")),
            "dir": g}
