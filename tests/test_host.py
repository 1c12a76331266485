"""Host-side tests that need no GPU: the C-ABI library loads and exports every
entry point of include/sptrsv_amd.h, input validation, the Matrix Market
loader against the oracle and the hand fixtures, and the synthetic
generators."""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest

import oracle_py as O
from sptrsv_amd import _lib as L
from sptrsv_amd import interface as I

ROOT = Path(__file__).resolve().parents[1]


def header_functions() -> set[str]:
    h = (ROOT / "include" / "sptrsv_amd.h").read_text()
    h = re.sub(r"/\*.*?\*/", "", h, flags=re.S)
    return set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_]\w*\s*\**\s+\**([A-Za-z_]\w*)\s*\(", h, flags=re.M))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    declared = header_functions()
    assert len(declared) > 40
    missing = [f for f in sorted(declared) if not hasattr(lib, f)]
    assert not missing, missing
    assert set(L.EXPORTED) == declared


def test_status_strings():
    lib = L.lib()
    for st in range(0, -12, -1):
        assert lib.sptrsv_status_string(st).decode() not in ("", "unknown status")
    assert lib.sptrsv_status_string(-99).decode() == "unknown status"


@pytest.mark.parametrize("ia,ja,why", [
    ([0, 1, 3], [0, 1, 0], "diagonal not last"),
    ([0, 1, 1], [0], "empty row"),
    ([0, 2], [1, 0], "upper entry"),
    ([0, 1, 3], [0, -1, 1], "negative column"),
])
def test_allocate_rejects_invalid_csr(ia, ja, why):
    # validation happens on the host before any device call
    with pytest.raises(L.SptrsvError) as e:
        I.allocate_memory(np.array(ia), np.array(ja), np.ones(len(ja)))
    assert e.value.status == L.E_INVALID_CSR, why


@pytest.mark.parametrize("ia,ja,why", [
    ([0, 2, 3], [1, 0, 1], "diagonal not first"),
    ([0, 1, 1], [0], "empty row"),
    ([0, 1, 3], [0, 0, 1], "lower entry"),
    ([0, 2, 3], [0, 2, 1], "column past the last row"),
])
def test_allocate_upper_rejects_invalid_csr(ia, ja, why):
    with pytest.raises(L.SptrsvError) as e:
        I.allocate_memory_upper(np.array(ia), np.array(ja), np.ones(len(ja)))
    assert e.value.status == L.E_INVALID_CSR, why


@pytest.mark.parametrize("name", ["tiny_general", "tiny_symmetric"])
@pytest.mark.parametrize("unit", [True, False])
def test_mtx_loader_matches_oracle_and_fixture(name, unit, golden):
    path = golden["dir"] / f"{name}.mtx"
    n, ia, ja, a = I.read_mtx_lower(str(path), unit=unit)
    m, ri, ci, v, sym = O.read_mtx_coo(path)
    rp, col, val = O.coo_to_csr(m, ri, ci, v, sym)
    lp, lc, lv = O.extract_lower(m, rp, col, val, unit=unit)
    assert n == m
    assert np.array_equal(ia, lp) and np.array_equal(ja, lc) and np.array_equal(a, lv)
    e = golden["hand"][name]
    assert ia.tolist() == e["L_ia"] and ja.tolist() == e["L_ja"]
    if not unit:
        assert a.tolist() == e["L_val_keep"]
    else:
        assert np.all(a == 1.0)


def test_mtx_loader_pattern(golden):
    n, ia, ja, a = I.read_mtx_lower(str(golden["dir"] / "tiny_pattern.mtx"))
    assert ia.tolist() == golden["hand"]["tiny_pattern"]["L_ia"]
    assert ja.tolist() == golden["hand"]["tiny_pattern"]["L_ja"]


def test_mtx_loader_errors(tmp_path):
    bad = tmp_path / "c.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 1.0 0.0\n")
    with pytest.raises(L.SptrsvError) as e:
        I.read_mtx_lower(str(bad))
    assert e.value.status == L.E_IO
    rect = tmp_path / "r.mtx"
    rect.write_text("%%MatrixMarket matrix coordinate real general\n2 3 1\n1 1 1.0\n")
    with pytest.raises(L.SptrsvError):
        I.read_mtx_lower(str(rect))
    oob = tmp_path / "o.mtx"
    oob.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n")
    with pytest.raises(L.SptrsvError):
        I.read_mtx_lower(str(oob))
    short = tmp_path / "s.mtx"
    short.write_text("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n")
    with pytest.raises(L.SptrsvError):
        I.read_mtx_lower(str(short))


def test_mtx_loader_random_symmetric_vs_scipy(tmp_path):
    import scipy.io
    import scipy.sparse as sp
    rng = np.random.default_rng(3)
    A = sp.random(60, 60, density=0.1, random_state=4, format="coo")
    A = (A + A.T).tocoo()
    path = tmp_path / "s.mtx"
    scipy.io.mmwrite(str(path), A, symmetry="symmetric")
    n, ia, ja, a = I.read_mtx_lower(str(path), unit=False)
    m, ri, ci, v, sym = O.read_mtx_coo(path)
    lp, lc, lv = O.extract_lower(m, *O.coo_to_csr(m, ri, ci, v, sym), unit=False)
    assert np.array_equal(ia, lp) and np.array_equal(ja, lc) and np.array_equal(a, lv)
    assert O.check_lower_csr(ia, ja) == 0
    del rng


def test_generator_config1_known_answer():
    s = I.generate(1)
    assert s.n == 4096 and s.nnz == 12160
    assert O.check_lower_csr(s.ia, s.ja) == 0
    assert O.validate_x(O.fwd_f64(s.ia, s.ja, s.a, s.b)) == 0
    _, nl = O.levels(s.ia, s.ja)
    assert nl == 127


def test_generator_config4_structure():
    s = I.generate(4, 16)
    assert s.n == 16 ** 3 and s.k == 16 and s.b.size == s.n * 16
    assert O.check_lower_csr(s.ia, s.ja) == 0
    _, nl = O.levels(s.ia, s.ja)
    assert nl == 3 * 15 + 1
    assert np.all(s.a[s.ia[1:] - 1] == 1.0)


def test_generator_config2_statistics_and_determinism():
    s = I.generate(2, 40000, seed=2)
    t = I.generate(2, 40000, seed=2)
    u = I.generate(2, 40000, seed=9)
    assert np.array_equal(s.ja, t.ja) and np.array_equal(s.a, t.a) and np.array_equal(s.b, t.b)
    assert not np.array_equal(s.a[:1000], u.a[:1000])
    assert O.check_lower_csr(s.ia, s.ja) == 0
    deg = np.diff(s.ia) - 1
    assert abs(deg[6000:].mean() - 8.0) < 0.1          # Poisson(8) away from the first rows
    rows = np.repeat(np.arange(s.n), np.diff(s.ia))
    off = s.ja != rows
    assert np.all(rows[off] - s.ja[off] <= 5000)         # window [i-5000, i)
    d = s.a[s.ia[1:] - 1]
    absum = np.add.reduceat(np.abs(s.a), s.ia[:-1]) - d
    assert np.allclose(d, 1 + absum)


def test_generator_config3_statistics():
    s = I.generate(3, 50000, seed=3)
    assert O.check_lower_csr(s.ia, s.ja) == 0
    deg = np.diff(s.ia) - 1
    assert deg[1:].min() >= 1 and deg.max() <= 2000
    assert 7.0 < deg[5000:].mean() < 13.0                 # Zipf(1.8) capped at 2000: mean ~9.8


def test_generator_batch_members():
    b = I.generate_batch_member(0, 2000, 5000, seed=5)
    r = I.generate_batch_member(1, 2000, 5000, seed=5)
    for s in (b, r):
        assert 2000 <= s.n <= 5000 and O.check_lower_csr(s.ia, s.ja) == 0
    deg = np.diff(b.ia) - 1
    bw = deg.max()
    assert 1 <= bw <= 32 and np.all(deg[bw:] == bw)      # dense band of width bw
    _, nl = O.levels(b.ia, b.ja)
    assert nl == b.n                                       # chain depth n


# ---- parallel COO reader and the binary CSR cache ---------------------------------

def write_mtx(path, n, ri, ci, v, sym=False, pattern=False, noise=False, crlf=False, seed=0):
    """A coordinate file of 1-based entries, optionally with comment and blank
    lines between entries and CRLF line ends (all skipped by the format)."""
    rng = np.random.default_rng(seed)
    field = "pattern" if pattern else "real"
    lines = [f"%%MatrixMarket matrix coordinate {field} {'symmetric' if sym else 'general'}",
             "% generated", f"{n} {n} {len(ri)}"]
    for r, c, x in zip(ri, ci, v):
        if noise and rng.random() < 0.01:
            lines.append("% a comment between entries" if rng.random() < 0.5 else "")
        lines.append(f"{r + 1} {c + 1}" if pattern else f"{r + 1} {c + 1} {float(x)!r}")
    end = "\r\n" if crlf else "\n"
    path.write_text(end.join(lines) + end)


def random_coo(n, nz, rng, sym=False):
    ri = rng.integers(0, n, nz).astype(np.int32)
    ci = rng.integers(0, n, nz).astype(np.int32)
    if sym:  # symmetric files hold one triangle
        ri, ci = np.maximum(ri, ci), np.minimum(ri, ci)
    return ri, ci, rng.uniform(-2, 2, nz)


@pytest.mark.parametrize("name", ["tiny_general", "tiny_symmetric", "tiny_pattern"])
@pytest.mark.parametrize("threads", [1, 3, 16])
def test_mtx_coo_reader_fixtures(name, threads, golden):
    path = golden["dir"] / f"{name}.mtx"
    n, r, c, v, sym = I.read_mtx_coo(str(path), threads)
    m, ri, ci, vi, s = O.read_mtx_coo(path)
    assert (n, sym) == (m, s)
    assert np.array_equal(r, ri) and np.array_equal(c, ci) and np.array_equal(v, vi)


@pytest.mark.parametrize("sym,pattern,crlf", [(False, False, False), (True, False, True), (False, True, False)])
def test_mtx_coo_reader_threads_agree(tmp_path, sym, pattern, crlf):
    rng = np.random.default_rng(7)
    n, nz = 3000, 60000          # ~1.3 MB: many line-aligned chunks
    ri, ci, v = random_coo(n, nz, rng, sym)
    path = tmp_path / "m.mtx"
    write_mtx(path, n, ri, ci, v, sym=sym, pattern=pattern, noise=True, crlf=crlf)
    want = v if not pattern else np.ones(nz)
    for threads in (1, 2, 7, 16):
        m, r, c, x, s = I.read_mtx_coo(str(path), threads)
        assert (m, s) == (n, sym)
        assert np.array_equal(r, ri) and np.array_equal(c, ci) and np.array_equal(x, want), threads
    # the lower triangle built from the parallel read matches the oracle's pipeline
    n2, ia, ja, a = I.read_mtx_lower(str(path), unit=False)
    rp, col, val = O.coo_to_csr(n, ri, ci, want, sym)
    lp, lc, lv = O.extract_lower(n, rp, col, val, unit=False)
    assert np.array_equal(ia, lp) and np.array_equal(ja, lc) and np.array_equal(a, lv)


def test_mtx_coo_reader_errors_in_later_chunks(tmp_path):
    rng = np.random.default_rng(2)
    n, nz = 2000, 40000
    ri, ci, v = random_coo(n, nz, rng)
    good = tmp_path / "g.mtx"
    write_mtx(good, n, ri, ci, v)
    text = good.read_text().splitlines()
    # a malformed line near the end, before the last entry: premature end
    bad = tmp_path / "b.mtx"
    bad.write_text("\n".join(text[:-10] + ["7 x 1.0"] + text[-10:]) + "\n")
    for threads in (1, 8):
        with pytest.raises(L.SptrsvError) as e:
            I.read_mtx_coo(str(bad), threads)
        assert e.value.status == L.E_IO and "premature" in L.last_error()
    # the same line after the last counted entry is never read
    after = tmp_path / "a.mtx"
    after.write_text("\n".join(text + ["7 x 1.0", "999999 1 1.0"]) + "\n")
    for threads in (1, 8):
        assert np.array_equal(I.read_mtx_coo(str(after), threads)[1], ri)
    # an index out of range deep in the file names its entry
    oob = tmp_path / "o.mtx"
    oob.write_text("\n".join(text[:-5] + [f"{n + 1} 1 1.0"] + text[-4:]) + "\n")
    for threads in (1, 8):
        with pytest.raises(L.SptrsvError):
            I.read_mtx_coo(str(oob), threads)
        assert f"entry {nz - 4}" in L.last_error()
    # fewer entries than announced
    short = tmp_path / "s.mtx"
    short.write_text("\n".join(text[:-3]) + "\n")
    with pytest.raises(L.SptrsvError):
        I.read_mtx_coo(str(short), 8)


def test_csr_cache_round_trip_and_corruption(tmp_path):
    s = I.generate(2, 20000)
    path = tmp_path / "m.csr"
    I.save_csr(str(path), s.ia, s.ja, s.a)
    n, ia, ja, a = I.load_csr(str(path))
    assert n == s.n and np.array_equal(ia, s.ia) and np.array_equal(ja, s.ja) and np.array_equal(a, s.a)
    raw = bytearray(path.read_bytes())
    raw[-3] ^= 0x40                       # one flipped bit in the values
    bad = tmp_path / "bad.csr"
    bad.write_bytes(bytes(raw))
    with pytest.raises(L.SptrsvError) as e:
        I.load_csr(str(bad))
    assert e.value.status == L.E_IO and "checksum" in L.last_error()
    cut = tmp_path / "cut.csr"
    cut.write_bytes(bytes(raw[:-100]))
    with pytest.raises(L.SptrsvError):
        I.load_csr(str(cut))
    notcsr = tmp_path / "x.csr"
    notcsr.write_bytes(b"%%MatrixMarket matrix coordinate real general\n")
    with pytest.raises(L.SptrsvError):
        I.load_csr(str(notcsr))
    empty = tmp_path / "e.csr"
    I.save_csr(str(empty), np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    n, ia, ja, a = I.load_csr(str(empty))
    assert n == 0 and ia.tolist() == [0] and len(ja) == 0
