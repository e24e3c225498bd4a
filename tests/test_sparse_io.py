"""Sparse inputs of the CG path: the Matrix Market reader and COO -> CSR (the
reference's CooMatrix::InitMarket / CsrMatrix::Init order), the reference's
random tridiagonal matrix against its committed output, CSR -> DIA, and the
oracle's FP32 CG."""
import ctypes
import pathlib

import numpy as np
import pytest

import oracle as O
import perks_amd as P

GOLD = np.load(pathlib.Path(__file__).resolve().parent / "golden" / "ref_gold.npz")


def csr_of(entries, n):
    """Expected CSR: entries (row, col, val) stable-sorted by (row, col)."""
    ent = sorted(enumerate(entries), key=lambda t: (t[1][0], t[1][1], t[0]))
    rowptr = np.zeros(n + 1, np.int32)
    for _, (r, _c, _v) in ent:
        rowptr[r + 1] += 1
    return np.cumsum(rowptr).astype(np.int32), np.array([e[1] for _, e in ent], np.int32), \
        np.array([e[2] for _, e in ent])


def test_market_general_with_duplicates_comments_and_unsorted_entries(tmp_path):
    f = tmp_path / "g.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n% a comment\n\n4 4 6\n"
                 "3 2 1.5\n1 1 -2\n3 2 0.25\n4 4 7\n1 3 1e-3\n2 1 3\n")
    rowptr, col, val = P.mm_read_csr(f)
    want = csr_of([(2, 1, 1.5), (0, 0, -2.0), (2, 1, 0.25), (3, 3, 7.0), (0, 2, 1e-3), (1, 0, 3.0)], 4)
    assert np.array_equal(rowptr, want[0]) and np.array_equal(col, want[1]) and np.array_equal(val, want[2])


@pytest.mark.parametrize("kind,sign", [("symmetric", 1.0), ("skew-symmetric", -1.0)])
def test_market_symmetric_and_skew_are_mirrored(tmp_path, kind, sign):
    f = tmp_path / "s.mtx"
    f.write_text(f"%%MatrixMarket matrix coordinate real {kind}\n3 3 3\n1 1 4\n2 1 -1\n3 2 2\n")
    rowptr, col, val = P.mm_read_csr(f)
    want = csr_of([(0, 0, 4.0), (1, 0, -1.0), (0, 1, -1.0 * sign), (2, 1, 2.0), (1, 2, 2.0 * sign)], 3)
    assert np.array_equal(rowptr, want[0]) and np.array_equal(col, want[1]) and np.array_equal(val, want[2])


def test_market_pattern_integer_and_array(tmp_path):
    f = tmp_path / "p.mtx"
    f.write_text("%%MatrixMarket matrix coordinate pattern general\n2 3 2\n2 3\n1 1\n")
    rowptr, col, val = P.mm_read_csr(f)
    assert list(rowptr) == [0, 1, 2] and list(col) == [0, 2] and list(val) == [1.0, 1.0]
    f = tmp_path / "a.mtx"
    f.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")  # column-major
    rowptr, col, val = P.mm_read_csr(f)
    assert list(rowptr) == [0, 2, 4] and list(col) == [0, 1, 0, 1] and list(val) == [1.0, 3.0, 2.0, 4.0]


def test_market_errors_are_returned(tmp_path):
    f = tmp_path / "bad.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 1\n2 2 2\n")  # more than declared
    with pytest.raises(P.PerksError):
        P.mm_read_csr(f)
    with pytest.raises(P.PerksError):
        P.mm_read_csr(tmp_path / "missing.mtx")


@pytest.mark.parametrize("tag,dt", [("f32", np.float32), ("f64", np.float64)])
def test_tridiagonal_generator_matches_reference_output(tag, dt):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(12345)
    before = libc.rand()
    libc.srand(12345)
    rowptr, col, val = P.gen_tridiag(50, dt)
    assert libc.rand() == before, "the generator leaves the process's rand() state alone"
    assert np.array_equal(rowptr, GOLD[f"tridiag_{tag}_rowptr"])
    assert np.array_equal(col, GOLD[f"tridiag_{tag}_col"])
    assert np.array_equal(val.view(np.uint8), GOLD[f"tridiag_{tag}_val"].view(np.uint8))


def test_csr_to_dia_products_agree():
    rng = np.random.default_rng(4)
    n = 200
    dense = np.zeros((n, n))
    for off in (-17, -1, 0, 3, 40):
        idx = np.arange(max(0, -off), min(n, n - off))
        dense[idx, idx + off] = rng.uniform(-1, 1, len(idx))
    rows, cols = np.nonzero(dense)
    rowptr = np.searchsorted(rows, np.arange(n + 1)).astype(np.int32)
    offs, dv = P.csr_to_dia(rowptr, cols.astype(np.int32), dense[rows, cols])
    assert list(offs) == [-17, -1, 0, 3, 40]
    x = rng.uniform(-1, 1, n)
    y = np.empty(n)
    O.lib().po_dia_spmv_f64(n, len(offs), O.p(offs.astype(np.int32)), O.p(dv), O.p(x), O.p(y))
    np.testing.assert_allclose(y, dense @ x, rtol=1e-13, atol=1e-13)
    with pytest.raises(P.PerksError):
        P.csr_to_dia(rowptr, cols.astype(np.int32), dense[rows, cols], max_diag=4)


def test_fp32_cg_oracle_solves_and_counts():
    rowptr, col, val64 = P.gen_tridiag(300, np.float64)
    val = val64.astype(np.float32)
    n = 300
    b = np.ones(n, np.float32)
    x, hist, k = O.cg_f32(0, n, rowptr, col, val, 0, b, float(np.float32(1e-5) * np.float32(1e-5)), False, 1000)
    dense = np.zeros((n, n))
    for r in range(n):
        dense[r, col[rowptr[r]:rowptr[r + 1]]] = val[rowptr[r]:rowptr[r + 1]]
    np.testing.assert_allclose(x, np.linalg.solve(dense, np.ones(n)), rtol=1e-5, atol=1e-5)
    assert hist[-1] < 1e-10 and k < 100
    _, hist2, k2 = O.cg_f32(0, n, rowptr, col, val, 0, b, 1e-30, False, 7, static_iter=True)
    assert k2 == 7 and len(hist2) == 8
