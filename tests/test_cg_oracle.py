"""Pins the oracle's CG before it is used as the checker.

The reference's CG orders its dot products through atomics (cg_kernels.cuh:
201-223, 340-372), so no reference run fixes a residual history; the oracle's
summation order is this build's own (documented in perks_oracle.h).  What the
mathematics fixes is checked here: the Poisson operator against a dense
assembly, CSR and DIA products agree bit for bit, the CG solution against a
dense solve and a manufactured solution, the fixed-order dot against an exactly
rounded sum, and CG's finite-termination property on a tiny system.
"""
import math

import numpy as np
import pytest

import oracle as O


def dense_poisson(nx, ny, nz):
    n = nx * ny * nz
    A = np.zeros((n, n))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = (z * ny + y) * nx + x
                A[i, i] = 6.0
                for dz, dy, dx in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                    zz, yy, xx = z + dz, y + dy, x + dx
                    if 0 <= zz < nz and 0 <= yy < ny and 0 <= xx < nx:
                        A[i, (zz * ny + yy) * nx + xx] = -1.0
    return A


def test_poisson_csr_matches_dense_assembly():
    nx, ny, nz = 5, 4, 3
    rowptr, col, val = O.poisson7_csr(nx, ny, nz)
    A = dense_poisson(nx, ny, nz)
    n = nx * ny * nz
    B = np.zeros((n, n))
    for i in range(n):
        cols = col[rowptr[i]:rowptr[i + 1]]
        assert np.all(np.diff(cols) > 0), "columns ascending"
        B[i, cols] = val[rowptr[i]:rowptr[i + 1]]
    assert np.array_equal(A, B)
    assert O.lib().po_poisson7_nnz(256, 256, 256) == 117047296  # 7N - 6*256^2


def test_csr_and_dia_products_are_bitwise_equal():
    nx, ny, nz = 9, 7, 6
    n = nx * ny * nz
    rowptr, col, val = O.poisson7_csr(nx, ny, nz)
    offs, dval = O.poisson7_dia(nx, ny, nz)
    x = O.uniform(n, 3, -1.0, 1.0)
    y1, y2 = np.empty(n), np.empty(n)
    O.lib().po_csr_spmv_f64(n, O.p(rowptr), O.p(col), O.p(val), O.p(x), O.p(y1))
    O.lib().po_dia_spmv_f64(n, 7, O.p(offs), O.p(dval), O.p(x), O.p(y2))
    assert np.array_equal(y1, y2)
    np.testing.assert_allclose(y1, dense_poisson(nx, ny, nz) @ x, rtol=1e-14, atol=1e-14)


def test_fixed_order_dot_is_close_to_exact_sum():
    for n in (1, 255, 4096, 4097, 100000):
        a = O.uniform(n, n, -1.0, 1.0)
        b = O.uniform(n, n + 1, -1.0, 1.0)
        exact = math.fsum(float(x) * float(y) for x, y in zip(a, b))
        got = O.dot(a, b)
        assert abs(got - exact) <= 1e-15 * max(1.0, math.fsum(abs(float(x) * float(y)) for x, y in zip(a, b)))
    assert O.dot(np.zeros(0), np.zeros(0)) == 0.0


def test_cg_matches_dense_solve_and_terminates():
    nx, ny, nz = 4, 3, 3
    n = nx * ny * nz
    rowptr, col, val = O.poisson7_csr(nx, ny, nz)
    b = O.uniform(n, 9, -1.0, 1.0)
    x, hist, k = O.cg(0, n, rowptr, col, val, 0, b, 1e-13, True, 200)
    want = np.linalg.solve(dense_poisson(nx, ny, nz), b)
    np.testing.assert_allclose(x, want, rtol=0, atol=1e-11)
    assert k <= n  # finite termination (exact arithmetic: at most n steps)
    assert hist[-1] < (1e-13) ** 2 * O.dot(b, b)


def test_cg_manufactured_solution_dia():
    nx, ny, nz = 10, 9, 8
    n = nx * ny * nz
    offs, dval = O.poisson7_dia(nx, ny, nz)
    xs = O.uniform(n, 5, -1.0, 1.0)
    b = np.empty(n)
    O.lib().po_dia_spmv_f64(n, 7, O.p(offs), O.p(dval), O.p(xs), O.p(b))
    x, hist, k = O.cg(1, n, None, offs, dval, 7, b, 1e-12, True, 1000)
    np.testing.assert_allclose(x, xs, rtol=0, atol=1e-9)
    # Poisson eigenvalues 6 - 2(cos + cos + cos): the iteration count stays within the sqrt(kappa) bound
    lam = [6 - 2 * sum(math.cos(math.pi * k_ / (m + 1)) for k_, m in zip(ks, (nx, ny, nz)))
           for ks in ((1, 1, 1), (nx, ny, nz))]
    kappa = lam[1] / lam[0]
    assert k <= math.ceil(0.5 * math.sqrt(kappa) * math.log(2 / 1e-12)) + 5


def test_refinement_reaches_fp64_tolerance():
    nx, ny, nz = 12, 10, 9
    n = nx * ny * nz
    rowptr, col, val = O.poisson7_csr(nx, ny, nz)
    b = O.uniform(n, 13, -1.0, 1.0)
    x, hist, total, outer = O.cg_f32_refine(0, n, rowptr, col, val, 0, b, 1e-10, True, 1e-4, 500, 20)
    assert hist[-1] < 1e-20 * O.dot(b, b)
    assert 2 <= outer <= 8
    want, _, _ = O.cg(0, n, rowptr, col, val, 0, b, 1e-13, True, 2000)
    np.testing.assert_allclose(x, want, rtol=0, atol=1e-8)


# ---- the CSR row order (po_csr_spmv_*) ----------------------------------------------
def _row(values):
    """One CSR row holding `values` against x = 1 (products = values)."""
    n = len(values)
    return (np.array([0, n], np.int32), np.arange(n, dtype=np.int32) % 1, np.asarray(values, np.float64),
            np.ones(1))


BIG = 2.0 ** 53  # next to it the spacing is 2, and BIG + 1 is a tie that rounds back to BIG


def _seq(vals):
    s = 0.0
    for v in vals:
        s = s + v
    return s


def test_short_rows_sum_in_storage_order():
    vals = [BIG] + [1.0] * 20 + [-BIG] + [1.0] * 5
    rp, col, val, x = _row(vals)
    assert len(val) <= 64
    assert O.csr_spmv(rp, col, val, x)[0] == 5.0 == _seq(vals)


def test_medium_rows_sum_lane_strided_then_pairwise():
    """100 entries: lane l sums entries l and l+64.  With BIG at entry 0 and
    -BIG at entry 64, lane 0 is exactly 0 and the other 98 ones survive: 98,
    the exact sum.  Storage order gives 35 (the 63 ones after BIG are lost)."""
    vals = [1.0] * 100
    vals[0], vals[64] = BIG, -BIG
    rp, col, val, x = _row(vals)
    assert _seq(vals) == 35.0
    assert O.csr_spmv(rp, col, val, x)[0] == 98.0


def test_long_rows_fold_chunk_values_in_order():
    """4096 + 64 entries, BIG at entry 0 (chunk 0, lane 0) and -BIG at entry
    4097 (chunk 1, lane 1), ones elsewhere.  Chunk 0 is BIG + 4032 (lane 0's
    63 ones are lost; the other lanes' 4032 ones fold in exactly), chunk 1 is
    -BIG + 63, exactly (magnitudes below 2^53), and the row is 4095, the exact
    sum.  Storage order gives 62."""
    vals = [1.0] * (4096 + 64)
    vals[0], vals[4097] = BIG, -BIG
    rp, col, val, x = _row(vals)
    assert _seq(vals) == 62.0
    assert O.csr_spmv(rp, col, val, x)[0] == 4095.0


def _row_order_numpy(v, short=64, chunk=4096):
    """The documented order restated with numpy arrays (lane sums as columns of
    a padded [k, 64] array, summed row by row)."""
    if len(v) <= short:
        s = v.dtype.type(0)
        for t in v:
            s = s + t
        return s
    out = None
    for c0 in range(0, len(v), chunk):
        part = v[c0:c0 + chunk]
        pad = np.zeros(-(-len(part) // 64) * 64, v.dtype)
        pad[:len(part)] = part
        m = pad.reshape(-1, 64)
        acc = np.zeros(64, v.dtype)
        for j in range(m.shape[0]):
            live = np.arange(64) + 64 * j < len(part)
            acc = np.where(live, acc + m[j], acc)
        s = 32
        while s >= 1:
            acc[:s] = acc[:s] + acc[s:2 * s]
            s //= 2
        out = acc[0] if out is None else out + acc[0]
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_row_order_against_a_numpy_restatement(dtype):
    rng = np.random.default_rng(11)
    lens = [0, 1, 31, 32, 33, 63, 64, 65, 100, 4095, 4096, 4097, 8192, 8193, 10000]
    rowptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    n = 12000
    col = rng.integers(0, n, size=rowptr[-1]).astype(np.int32)
    val = rng.uniform(-1, 1, size=rowptr[-1]).astype(dtype)
    x = rng.uniform(-1, 1, size=n).astype(dtype)
    y = O.csr_spmv(rowptr, col, val, x)
    for i in range(len(lens)):
        k0, k1 = rowptr[i], rowptr[i + 1]
        want = _row_order_numpy(val[k0:k1] * x[col[k0:k1]])
        assert y[i].tobytes() == np.asarray(want, dtype).tobytes(), (i, lens[i])


def test_skewed_matrix_is_spd_and_cg_solves_it():
    rowptr, col, val = O.skewed_spd_csr(3000, 5, hubs=(2500,))
    n = len(rowptr) - 1
    A = np.zeros((n, n))
    for i in range(n):
        A[i, col[rowptr[i]:rowptr[i + 1]]] = val[rowptr[i]:rowptr[i + 1]]
    assert np.array_equal(A, A.T)
    assert np.diff(rowptr).max() > 4096 // 2
    b = O.uniform(n, 3, -1.0, 1.0)
    x, iters, hist = O.cg(0, n, rowptr, col, val, 0, b, 1e-10, 1, 1000)
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-9
