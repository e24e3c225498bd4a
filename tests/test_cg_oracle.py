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
