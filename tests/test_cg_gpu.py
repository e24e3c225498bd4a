"""Persistent CG on the GPU through the C ABI, against the oracle.

The 7-point Poisson operator is generated on the device (CSR and DIA) and must
equal the oracle's arrays; the FP64 solve must reproduce the oracle's residual
history r.r bit for bit (the reduction order is fixed), and so must the
solution; the FP32-inner / FP64-refinement solve must reproduce the oracle's
outer residual history and inner iteration count.
"""
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def device_poisson(torch, fmt, nx, ny, nz):
    n = nx * ny * nz
    if fmt == "csr":
        nnz = P.poisson7_nnz(nx, ny, nz)
        rowptr = torch.empty(n + 1, dtype=torch.int32, device="cuda")
        col = torch.empty(nnz, dtype=torch.int32, device="cuda")
        val = torch.empty(nnz, dtype=torch.float64, device="cuda")
        P.poisson7_csr(nx, ny, nz, rowptr, col, val)
        return dict(rowptr=rowptr, col=col, val=val, nnz=nnz)
    val = torch.empty(7 * n, dtype=torch.float64, device="cuda")
    offs = P.poisson7_dia(nx, ny, nz, val)
    return dict(offsets=offs, val=val)


def oracle_args(fmt, nx, ny, nz):
    if fmt == "csr":
        rowptr, col, val = O.poisson7_csr(nx, ny, nz)
        return (0, rowptr, col, val, 0)
    offs, val = O.poisson7_dia(nx, ny, nz)
    return (1, None, offs, val, 7)


@pytest.mark.parametrize("dims", [(1, 1, 1), (5, 1, 3), (17, 9, 4), (64, 48, 40)])
def test_poisson_generators_match_oracle(gpu, dims):
    torch = gpu
    say(f"[cg] poisson generators {dims}")
    m = device_poisson(torch, "csr", *dims)
    rowptr, col, val = O.poisson7_csr(*dims)
    assert np.array_equal(m["rowptr"].cpu().numpy(), rowptr)
    assert np.array_equal(m["col"].cpu().numpy(), col)
    assert np.array_equal(m["val"].cpu().numpy(), val)
    d = device_poisson(torch, "dia", *dims)
    offs, dval = O.poisson7_dia(*dims)
    assert np.array_equal(d["offsets"], offs)
    assert np.array_equal(d["val"].cpu().numpy(), dval)


@pytest.mark.parametrize("fmt", ["csr", "dia"])
@pytest.mark.parametrize("dims,tol", [((3, 2, 2), 1e-12), ((24, 20, 18), 1e-10), ((40, 40, 40), 1e-10),
                                      ((70, 33, 17), 1e-8)])
def test_cg_f64_history_bitwise(gpu, fmt, dims, tol):
    torch = gpu
    nx, ny, nz = dims
    n = nx * ny * nz
    say(f"[cg] f64 {fmt} dims={dims} tol={tol}")
    m = device_poisson(torch, fmt, *dims)
    b_host = O.uniform(n, 7, -1.0, 1.0)
    b = torch.from_numpy(b_host).cuda()
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    res = P.cg_solve(fmt, n, b, x, tol=tol, max_iter=5000, **m)
    f, rowptr, col, val, ndiag = oracle_args(fmt, *dims)
    want_x, want_hist, want_k = O.cg(f, n, rowptr, col, val, ndiag, b_host, tol, True, 5000)
    assert res.code == 0
    assert res.iterations == want_k
    assert np.array_equal(res.history, want_hist), np.flatnonzero(res.history != want_hist)[:5]
    assert np.array_equal(x.cpu().numpy(), want_x)
    assert res.final_rr == want_hist[-1]
    assert res.bb == O.dot(b_host, b_host)


def test_cg_csr_and_dia_agree(gpu):
    torch = gpu
    dims = (33, 31, 29)
    n = dims[0] * dims[1] * dims[2]
    say(f"[cg] csr vs dia {dims}")
    b = torch.from_numpy(O.uniform(n, 2, -1.0, 1.0)).cuda()
    out = []
    for fmt in ("csr", "dia"):
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        res = P.cg_solve(fmt, n, b, x, tol=1e-10, **device_poisson(torch, fmt, *dims))
        out.append((res.history, x.cpu().numpy()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_cg_nonzero_start_and_absolute_tolerance(gpu):
    torch = gpu
    dims = (20, 16, 12)
    n = dims[0] * dims[1] * dims[2]
    say(f"[cg] x0 != 0, absolute tol {dims}")
    b_host = O.uniform(n, 4, -1.0, 1.0)
    x0 = O.uniform(n, 5, -0.5, 0.5)
    x = torch.from_numpy(x0.copy()).cuda()
    res = P.cg_solve("csr", n, torch.from_numpy(b_host).cuda(), x, tol=1e-9, relative=False,
                     **device_poisson(torch, "csr", *dims))
    rowptr, col, val = O.poisson7_csr(*dims)
    want_x, want_hist, want_k = O.cg(0, n, rowptr, col, val, 0, b_host, 1e-9, False, 10000, x0=x0)
    assert res.iterations == want_k
    assert np.array_equal(res.history, want_hist)
    assert np.array_equal(x.cpu().numpy(), want_x)


def test_cg_not_converged_code(gpu):
    torch = gpu
    dims = (30, 30, 30)
    n = dims[0] * dims[1] * dims[2]
    say(f"[cg] max_iter cut {dims}")
    b_host = O.uniform(n, 8, -1.0, 1.0)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    res = P.cg_solve("dia", n, torch.from_numpy(b_host).cuda(), x, tol=1e-12, max_iter=7,
                     **device_poisson(torch, "dia", *dims))
    assert res.code == -14 and res.iterations == 7
    offs, val = O.poisson7_dia(*dims)
    _, want_hist, _ = O.cg(1, n, None, offs, val, 7, b_host, 1e-12, True, 7)
    assert np.array_equal(res.history, want_hist)


@pytest.mark.parametrize("fmt", ["csr", "dia"])
@pytest.mark.parametrize("dims", [(12, 10, 9), (48, 40, 36)])
def test_cg_f32_refinement_matches_oracle(gpu, fmt, dims):
    torch = gpu
    nx, ny, nz = dims
    n = nx * ny * nz
    say(f"[cg] f32 refinement {fmt} dims={dims}")
    m = device_poisson(torch, fmt, *dims)
    b_host = O.uniform(n, 13, -1.0, 1.0)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    res = P.cg_solve(fmt, n, torch.from_numpy(b_host).cuda(), x, tol=1e-10, precision="f32_refine",
                     inner_tol=1e-4, max_inner=500, max_outer=20, **m)
    f, rowptr, col, val, ndiag = oracle_args(fmt, *dims)
    want_x, want_hist, want_total, want_outer = O.cg_f32_refine(f, n, rowptr, col, val, ndiag, b_host, 1e-10,
                                                                True, 1e-4, 500, 20)
    assert res.code == 0
    assert res.outer_iterations == want_outer
    assert res.iterations == want_total
    assert np.array_equal(res.history, want_hist)
    assert np.array_equal(x.cpu().numpy(), want_x)
    assert res.relative_residual < 1e-10


@pytest.mark.parametrize("fmt", ["csr", "dia"])
@pytest.mark.parametrize("relative", [True, False])
def test_cg_f32_matches_oracle(gpu, fmt, relative):
    """FP32 CG (the reference's myTest<float,...>): float data, double dots."""
    torch = gpu
    dims = (30, 26, 22)
    nx, ny, nz = dims
    n = nx * ny * nz
    say(f"[cg] f32 {fmt} relative={relative}")
    tol = 1e-5 if relative else 1e-3
    b_host = O.uniform(n, 17, -1.0, 1.0, np.float32)
    if fmt == "csr":
        rowptr, col, val = O.poisson7_csr(*dims)
        val32 = val.astype(np.float32)
        m = dict(rowptr=torch.from_numpy(rowptr).cuda(), col=torch.from_numpy(col).cuda(),
                 val=torch.from_numpy(val32).cuda(), nnz=len(col))
        oargs = (0, rowptr, col, val32, 0)
    else:
        offs, dval = O.poisson7_dia(*dims)
        val32 = dval.astype(np.float32)
        m = dict(offsets=offs, val=torch.from_numpy(val32).cuda())
        oargs = (1, None, offs, val32, 7)
    x = torch.zeros(n, dtype=torch.float32, device="cuda")
    res = P.cg_solve(fmt, n, torch.from_numpy(b_host).cuda(), x, tol=tol, relative=relative, max_iter=3000,
                     precision="f32", **m)
    f, rp, cl, vv, nd = oargs
    thr2 = float(np.float32(tol) * np.float32(tol))
    want_x, want_hist, want_k = O.cg_f32(f, n, rp, cl, vv, nd, b_host, thr2, relative, 3000)
    assert res.code == 0 and res.iterations == want_k
    assert np.array_equal(res.history, want_hist)
    assert np.array_equal(x.cpu().numpy(), want_x)


def test_cg_static_iterations(gpu):
    torch = gpu
    dims = (16, 16, 16)
    n = 16 ** 3
    say("[cg] static iterations")
    b = torch.from_numpy(O.uniform(n, 3, -1.0, 1.0)).cuda()
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    res = P.cg_solve("dia", n, b, x, tol=1e-3, max_iter=12, static_iter=True, **device_poisson(torch, "dia", *dims))
    assert res.code == 0 and res.iterations == 12 and len(res.history) == 13


@pytest.mark.parametrize("value_bytes", [8, 4])
def test_cg_driver_virtual_tridiagonal(gpu, value_bytes):
    """The reference's driver on its virtual matrix: rhs = 1, x = 0, absolute tol 1e-10."""
    say(f"[cg] driver vdata value_bytes={value_bytes}")
    r = P.cg_driver(value_bytes=value_bytes, vdata_n=4096, check=True, max_iter=200)
    assert r.n == 4096 and r.nnz == 3 * 4096 - 2
    assert r.iterations > 0 and r.elapsed_ms > 0
    if value_bytes == 8:
        assert r.code == 0 and r.residual < 1e-10 and r.check_error < 1e-9
    else:
        assert r.check_error < 1e-4


def test_cg_driver_matrix_market_file(gpu, tmp_path):
    """A Matrix Market file of the Poisson operator through the driver: the same
    iterations and residual as the CSR solve on the device-generated matrix."""
    torch = gpu
    dims = (12, 10, 9)
    n = 12 * 10 * 9
    rowptr, col, val = O.poisson7_csr(*dims)
    f = tmp_path / "poisson.mtx"
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    lower = col <= rows
    lines = [f"{c + 1} {r + 1} {float(v)!r}" for r, c, v in zip(rows[lower], col[lower], val[lower])]
    f.write_text("%%MatrixMarket matrix coordinate real symmetric\n" + f"{n} {n} {len(lines)}\n" + "\n".join(lines)
                 + "\n")
    say("[cg] driver mtx")
    r = P.cg_driver(mtx=str(f), value_bytes=8, check=True)
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    res = P.cg_solve("csr", n, b, x, tol=1e-10, relative=False, max_iter=10 * n, x0_is_zero=True,
                     **device_poisson(torch, "csr", *dims))
    assert r.code == 0 and r.iterations == res.iterations
    assert abs(r.residual - np.sqrt(res.final_rr)) == 0.0
    assert r.check_error < 1e-9
