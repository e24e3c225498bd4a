"""Persistent CG's work split and on-chip caches on the GPU, against the oracle.

* Skewed CSR matrices (power-law row lengths, hub rows of thousands of
  entries): rows of up to 64 entries are summed by their lane, longer ones
  in chunks of up to 4096 entries by any wave of the grid with a fix-up, and
  the residual history, iteration count and solution still equal the
  oracle's bit for bit, in FP64 and FP32, from x = 0 and from a nonzero
  start.
* The LDS caches (Ap, x, p, r and the matrix) change no arithmetic: every
  solve gives the same bits with them on and off, and the report shows what
  was cached.  A domain of one segment per workgroup keeps all four vectors
  and part of the matrix on chip.
* The 256^3 Poisson configuration in full: CSR and DIA reach
  ||r|| / ||b|| < 1e-10 with bitwise-equal histories, checked by an
  independent residual computed with torch on the device; so does FP32-inner /
  FP64-refinement.
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


def device_csr(torch, rowptr, col, val):
    return dict(rowptr=torch.from_numpy(rowptr).cuda(), col=torch.from_numpy(col).cuda(),
                val=torch.from_numpy(val).cuda(), nnz=len(col))


@pytest.mark.parametrize("x0", ["zero", "nonzero"])
def test_skewed_csr_f64_matches_oracle(gpu, x0):
    torch = gpu
    rowptr, col, val = O.skewed_spd_csr(20000, 3, hubs=(9000, 5000, 4100))
    n = len(rowptr) - 1
    lens = np.diff(rowptr)
    say(f"[cgc] skewed f64 x0={x0} n={n} nnz={len(col)} max row {lens.max()} "
        f"medium {(lens > 32).sum()} long {(lens > 4096).sum()}")
    assert (lens > 4096).sum() >= 2 and ((lens > 64) & (lens <= 4096)).sum() > 100
    b_host = O.uniform(n, 21, -1.0, 1.0)
    x_init = O.uniform(n, 22, -0.5, 0.5) if x0 == "nonzero" else np.zeros(n)
    x = torch.from_numpy(x_init.copy()).cuda()
    res = P.cg_solve("csr", n, torch.from_numpy(b_host).cuda(), x, tol=1e-10, max_iter=2000,
                     x0_is_zero=(x0 == "zero"), **device_csr(torch, rowptr, col, val))
    want_x, want_hist, want_k = O.cg(0, n, rowptr, col, val, 0, b_host, 1e-10, True, 2000, x0=x_init)
    assert res.code == 0 and res.long_rows == (lens > 64).sum()
    assert res.iterations == want_k
    assert np.array_equal(res.history, want_hist), np.flatnonzero(res.history != want_hist)[:5]
    assert np.array_equal(x.cpu().numpy(), want_x)


def test_skewed_csr_f32_matches_oracle(gpu):
    torch = gpu
    rowptr, col, val = O.skewed_spd_csr(12000, 9, hubs=(8500,), dtype=np.float32)
    n = len(rowptr) - 1
    say(f"[cgc] skewed f32 n={n} nnz={len(col)}")
    b_host = O.uniform(n, 23, -1.0, 1.0, np.float32)
    x = torch.zeros(n, dtype=torch.float32, device="cuda")
    res = P.cg_solve("csr", n, torch.from_numpy(b_host).cuda(), x, tol=1e-5, max_iter=1000, precision="f32",
                     **device_csr(torch, rowptr, col, val))
    thr2 = float(np.float32(1e-5) * np.float32(1e-5))
    want_x, want_hist, want_k = O.cg_f32(0, n, rowptr, col, val, 0, b_host, thr2, True, 1000)
    assert res.code == 0 and res.long_rows == (np.diff(rowptr) > 64).sum() and res.iterations == want_k
    assert np.array_equal(res.history, want_hist)
    assert np.array_equal(x.cpu().numpy(), want_x)


def test_skewed_csr_refinement_matches_oracle(gpu):
    torch = gpu
    rowptr, col, val = O.skewed_spd_csr(15000, 4, hubs=(6000,))
    n = len(rowptr) - 1
    say(f"[cgc] skewed refinement n={n}")
    b_host = O.uniform(n, 24, -1.0, 1.0)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    res = P.cg_solve("csr", n, torch.from_numpy(b_host).cuda(), x, tol=1e-10, precision="f32_refine",
                     inner_tol=1e-4, max_inner=500, max_outer=20, **device_csr(torch, rowptr, col, val))
    want_x, want_hist, want_total, want_outer = O.cg_f32_refine(0, n, rowptr, col, val, 0, b_host, 1e-10, True,
                                                                1e-4, 500, 20)
    assert res.code == 0 and res.outer_iterations == want_outer and res.iterations == want_total
    assert np.array_equal(res.history, want_hist)
    assert np.array_equal(x.cpu().numpy(), want_x)


def poisson(torch, fmt, dims, dtype=np.float64):
    n = int(np.prod(dims))
    if fmt == "csr":
        rowptr, col, val = O.poisson7_csr(*dims)
        return device_csr(torch, rowptr, col, val.astype(dtype)), (0, rowptr, col, val.astype(dtype), 0), n
    offs, val = O.poisson7_dia(*dims)
    return dict(offsets=offs, val=torch.from_numpy(val.astype(dtype)).cuda()), (1, None, offs, val.astype(dtype), 7), n


@pytest.mark.parametrize("fmt", ["csr", "dia"])
@pytest.mark.parametrize("precision", ["f64", "f32_refine", "f32"])
def test_caches_change_no_bits(gpu, fmt, precision):
    torch = gpu
    dims = (40, 36, 30)
    dt = np.float32 if precision == "f32" else np.float64
    m, _, n = poisson(torch, fmt, dims, dt)
    b_host = O.uniform(n, 31, -1.0, 1.0, dt)
    say(f"[cgc] caches on/off {fmt} {precision}")
    out = []
    for cache in (True, False):
        x = torch.zeros(n, dtype=torch.float32 if dt == np.float32 else torch.float64, device="cuda")
        res = P.cg_solve(fmt, n, torch.from_numpy(b_host).cuda(), x, tol=1e-5 if dt == np.float32 else 1e-10,
                         max_iter=3000, precision=precision, inner_tol=1e-4, max_inner=500, cache_vector=cache,
                         cache_matrix=cache, **m)
        assert res.code == 0
        out.append((res, x.cpu().numpy()))
    (on, x_on), (off, x_off) = out
    say(f"[cgc]   cached: {on.lds_bytes} B LDS per workgroup, {on.cache_units} vector units, "
        f"{on.cache_matrix} matrix entries/rows; grid {on.grid}")
    assert on.cache_units > 0 and on.lds_bytes > 0
    assert off.cache_units == 0 and off.cache_matrix == 0 and off.lds_bytes == 0
    assert on.iterations == off.iterations
    assert np.array_equal(on.history, off.history)
    assert np.array_equal(x_on, x_off)


@pytest.mark.parametrize("fmt", ["csr", "dia"])
def test_one_segment_per_workgroup_caches_everything(gpu, fmt):
    """16^3 rows = one segment: one workgroup holding the vectors and, where
    its LDS has room, the first part of the matrix; still the oracle's bits."""
    torch = gpu
    dims = (16, 16, 16)
    m, oargs, n = poisson(torch, fmt, dims)
    b_host = O.uniform(n, 41, -1.0, 1.0)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    res = P.cg_solve(fmt, n, torch.from_numpy(b_host).cuda(), x, tol=1e-10, max_iter=500, **m)
    say(f"[cgc] one segment {fmt}: grid {res.grid}, {res.lds_bytes} B, units {res.cache_units}, "
        f"matrix {res.cache_matrix}")
    lds = P.query_device().lds_per_block
    assert res.grid == 1 and res.cache_units >= min(4 * 16, (lds - 4096) // 2048)
    if lds >= 4 * 16 * 2048 + 4096:  # room for all four vectors and some of the matrix
        assert res.cache_units == 4 * 16 and res.cache_matrix > 0
    f, rowptr, col, val, ndiag = oargs
    want_x, want_hist, want_k = O.cg(f, n, rowptr, col, val, ndiag, b_host, 1e-10, True, 500)
    assert res.iterations == want_k
    assert np.array_equal(res.history, want_hist)
    assert np.array_equal(x.cpu().numpy(), want_x)


def _poisson_residual(torch, x, b, m):
    """||b - A x|| / ||b|| with A the 7-point Poisson operator applied by torch
    (zero-padded shifts in FP64): independent of the library's SpMV."""
    u = torch.nn.functional.pad(x.view(m, m, m), (1, 1, 1, 1, 1, 1))
    ax = 6.0 * u[1:-1, 1:-1, 1:-1] - u[:-2, 1:-1, 1:-1] - u[2:, 1:-1, 1:-1] - u[1:-1, :-2, 1:-1] \
        - u[1:-1, 2:, 1:-1] - u[1:-1, 1:-1, :-2] - u[1:-1, 1:-1, 2:]
    r = b.view(m, m, m) - ax
    return float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b))


def test_poisson_256_full_configuration(gpu):
    """The 256^3 CG configuration: rhs uniform[-1,1), x0 = 0, ||r||/||b|| < 1e-10."""
    torch = gpu
    m = 256
    n = m ** 3
    b = torch.empty(n, dtype=torch.float64, device="cuda")
    P.fill_uniform(b, n, 1, -1.0, 1.0)
    hist = {}
    for fmt in ("csr", "dia"):
        if fmt == "csr":
            nnz = P.poisson7_nnz(m, m, m)
            rowptr = torch.empty(n + 1, dtype=torch.int32, device="cuda")
            col = torch.empty(nnz, dtype=torch.int32, device="cuda")
            val = torch.empty(nnz, dtype=torch.float64, device="cuda")
            P.poisson7_csr(m, m, m, rowptr, col, val)
            mat = dict(rowptr=rowptr, col=col, val=val, nnz=nnz)
        else:
            val = torch.empty(7 * n, dtype=torch.float64, device="cuda")
            mat = dict(offsets=P.poisson7_dia(m, m, m, val), val=val)
        for precision in ("f64", "f32_refine"):
            x = torch.zeros(n, dtype=torch.float64, device="cuda")
            res = P.cg_solve(fmt, n, b, x, tol=1e-10, max_iter=5000, precision=precision, x0_is_zero=True,
                             max_outer=20, inner_tol=1e-5, **mat)
            rel = _poisson_residual(torch, x, b, m)
            say(f"[cgc] 256^3 {fmt} {precision}: {res.iterations} iterations ({res.outer_iterations} outer), "
                f"{res.elapsed_ms:.1f} ms, library ||r||/||b|| {res.relative_residual:.3e}, torch {rel:.3e}")
            assert res.code == 0 and res.relative_residual < 1e-10 and rel < 1e-10
            if precision == "f64":
                hist[fmt] = res.history
        del mat, val
        torch.cuda.empty_cache()
    assert len(hist["csr"]) > 500
    assert np.array_equal(hist["csr"], hist["dia"])


@pytest.mark.parametrize("fmt", ["csr", "dia"])
def test_every_workgroups_per_cu_fits_at_256(gpu, fmt):
    """256^3 with 1 to 4 workgroups per CU (opts.blkpsm): the planner sizes each
    workgroup's LDS caches to the CU's LDS divided by that count, rounded to the
    allocation granule, so the whole cooperative grid is resident at every
    count (three per CU once took 54784 B each in 512-byte granules, more than
    a CU holds, and the grid barrier timed out).  Every count gives the same
    residual history."""
    torch = gpu
    m = 256
    n = m ** 3
    b = torch.empty(n, dtype=torch.float64, device="cuda")
    P.fill_uniform(b, n, 1, -1.0, 1.0)
    if fmt == "csr":
        nnz = P.poisson7_nnz(m, m, m)
        rowptr = torch.empty(n + 1, dtype=torch.int32, device="cuda")
        col = torch.empty(nnz, dtype=torch.int32, device="cuda")
        val = torch.empty(nnz, dtype=torch.float64, device="cuda")
        P.poisson7_csr(m, m, m, rowptr, col, val)
        mat = dict(rowptr=rowptr, col=col, val=val, nnz=nnz)
    else:
        val = torch.empty(7 * n, dtype=torch.float64, device="cuda")
        mat = dict(offsets=P.poisson7_dia(m, m, m, val), val=val)
    ref = None
    for blk in (4, 3, 2, 1):
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        say(f"[cgc] 256^3 {fmt} blkpsm={blk}")
        res = P.cg_solve(fmt, n, b, x, tol=1e-10, max_iter=5000, x0_is_zero=True, blkpsm=blk, **mat)
        say(f"[cgc]   grid {res.grid}, {res.iterations} iterations, {res.elapsed_ms:.1f} ms, LDS {res.lds_bytes} B")
        assert res.code == 0 and res.grid == 256 * blk and res.relative_residual < 1e-10
        if ref is None:
            ref = res.history
        assert np.array_equal(res.history, ref), blk
