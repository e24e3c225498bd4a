"""Host-side mirror of the reference's conjugate-gradient entry over the C ABI.

The reference drives its persistent CG from myTest<ValueT,...>
(conjugateGradient/cg_driver.cu:53-428) through DispatchCG::ProcessDispatchCG
(cg_dispatcher.cuh:559-916) on a CSR matrix in device memory.  Here the
matrix (CSR or 7-diagonal DIA), b and x live in device memory (torch tensors
or raw pointers) and one call runs the whole solve.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .stencil import _ptr, _ptr_stream


@dataclass
class CgResult:
    code: int
    iterations: int
    outer_iterations: int
    final_rr: float
    bb: float
    elapsed_ms: float
    grid: int
    history: np.ndarray = field(repr=False)
    lds_bytes: int = 0     # LDS cache per workgroup
    cache_units: int = 0   # 256-row vector units (Ap, x, p, r) cached per workgroup
    cache_matrix: int = 0  # CSR entries / DIA rows cached per workgroup
    long_rows: int = 0     # CSR rows of more than 64 entries, split across the grid

    @property
    def relative_residual(self) -> float:
        return float(np.sqrt(self.final_rr / self.bb)) if self.bb > 0 else float(np.sqrt(self.final_rr))


def poisson7_nnz(nx, ny, nz) -> int:
    return int(L.load().perks_poisson7_nnz(nx, ny, nz))


def poisson7_csr(nx, ny, nz, rowptr, col, val, stream=None) -> None:
    s = None if stream is None else _ptr_stream(stream)
    L.check(L.load().perks_poisson7_csr(nx, ny, nz, _ptr(rowptr), _ptr(col), _ptr(val), s), "poisson7_csr")


def poisson7_dia(nx, ny, nz, val, stream=None) -> np.ndarray:
    offs = np.zeros(7, np.int32)
    s = None if stream is None else _ptr_stream(stream)
    L.check(L.load().perks_poisson7_dia(nx, ny, nz, offs.ctypes.data, _ptr(val), s), "poisson7_dia")
    return offs


def _opts(tol, relative, max_iter, precision, inner_tol, max_inner, max_outer, x0_is_zero, blkpsm, static_iter=0,
          cache_vector=True, cache_matrix=True):
    o = L.CgOpts()
    L.load().perks_cg_opts_init(C.byref(o))
    o.tol, o.relative, o.max_iter = float(tol), int(relative), int(max_iter)
    o.precision = L.CG_F32_REFINE if precision in ("f32_refine", L.CG_F32_REFINE) else L.CG_F64
    o.inner_tol, o.max_inner, o.max_outer = float(inner_tol), int(max_inner), int(max_outer)
    o.x0_is_zero, o.blkpsm, o.static_iter = int(x0_is_zero), int(blkpsm), int(static_iter)
    o.cache_vector, o.cache_matrix = (0 if cache_vector else -1), (0 if cache_matrix else -1)
    return o


def cg_solve(fmt, n, b, x, *, rowptr=None, col=None, val=None, nnz=None, offsets=None, tol=1e-10, relative=True,
             max_iter=10000, precision="f64", inner_tol=1e-5, max_inner=0, max_outer=0, x0_is_zero=False,
             blkpsm=0, stream=None, hist_cap=None, static_iter=False, cache_vector=True,
             cache_matrix=True) -> CgResult:
    """fmt 'csr' (rowptr, col, val, nnz) or 'dia' (offsets: host int32 array,
    val: device ndiag*n).  precision 'f64', 'f32_refine', or 'f32' (FP32
    matrix and vectors: val, b, x float32).  Returns a CgResult; raises on
    argument errors, returns code PERKS_ERR_NOT_CONVERGED (-14) without raising.
    cache_vector / cache_matrix: the LDS caches across iterations (the
    reference's cacheVector / cacheMatrix); they change no arithmetic."""
    lib = L.load()
    o = _opts(tol, relative, max_iter, "f64" if precision == "f32" else precision, inner_tol, max_inner, max_outer,
              x0_is_zero, blkpsm, static_iter, cache_vector, cache_matrix)
    cap = hist_cap if hist_cap is not None else max_iter + 2
    hist = np.zeros(max(cap, 1))
    rep = L.CgReport()
    rep.hist_cap = hist.size
    rep.history = hist.ctypes.data_as(C.POINTER(C.c_double))
    s = None if stream is None else _ptr_stream(stream)
    if fmt == "csr":
        A = L.Csr()
        A.n, A.nnz, A.rowptr, A.col, A.val = int(n), int(nnz), _ptr(rowptr), _ptr(col), _ptr(val)
        fn = lib.perks_cg_solve_csr_f32 if precision == "f32" else lib.perks_cg_solve_csr
        rc = fn(C.byref(A), _ptr(b), _ptr(x), C.byref(o), s, C.byref(rep))
    else:
        offs = np.ascontiguousarray(offsets, dtype=np.int32)
        A = L.Dia()
        A.n, A.ndiag, A.offsets, A.val = int(n), int(offs.size), offs.ctypes.data, _ptr(val)
        fn = lib.perks_cg_solve_dia_f32 if precision == "f32" else lib.perks_cg_solve_dia
        rc = fn(C.byref(A), _ptr(b), _ptr(x), C.byref(o), s, C.byref(rep))
    if rc < 0 and rc != L.ERR_NOT_CONVERGED:
        raise L.PerksError(rc, "cg_solve")
    return CgResult(rc, rep.iterations, rep.outer_iterations, rep.final_rr, rep.bb, rep.elapsed_ms, rep.grid,
                    hist[: min(rep.hist_len, hist.size)].copy(), rep.lds_bytes, rep.cache_units, rep.cache_matrix,
                    rep.long_rows)


# ---- the reference's driver and sparse inputs ----------------------------------
def skewed_spd_csr(n, seed, hubs=(), max_deg=400, dtype=np.float64, degree=None):
    """A symmetric, diagonally dominant (so SPD) matrix with power-law row
    lengths: every row links to a Zipf-distributed number of random columns,
    each hub row to `hubs[i]` columns; the pattern is symmetrised, off-diagonal
    values are uniform in [-1, 0) and the diagonal is 1 + the row's absolute sum.
    Returns (rowptr, col, val) with ascending columns.  degree: every row
    links to that many random columns instead (the same structure without
    the skew)."""
    rng = np.random.default_rng(seed)
    deg = np.full(n, degree) if degree else np.minimum(rng.zipf(1.6, size=n), max_deg)
    rows = [np.repeat(np.arange(n), deg)]
    cols = [rng.integers(0, n, size=int(deg.sum()))]
    for h, d in enumerate(hubs):
        hub = (h * 7919 + 13) % n
        rows.append(np.full(d, hub))
        cols.append(rng.choice(n, size=d, replace=False))
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    keep = r != c
    r, c = r[keep], c[keep]
    r, c = np.concatenate([r, c]), np.concatenate([c, r])
    key = np.unique(r.astype(np.int64) * n + c)
    r, c = (key // n).astype(np.int64), (key % n).astype(np.int64)
    # symmetric values: (i, j) and (j, i) take the value drawn for the unordered pair
    lo, hi = np.minimum(r, c), np.maximum(r, c)
    pair = lo * n + hi
    uniq, inv = np.unique(pair, return_inverse=True)
    pv = -rng.uniform(1e-3, 1.0, size=len(uniq))
    v = pv[inv]
    diag = np.zeros(n)
    np.add.at(diag, r, np.abs(v))
    r = np.concatenate([r, np.arange(n)])
    c = np.concatenate([c, np.arange(n)])
    v = np.concatenate([v, diag + 1.0])
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    rowptr = np.zeros(n + 1, dtype=np.int32)
    np.add.at(rowptr, r + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int32)
    return rowptr, c.astype(np.int32), v.astype(dtype)



def cg_driver(mtx=None, value_bytes=8, vdata_n=0, max_iter=0, static_iter=False, warmup=False, check=False,
              verbose=False, tol=0.0) -> L.CgDriverResult:
    """myTest<ValueT,int,...> (conjugateGradient/cg_driver.cu:53-428) on this
    library: a Matrix Market file (mtx) or the virtual tridiagonal matrix."""
    o = L.CgDriverOpts()
    o.mtx_path = (mtx or "").encode()
    o.value_bytes, o.vdata_n, o.max_iter = int(value_bytes), int(vdata_n), int(max_iter)
    o.static_iter, o.warmup, o.check, o.verbose = int(static_iter), int(warmup), int(check), int(verbose)
    o.tol = float(tol)
    r = L.CgDriverResult()
    L.check(L.load().perks_cg_driver(C.byref(o), C.byref(r)), "cg_driver")
    return r


def mm_read_csr(path):
    """Matrix Market file -> (rowptr, col, val) numpy arrays (the reference's
    reader and COO -> CSR order)."""
    lib = L.load()
    n, nnz = C.c_int32(), C.c_int64()
    rp, cl, vl = C.c_void_p(), C.c_void_p(), C.c_void_p()
    L.check(lib.perks_mm_read_csr(str(path).encode(), C.byref(n), C.byref(nnz), C.byref(rp), C.byref(cl),
                                  C.byref(vl)), "mm_read_csr")
    try:
        rowptr = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int32)), (n.value + 1,)).copy()
        col = np.ctypeslib.as_array(C.cast(cl, C.POINTER(C.c_int32)), (max(nnz.value, 1),))[: nnz.value].copy()
        val = np.ctypeslib.as_array(C.cast(vl, C.POINTER(C.c_double)), (max(nnz.value, 1),))[: nnz.value].copy()
    finally:
        for q in (rp, cl, vl):
            lib.perks_free_host(q)
    return rowptr, col, val


def gen_tridiag(n, dtype=np.float64):
    """The reference's random symmetric tridiagonal matrix (tridiag.cpp:11-47)."""
    nnz = 3 * n - 2
    rowptr = np.empty(n + 1, np.int32)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, dtype)
    fn = L.load().perks_gen_tridiag_f64 if np.dtype(dtype) == np.float64 else L.load().perks_gen_tridiag_f32
    L.check(fn(n, rowptr.ctypes.data, col.ctypes.data, val.ctypes.data), "gen_tridiag")
    return rowptr, col, val


def csr_to_dia(rowptr, col, val, max_diag=16):
    """Host CSR -> DIA: (offsets, values[ndiag*n])."""
    n = len(rowptr) - 1
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    offs = np.empty(max_diag, np.int32)
    nd = L.load().perks_csr_to_dia(n, rowptr.ctypes.data, col.ctypes.data, val.ctypes.data, max_diag,
                                   offs.ctypes.data, None)
    if nd < 0:
        raise L.PerksError(nd, "csr_to_dia")
    dv = np.empty(nd * n, np.float64)
    L.load().perks_csr_to_dia(n, rowptr.ctypes.data, col.ctypes.data, val.ctypes.data, max_diag, offs.ctypes.data,
                              dv.ctypes.data)
    return offs[:nd].copy(), dv
