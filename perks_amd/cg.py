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


def _opts(tol, relative, max_iter, precision, inner_tol, max_inner, max_outer, x0_is_zero, blkpsm):
    o = L.CgOpts()
    L.load().perks_cg_opts_init(C.byref(o))
    o.tol, o.relative, o.max_iter = float(tol), int(relative), int(max_iter)
    o.precision = L.CG_F32_REFINE if precision in ("f32_refine", L.CG_F32_REFINE) else L.CG_F64
    o.inner_tol, o.max_inner, o.max_outer = float(inner_tol), int(max_inner), int(max_outer)
    o.x0_is_zero, o.blkpsm = int(x0_is_zero), int(blkpsm)
    return o


def cg_solve(fmt, n, b, x, *, rowptr=None, col=None, val=None, nnz=None, offsets=None, tol=1e-10, relative=True,
             max_iter=10000, precision="f64", inner_tol=1e-5, max_inner=0, max_outer=0, x0_is_zero=False,
             blkpsm=0, stream=None, hist_cap=None) -> CgResult:
    """fmt 'csr' (rowptr, col, val, nnz) or 'dia' (offsets: host int32 array,
    val: device ndiag*n).  Returns a CgResult; raises on argument errors,
    returns code PERKS_ERR_NOT_CONVERGED (-14) without raising."""
    lib = L.load()
    o = _opts(tol, relative, max_iter, precision, inner_tol, max_inner, max_outer, x0_is_zero, blkpsm)
    cap = hist_cap if hist_cap is not None else max_iter + 2
    hist = np.zeros(max(cap, 1))
    rep = L.CgReport()
    rep.hist_cap = hist.size
    rep.history = hist.ctypes.data_as(C.POINTER(C.c_double))
    s = None if stream is None else _ptr_stream(stream)
    if fmt == "csr":
        A = L.Csr()
        A.n, A.nnz, A.rowptr, A.col, A.val = int(n), int(nnz), _ptr(rowptr), _ptr(col), _ptr(val)
        rc = lib.perks_cg_solve_csr(C.byref(A), _ptr(b), _ptr(x), C.byref(o), s, C.byref(rep))
    else:
        offs = np.ascontiguousarray(offsets, dtype=np.int32)
        A = L.Dia()
        A.n, A.ndiag, A.offsets, A.val = int(n), int(offs.size), offs.ctypes.data, _ptr(val)
        rc = lib.perks_cg_solve_dia(C.byref(A), _ptr(b), _ptr(x), C.byref(o), s, C.byref(rep))
    if rc < 0 and rc != L.ERR_NOT_CONVERGED:
        raise L.PerksError(rc, "cg_solve")
    return CgResult(rc, rep.iterations, rep.outer_iterations, rep.final_rr, rep.bb, rep.elapsed_ms, rep.grid,
                    hist[: min(rep.hist_len, hist.size)].copy())
