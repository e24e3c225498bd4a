"""ctypes access to the CPU oracle (oracle/libperks_oracle.so) and to the
reference's own CPU gold routines built from their sources (oracle/_ref/*.so).

Test infrastructure: used as the checker only.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "libperks_oracle.so"
REF_DIR = ROOT / "oracle" / "_ref"

PO_MAX_GROUPS = 64
PO_MAX_POINTS = 12
BC = {"clamp": 0, "zero": 1}


class PoGroup(C.Structure):
    _fields_ = [("coef", C.c_double), ("npts", C.c_int32), ("prod_f64", C.c_int32),
                ("off", (C.c_int32 * 3) * PO_MAX_POINTS)]


class PoStencil(C.Structure):
    _fields_ = [("ngroups", C.c_int32), ("boundary", C.c_int32), ("divisor", C.c_double),
                ("groups", PoGroup * PO_MAX_GROUPS)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            subprocess.run(["make", "-C", str(ROOT / "oracle"), "oracle"], check=True, capture_output=True)
        l = C.CDLL(str(ORACLE_SO))
        P, I, D, S = C.c_void_p, C.c_int, C.c_double, C.c_size_t
        sig = {
            "po_splitmix64": (C.c_uint64, [C.c_uint64, C.c_uint64]),
            "po_fill_uniform_f64": (None, [P, S, C.c_uint64, D, D]),
            "po_fill_uniform_f32": (None, [P, S, C.c_uint64, D, D]),
            "po_fill_libc_rand_f64": (None, [P, S, C.c_uint]),
            "po_fill_libc_rand_f32": (None, [P, S, C.c_uint]),
            "po_stencil_step_f64": (None, [P, P, I, I, I, P]),
            "po_stencil_step_f32": (None, [P, P, I, I, I, P]),
            "po_stencil_iterate_f64": (None, [P, P, I, I, I, P, I]),
            "po_stencil_iterate_f32": (None, [P, P, I, I, I, P, I]),
            "po_stencil_iterate_omp_f64": (None, [P, P, I, I, I, P, I, I]),
            "po_stencil_iterate_omp_f32": (None, [P, P, I, I, I, P, I, I]),
            "po_star7_iterate_omp_f64": (None, [P, P, I, I, I, P, I, I, I]),
            "po_stencil_iterate_fused_f64": (I, [P, P, I, I, I, P, I]),
            "po_stencil_iterate_fused_f32": (I, [P, P, I, I, I, P, I]),
            "po_f32_to_f16": (C.c_uint16, [C.c_float]),
            "po_f16_to_f32": (C.c_float, [C.c_uint16]),
            "po_f32_to_bf16": (C.c_uint16, [C.c_float]),
            "po_bf16_to_f32": (C.c_float, [C.c_uint16]),
            "po_stencil_iterate_f16s": (None, [P, P, I, I, I, P, I]),
            "po_stencil_iterate_bf16s": (None, [P, P, I, I, I, P, I]),
            "po_lowbias32": (C.c_uint32, [C.c_uint32]),
            "po_cg_f32": (I, [I, I, P, P, P, I, P, P, C.c_double, I, I, I, P, I]),
            "po_sr_bits": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
            "po_f32_to_bf16_sr": (C.c_uint16, [C.c_float, C.c_uint32]),
            "po_f32_to_f16_sr": (C.c_uint16, [C.c_float, C.c_uint32]),
            "po_stencil_iterate_sr16": (None, [P, P, I, I, I, P, I, I, C.c_uint32]),
            "po_rms_error_f64": (D, [P, P, I, I, I, I, I, I, I, I, P]),
            "po_rms_error_f32": (D, [P, P, I, I, I, I, I, I, I, I, P]),
            "po_poisson7_nnz": (C.c_int64, [I, I, I]),
            "po_poisson7_csr": (None, [I, I, I, P, P, P]),
            "po_poisson7_dia": (None, [I, I, I, P, P]),
            "po_csr_spmv_f64": (None, [I, P, P, P, P, P]),
            "po_csr_spmv_f32": (None, [I, P, P, P, P, P]),
            "po_dia_spmv_f64": (None, [I, I, P, P, P, P]),
            "po_dot_f64": (D, [P, P, S]),
            "po_dot_f32": (D, [P, P, S]),
            "po_cg_f64": (I, [I, I, P, P, P, I, P, P, D, I, I, P, I]),
            "po_cg_f32_refine": (I, [I, I, P, P, P, I, P, P, D, I, D, I, I, P, I, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(l, name)
            f.restype, f.argtypes = res, args
        _lib = l
    return _lib


def p(a: np.ndarray) -> int:
    return a.ctypes.data


# ---- inputs -----------------------------------------------------------------
def uniform(n: int, seed: int, lo: float, hi: float, dtype=np.float64) -> np.ndarray:
    out = np.empty(n, dtype=dtype)
    fn = lib().po_fill_uniform_f64 if dtype == np.float64 else lib().po_fill_uniform_f32
    fn(p(out), n, seed, lo, hi)
    return out


def libc_rand(n: int, seed: int, dtype=np.float64) -> np.ndarray:
    out = np.empty(n, dtype=dtype)
    fn = lib().po_fill_libc_rand_f64 if dtype == np.float64 else lib().po_fill_libc_rand_f32
    fn(p(out), n, seed)
    return out


# ---- stencil ----------------------------------------------------------------
def stencil(groups, boundary="clamp", divisor=0.0) -> PoStencil:
    """groups: list of (coef, [(dz,dy,dx), ...], prod_f64)."""
    st = PoStencil()
    st.ngroups = len(groups)
    st.boundary = BC[boundary]
    st.divisor = divisor
    for g, (coef, pts, pf) in enumerate(groups):
        st.groups[g].coef = coef
        st.groups[g].npts = len(pts)
        st.groups[g].prod_f64 = pf
        for k, (dz, dy, dx) in enumerate(pts):
            st.groups[g].off[k][0], st.groups[g].off[k][1], st.groups[g].off[k][2] = dz, dy, dx
    return st


def stencil_from_shape(si, boundary="clamp", coefs=None, divisor=None) -> PoStencil:
    """Builds the oracle's stencil from a perks_amd ShapeInfo (reference weights
    unless coefs is given)."""
    c = si.coef if coefs is None else list(coefs)
    d = si.divisor if coefs is None else (0.0 if divisor is None else divisor)
    return stencil([(c[g], si.offsets[g], si.prod_f64[g]) for g in range(len(si.offsets))], boundary, d)


def iterate(field: np.ndarray, st: PoStencil, steps: int, threads: int = 1) -> np.ndarray:
    field = np.ascontiguousarray(field)
    if field.ndim == 2:
        nz, (ny, nx) = 1, field.shape
    else:
        nz, ny, nx = field.shape
    out = np.empty_like(field)
    f64 = field.dtype == np.float64
    if threads > 1:
        fn = lib().po_stencil_iterate_omp_f64 if f64 else lib().po_stencil_iterate_omp_f32
        fn(p(field), p(out), nz, ny, nx, C.byref(st), steps, threads)
    else:
        fn = lib().po_stencil_iterate_f64 if f64 else lib().po_stencil_iterate_f32
        fn(p(field), p(out), nz, ny, nx, C.byref(st), steps)
    return out


def star7_iterate(field: np.ndarray, k, boundary: str, steps: int, threads: int = 1) -> np.ndarray:
    """The 3D 7-point star as the gold's own loop (po_star7_iterate_omp_f64):
    weights k in the term order e, w, n, s, t, b, c."""
    field = np.ascontiguousarray(field, dtype=np.float64)
    nz, ny, nx = field.shape
    kk = np.ascontiguousarray(k, dtype=np.float64)
    assert kk.shape == (7,)
    out = np.empty_like(field)
    lib().po_star7_iterate_omp_f64(p(field), p(out), nz, ny, nx, p(kk), 1 if boundary == "zero" else 0, steps,
                                   max(1, threads))
    return out


def iterate_fused(field: np.ndarray, st: PoStencil, steps: int) -> np.ndarray:
    """The contracted form: each group's product after the first fused into the
    running sum (po_stencil_iterate_fused_*)."""
    field = np.ascontiguousarray(field)
    nz, ny, nx = (1,) + field.shape if field.ndim == 2 else field.shape
    out = np.empty_like(field)
    fn = lib().po_stencil_iterate_fused_f64 if field.dtype == np.float64 else lib().po_stencil_iterate_fused_f32
    if fn(p(field), p(out), nz, ny, nx, C.byref(st), steps) != 0:
        raise ValueError("contracted form: FP32 stencil with double-formed groups")
    return out


def iterate16(field_bits: np.ndarray, st: PoStencil, steps: int, fmt: str) -> np.ndarray:
    """Reduced-precision storage: uint16 bit patterns of binary16 ('f16') or
    bfloat16 ('bf16'); each step in float, rounded to the format on store."""
    field_bits = np.ascontiguousarray(field_bits, dtype=np.uint16)
    if field_bits.ndim == 2:
        nz, (ny, nx) = 1, field_bits.shape
    else:
        nz, ny, nx = field_bits.shape
    out = np.empty_like(field_bits)
    fn = lib().po_stencil_iterate_f16s if fmt == "f16" else lib().po_stencil_iterate_bf16s
    fn(p(field_bits), p(out), nz, ny, nx, C.byref(st), steps)
    return out


def iterate16_sr(field_bits: np.ndarray, st: PoStencil, steps: int, fmt: str, seed: int) -> np.ndarray:
    """As iterate16, with stochastic rounding keyed by (cell, step, seed)."""
    field_bits = np.ascontiguousarray(field_bits, dtype=np.uint16)
    if field_bits.ndim == 2:
        nz, (ny, nx) = 1, field_bits.shape
    else:
        nz, ny, nx = field_bits.shape
    out = np.empty_like(field_bits)
    lib().po_stencil_iterate_sr16(p(field_bits), p(out), nz, ny, nx, C.byref(st), steps, 1 if fmt == "bf16" else 0,
                                  seed & 0xFFFFFFFF)
    return out


def to16(x: np.ndarray, fmt: str) -> np.ndarray:
    f = lib().po_f32_to_f16 if fmt == "f16" else lib().po_f32_to_bf16
    flat = np.asarray(x, dtype=np.float32).ravel()
    return np.array([f(float(v)) for v in flat], dtype=np.uint16).reshape(np.shape(x))


def from16(bits: np.ndarray, fmt: str) -> np.ndarray:
    if fmt == "bf16":
        return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32)
    return np.asarray(bits, dtype=np.uint16).view(np.float16).astype(np.float32)


def rms_error(a: np.ndarray, b: np.ndarray, box=None) -> tuple[float, float]:
    a3 = a.reshape((1,) + a.shape) if a.ndim == 2 else a
    b3 = np.ascontiguousarray(b.reshape(a3.shape))
    a3 = np.ascontiguousarray(a3)
    nz, ny, nx = a3.shape
    z0, z1, y0, y1, x0, x1 = box if box else (0, nz, 0, ny, 0, nx)
    mx = C.c_double()
    fn = lib().po_rms_error_f64 if a3.dtype == np.float64 else lib().po_rms_error_f32
    r = fn(p(a3), p(b3), ny, nx, z0, z1, y0, y1, x0, x1, C.byref(mx))
    return r, mx.value


# ---- sparse / CG ----------------------------------------------------------------
def poisson7_csr(nx, ny, nz):
    n = nx * ny * nz
    nnz = lib().po_poisson7_nnz(nx, ny, nz)
    rowptr = np.empty(n + 1, np.int32)
    col = np.empty(nnz, np.int32)
    val = np.empty(nnz, np.float64)
    lib().po_poisson7_csr(nx, ny, nz, p(rowptr), p(col), p(val))
    return rowptr, col, val


def poisson7_dia(nx, ny, nz):
    n = nx * ny * nz
    offs = np.empty(7, np.int32)
    val = np.empty(7 * n, np.float64)
    lib().po_poisson7_dia(nx, ny, nz, p(offs), p(val))
    return offs, val


def dot(a, b) -> float:
    if a.dtype == np.float32:
        return lib().po_dot_f32(p(a), p(b), a.size)
    return lib().po_dot_f64(p(a), p(b), a.size)


def cg(fmt, n, rowptr, col, val, ndiag, b, tol, relative, max_iter, x0=None):
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    hist = np.zeros(max_iter + 1)
    k = lib().po_cg_f64(fmt, n, p(rowptr) if rowptr is not None else None, p(col), p(val), ndiag, p(b), p(x),
                        tol, int(relative), max_iter, p(hist), hist.size)
    return x, hist[: k + 1], k


# ---- the reference's own gold (built from its sources into oracle/_ref) -------
REF_SYMBOLS = {
    ("3d", np.float64): "_Z18j3d_gold_iterativeIdEvPT_iiiS1_i",
    ("3d", np.float32): "_Z18j3d_gold_iterativeIfEvPT_iiiS1_i",
    ("2d", np.float64): "_Z21jacobi_gold_iterativeIdEvPT_iiS1_i",
    ("2d", np.float32): "_Z21jacobi_gold_iterativeIfEvPT_iiS1_i",
    ("3d4", np.float64): "_Z8j3d_goldIdEvPT_iiiS1_",
    ("3d4", np.float32): "_Z8j3d_goldIfEvPT_iiiS1_",
}


def ref_available(name: str) -> bool:
    return (REF_DIR / f"libref_{name}.so").exists()


def ref_gold_iterative(name: str, field: np.ndarray, steps: int) -> np.ndarray:
    """Runs the reference's j3d_gold_iterative / jacobi_gold_iterative from
    oracle/_ref/libref_<name>.so (e.g. name='3d7pt_gold_gen')."""
    l = C.CDLL(str(REF_DIR / f"libref_{name}.so"))
    field = np.ascontiguousarray(field).copy()
    out = np.zeros_like(field)
    if field.ndim == 3:
        f = getattr(l, REF_SYMBOLS[("3d", field.dtype.type)])
        f.restype = None
        f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int]
        f(p(field), *field.shape, p(out), steps)
    else:
        f = getattr(l, REF_SYMBOLS[("2d", field.dtype.type)])
        f.restype = None
        f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int]
        f(p(field), *field.shape, p(out), steps)
    return out


def ref_j3d_gold_4step(field: np.ndarray) -> np.ndarray:
    """The reference's interior-only 4-step j3d_gold (3d7pt_gold.cpp:28-45)."""
    l = C.CDLL(str(REF_DIR / "libref_3d7pt_gold.so"))
    field = np.ascontiguousarray(field).copy()
    out = np.zeros_like(field)
    f = getattr(l, REF_SYMBOLS[("3d4", field.dtype.type)])
    f.restype = None
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
    f(p(field), *field.shape, p(out))
    return out


def cg_f32(fmt, n, rowptr, col, val32, ndiag, b32, thr2, relative, max_iter, static_iter=False, x0=None):
    """FP32 CG (float data, double dots); thr2 is the squared threshold."""
    x = np.zeros(n, np.float32) if x0 is None else np.array(x0, dtype=np.float32)
    hist = np.zeros(max_iter + 1)
    k = lib().po_cg_f32(fmt, n, p(rowptr) if rowptr is not None else None, p(col), p(val32), ndiag, p(b32), p(x),
                        thr2, int(relative), max_iter, int(static_iter), p(hist), hist.size)
    return x, hist[: k + 1], k


def cg_f32_refine(fmt, n, rowptr, col, val, ndiag, b, tol, relative, inner_tol, max_inner, max_outer):
    x = np.zeros(n)
    hist = np.zeros(max_outer + 2)
    outer = C.c_int()
    total = lib().po_cg_f32_refine(fmt, n, p(rowptr) if rowptr is not None else None, p(col), p(val), ndiag, p(b),
                                   p(x), tol, int(relative), inner_tol, max_inner, max_outer, p(hist), hist.size,
                                   C.byref(outer))
    return x, hist[: outer.value + 1], total, outer.value


def csr_spmv(rowptr, col, val, x):
    """The oracle's CSR SpMV (fixed per-row order), FP64 or FP32 by val's dtype."""
    n = len(rowptr) - 1
    y = np.empty(n, dtype=val.dtype)
    f = lib().po_csr_spmv_f64 if val.dtype == np.float64 else lib().po_csr_spmv_f32
    f(n, p(np.ascontiguousarray(rowptr, np.int32)), p(np.ascontiguousarray(col, np.int32)), p(val), p(x), p(y))
    return y


from perks_amd.cg import skewed_spd_csr  # noqa: E402,F401  (test matrices; numpy only)
