"""ctypes binding of libperks_amd.so (the C ABI declared in include/perks_amd.h).

The library is built in-tree by ``python -m perks_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libperks_amd.so"

# ---- constants mirrored from include/perks_amd.h --------------------------------
PERKS_OK = 0
ERR_NO_LDS_FOR_CACHE = -1
ERR_UNSUPPORTED_SIZE = -2
ERR_NOT_LAUNCHABLE = -3
ERR_HEIGHT_TOO_SMALL = -4
ERR_SPILL = -5
ERR_INVALID_ARG = -10
ERR_HIP = -11
ERR_NO_DEVICE = -12
ERR_UNSUPPORTED = -13
ERR_NOT_CONVERGED = -14
ERR_BARRIER_TIMEOUT = -15

F64, F32, F16, BF16 = 0, 1, 2, 3
BC_CLAMP, BC_ZERO = 0, 1
METHOD_PLAIN, METHOD_PERSISTENT = 0, 1

SHAPES = {
    "3d_star_r1": 0, "3d_star_r2": 1, "3d_type0_r1": 2, "3d_box_r1": 3, "3d_poisson_r1": 4,
    "2d_star_r1": 10, "2d_star_r2": 11, "2d_star_r3": 12, "2d_star_r4": 13, "2d_star_r5": 14,
    "2d_star_r6": 15, "2d_box_r1": 16, "2d_box_r2": 17,
}
# reference directory name -> shape (the reference's 2D names clash: 2ds9pt is a
# radius-2 star, 2d9pt a radius-1 box)
REF_DIR_SHAPE = {
    "3d7pt": "3d_star_r1", "3d13pt": "3d_star_r2", "3d17pt": "3d_type0_r1", "3d27pt": "3d_box_r1",
    "poisson": "3d_poisson_r1", "2d5pt": "2d_star_r1", "2ds9pt": "2d_star_r2", "2d13pt": "2d_star_r3",
    "2d17pt": "2d_star_r4", "2d21pt": "2d_star_r5", "2ds25pt": "2d_star_r6", "2d9pt": "2d_box_r1",
    "2d25pt": "2d_box_r2",
}
MAX_COEFS = 27
MAX_PTS = 12


class StencilOpts(C.Structure):
    _fields_ = [
        ("shape", C.c_int32), ("boundary", C.c_int32), ("method", C.c_int32), ("use_ref_coefs", C.c_int32),
        ("coef", C.c_double * MAX_COEFS), ("divisor", C.c_double),
        ("bdimx", C.c_int32), ("blkpsm", C.c_int32), ("use_lds_cache", C.c_int32),
        ("reg_cache_planes", C.c_int32), ("double_tile", C.c_int32), ("warmup", C.c_int32),
        ("warmup_steps", C.c_int32), ("tune", C.c_int32 * 8),
        ("rounding", C.c_int32), ("round_seed", C.c_uint32),
    ]


class RunReport(C.Structure):
    _fields_ = [
        ("elapsed_ms", C.c_double), ("gcells_per_s", C.c_double), ("grid", C.c_int32 * 3),
        ("block", C.c_int32), ("lds_bytes", C.c_int32), ("lds_planes", C.c_int32),
        ("reg_planes", C.c_int32), ("result_in_b", C.c_int32), ("occupancy", C.c_int32),
        ("reserved", C.c_int32 * 7),
    ]


class DeviceCaps(C.Structure):
    _fields_ = [
        ("cu_count", C.c_int32), ("lds_per_cu", C.c_int32), ("lds_per_block", C.c_int32),
        ("vgprs_per_simd", C.c_int32), ("simds_per_cu", C.c_int32), ("warp_size", C.c_int32),
        ("cooperative", C.c_int32), ("reserved", C.c_int32 * 5),
    ]


class PlanInfo(C.Structure):
    _fields_ = [
        ("grid", C.c_int32 * 3), ("block", C.c_int32), ("occupancy", C.c_int32), ("lds_bytes", C.c_int32),
        ("lds_planes", C.c_int32), ("reg_planes", C.c_int32), ("min_height", C.c_int32),
        ("tile_x", C.c_int32), ("tile_y", C.c_int32), ("reserved", C.c_int32 * 5),
    ]


class KernelConfig(C.Structure):
    _fields_ = [
        ("method", C.c_int32), ("variant", C.c_int32), ("levels", C.c_int32), ("block", C.c_int32),
        ("vgprs", C.c_int32), ("spill_bytes", C.c_int32), ("lds_bytes", C.c_int32), ("occupancy", C.c_int32),
        ("tile_x", C.c_int32), ("tile_y", C.c_int32), ("status", C.c_int32), ("reserved0", C.c_int32),
        ("gcells_per_s", C.c_double), ("reserved", C.c_int32 * 4),
    ]


class Grid2d(C.Structure):
    _fields_ = [("a", C.c_void_p), ("b", C.c_void_p), ("width_y", C.c_int32), ("width_x", C.c_int32),
                ("steps", C.c_int32), ("result_in_b", C.c_int32)]


class Csr(C.Structure):
    _fields_ = [("n", C.c_int32), ("reserved0", C.c_int32), ("nnz", C.c_int64), ("rowptr", C.c_void_p),
                ("col", C.c_void_p), ("val", C.c_void_p)]


class Dia(C.Structure):
    _fields_ = [("n", C.c_int32), ("ndiag", C.c_int32), ("offsets", C.c_void_p), ("val", C.c_void_p)]


CG_F64, CG_F32_REFINE = 0, 1


class CgOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("inner_tol", C.c_double), ("relative", C.c_int32), ("max_iter", C.c_int32),
                ("precision", C.c_int32), ("max_inner", C.c_int32), ("max_outer", C.c_int32),
                ("x0_is_zero", C.c_int32), ("blkpsm", C.c_int32), ("static_iter", C.c_int32),
                ("cache_vector", C.c_int32), ("cache_matrix", C.c_int32), ("reserved", C.c_int32 * 2)]


class CgDriverOpts(C.Structure):
    _fields_ = [("mtx_path", C.c_char_p), ("value_bytes", C.c_int32), ("vdata_n", C.c_int32),
                ("max_iter", C.c_int32), ("static_iter", C.c_int32), ("warmup", C.c_int32), ("check", C.c_int32),
                ("verbose", C.c_int32), ("reserved0", C.c_int32), ("tol", C.c_double)]


class CgDriverResult(C.Structure):
    _fields_ = [("n", C.c_int32), ("iterations", C.c_int32), ("nnz", C.c_int64), ("residual", C.c_double),
                ("elapsed_ms", C.c_double), ("ms_per_iter", C.c_double), ("gbps", C.c_double),
                ("check_error", C.c_double), ("grid", C.c_int32), ("code", C.c_int32)]


class CgReport(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("outer_iterations", C.c_int32), ("final_rr", C.c_double),
                ("bb", C.c_double), ("elapsed_ms", C.c_double), ("grid", C.c_int32), ("hist_cap", C.c_int32),
                ("history", C.POINTER(C.c_double)), ("hist_len", C.c_int32), ("lds_bytes", C.c_int32),
                ("cache_units", C.c_int32), ("cache_matrix", C.c_int32), ("long_rows", C.c_int32),
                ("reserved", C.c_int32 * 1)]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double

# name -> (restype, argtypes); every symbol include/perks_amd.h declares.
SIGNATURES = {
    "perks_version": (C.c_char_p, []),
    "perks_error_string": (C.c_char_p, [_I]),
    "perks_query_device": (_I, [_I, C.POINTER(DeviceCaps)]),
    "perks_release": (_I, []),
    "perks_stencil_opts_init": (None, [C.POINTER(StencilOpts), _I]),
    "perks_shape_info": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _P]),
    "perks_j3d_plan": (_I, [_I, _I, _I, _I, C.POINTER(StencilOpts), C.POINTER(DeviceCaps), _I,
                            C.POINTER(PlanInfo)]),
    "perks_j3d_min_height": (_I, [_I, _I, _I, C.POINTER(StencilOpts)]),
    "perks_get_min_width_y": (_I, [_I, _I, _I, _I, _I]),
    "perks_get_min_width_y_2d": (_I, [_I, _I, _I, _I, _I, _I, _I]),
    "perks_jacobi2d_batch_noalloc": (_I, [_I, C.POINTER(Grid2d), _I, C.POINTER(StencilOpts), _P,
                                          C.POINTER(RunReport)]),
    "perks_j3d_iterative_f64": (_I, [_P, _I, _I, _I, _P, _I, C.POINTER(StencilOpts), C.POINTER(RunReport)]),
    "perks_j3d_iterative_f32": (_I, [_P, _I, _I, _I, _P, _I, C.POINTER(StencilOpts), C.POINTER(RunReport)]),
    "perks_j3d_device_f64": (_I, [_P, _P, _I, _I, _I, _I, C.POINTER(StencilOpts), _P, C.POINTER(RunReport)]),
    "perks_j3d_device_f32": (_I, [_P, _P, _I, _I, _I, _I, C.POINTER(StencilOpts), _P, C.POINTER(RunReport)]),
    "perks_j3d_device": (_I, [_I, _P, _P, _I, _I, _I, _I, C.POINTER(StencilOpts), _P, C.POINTER(RunReport)]),
    "perks_j3d_kernel_table": (_I, [_I, _I, _I, _I, C.POINTER(StencilOpts), C.POINTER(KernelConfig), _I,
                                    C.POINTER(C.c_int)]),
    "perks_j3d_autotune": (_I, [_I, _P, _I, _I, _I, _I, C.POINTER(StencilOpts), C.POINTER(KernelConfig), _I,
                                C.POINTER(C.c_int), _P]),
    "perks_fill_uniform_f64": (_I, [_P, C.c_size_t, C.c_uint64, _D, _D, _P]),
    "perks_fill_uniform_f32": (_I, [_P, C.c_size_t, C.c_uint64, _D, _D, _P]),
    "perks_fill_uniform_f16": (_I, [_P, C.c_size_t, C.c_uint64, _D, _D, _P]),
    "perks_fill_uniform_bf16": (_I, [_P, C.c_size_t, C.c_uint64, _D, _D, _P]),
    "perks_jacobi2d_iterative_f64": (_I, [_P, _I, _I, _P, _I, C.POINTER(StencilOpts), C.POINTER(RunReport)]),
    "perks_jacobi2d_iterative_f32": (_I, [_P, _I, _I, _P, _I, C.POINTER(StencilOpts), C.POINTER(RunReport)]),
    "perks_jacobi2d_device": (_I, [_I, _P, _P, _I, _I, _I, C.POINTER(StencilOpts), _P, C.POINTER(RunReport)]),
    "perks_jacobi2d_batch": (_I, [_I, C.POINTER(Grid2d), _I, C.POINTER(StencilOpts), _P, C.POINTER(RunReport)]),
    "perks_jacobi2d_batch_plan": (_I, [_I, C.POINTER(Grid2d), _I, C.POINTER(StencilOpts), _I, _P,
                                       C.POINTER(C.c_void_p)]),
    "perks_jacobi2d_batch_run": (_I, [C.c_void_p, _P, C.POINTER(RunReport)]),
    "perks_jacobi2d_batch_plan_info": (_I, [C.c_void_p, C.POINTER(C.c_int32), _I, C.POINTER(C.c_int32)]),
    "perks_jacobi2d_batch_plan_destroy": (None, [C.c_void_p]),
    "perks_cg_opts_init": (None, [C.POINTER(CgOpts)]),
    "perks_cg_solve_csr": (_I, [C.POINTER(Csr), _P, _P, C.POINTER(CgOpts), _P, C.POINTER(CgReport)]),
    "perks_cg_solve_dia": (_I, [C.POINTER(Dia), _P, _P, C.POINTER(CgOpts), _P, C.POINTER(CgReport)]),
    "perks_cg_solve_csr_f32": (_I, [C.POINTER(Csr), _P, _P, C.POINTER(CgOpts), _P, C.POINTER(CgReport)]),
    "perks_cg_solve_dia_f32": (_I, [C.POINTER(Dia), _P, _P, C.POINTER(CgOpts), _P, C.POINTER(CgReport)]),
    "perks_cg_driver": (_I, [C.POINTER(CgDriverOpts), C.POINTER(CgDriverResult)]),
    "perks_mm_read_csr": (_I, [C.c_char_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_void_p),
                               C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "perks_free_host": (None, [C.c_void_p]),
    "perks_gen_tridiag_f64": (_I, [C.c_int32, _P, _P, _P]),
    "perks_gen_tridiag_f32": (_I, [C.c_int32, _P, _P, _P]),
    "perks_csr_to_dia": (_I, [C.c_int32, _P, _P, _P, C.c_int32, _P, _P]),
    "perks_poisson7_nnz": (C.c_int64, [_I, _I, _I]),
    "perks_poisson7_csr": (_I, [_I, _I, _I, _P, _P, _P, _P]),
    "perks_poisson7_dia": (_I, [_I, _I, _I, _P, _P, _P]),
}

_lib = None


class PerksError(RuntimeError):
    def __init__(self, code: int, what: str = ""):
        self.code = code
        msg = error_string(code) if _lib is not None else str(code)
        super().__init__(f"{what}: {msg} (code {code})" if what else f"{msg} (code {code})")


def load() -> C.CDLL:
    """Loads libperks_amd.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m perks_amd.build` "
            "(the package has no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def error_string(code: int) -> str:
    return load().perks_error_string(code).decode()


def check(code: int, what: str = "") -> int:
    if code < 0:
        raise PerksError(code, what)
    return code
