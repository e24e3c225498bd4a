"""perks_amd -- persistent Jacobi stencils and CG for MI355X (gfx950).

The compute lives in libperks_amd.so (hand-written HIP kernels behind a C ABI,
include/perks_amd.h).  This package is the host-side mirror of the reference's
entry points; it loads the library and fails loudly if it is missing.
"""
from . import _lib
from .cg import (CgResult, cg_driver, cg_solve, csr_to_dia, gen_tridiag, mm_read_csr, poisson7_csr, poisson7_dia,
                 poisson7_nnz, skewed_spd_csr)
from ._lib import PerksError, check, load


def release() -> None:
    """perks_release(): frees the library's pooled device scratch on the
    current device (batch plans keep what they own)."""
    check(load().perks_release(), "perks_release")

from .stencil import (BatchPlan, fill_uniform, j3d_autotune, j3d_device, j3d_iterative, j3d_kernel_table, j3d_run,
                      jacobi2d_batch, jacobi2d_device, jacobi2d_run, jacobi_iterative, make_opts, plan, query_device, shape_id, shape_info)

__all__ = [
    "BatchPlan", "CgResult", "cg_driver", "cg_solve", "csr_to_dia", "gen_tridiag", "mm_read_csr", "poisson7_csr", "poisson7_dia", "poisson7_nnz", "skewed_spd_csr",
    "PerksError", "check", "load", "release", "fill_uniform", "j3d_autotune", "j3d_device", "j3d_kernel_table", "j3d_iterative", "j3d_run", "jacobi2d_batch",
    "jacobi2d_device", "jacobi2d_run", "jacobi_iterative", "make_opts", "plan", "query_device", "shape_id",
    "shape_info", "_lib",
]
