#!/usr/bin/env python3
"""Writes the golden vectors of the reference's own CPU gold.

Runs the reference's j3d_gold_iterative / jacobi_gold_iterative (every 3D and
2D shape, FP64 and FP32), compiled from the reference sources into oracle/_ref
by oracle/Makefile, on small glibc-rand() inputs -- the reference's input form
(stencil/common/common.hpp:24-27) -- and stores input and output as data in
tests/golden/ref_gold.npz.  Needs oracle/_ref (this container only); the
fixture itself travels with the tests.

    make -C oracle ref && python3 tests/golden/make_golden.py
"""
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import oracle as O  # noqa: E402

CASES_3D = {"3d7pt": "3d7pt_gold_gen", "3d13pt": "3d13pt_gold_gen", "3d17pt": "3d17pt_gold_gen",
            "3d27pt": "3d27pt_gold_gen", "poisson": "poisson_gold_gen"}
CASES_2D = {"2d5pt": "2d5pt_gold_gen", "2ds9pt": "2ds9pt_gold_gen", "2d13pt": "2d13pt_gold_gen",
            "2d17pt": "2d17pt_gold_gen", "2d21pt": "2d21pt_gold_gen", "2ds25pt": "2ds25pt_gold_gen",
            "2d9pt": "2d9pt_gold_gen", "2d25pt": "2d25pt_gold_gen"}
DIMS_3D, STEPS_3D, SEED_3D = (8, 9, 10), 3, 1
DIMS_2D, STEPS_2D, SEED_2D = (24, 20), 5, 3


def main():
    out = {}
    for cases, dims, steps, seed in ((CASES_3D, DIMS_3D, STEPS_3D, SEED_3D), (CASES_2D, DIMS_2D, STEPS_2D, SEED_2D)):
        for shape, lib in cases.items():
            if not O.ref_available(lib):
                sys.exit(f"oracle/_ref/libref_{lib}.so missing: run `make -C oracle ref` first")
            for dt, tag in ((np.float64, "f64"), (np.float32, "f32")):
                field = O.libc_rand(int(np.prod(dims)), seed, dt).reshape(dims)
                out[f"{shape}_{tag}_in"] = field
                out[f"{shape}_{tag}_out"] = O.ref_gold_iterative(lib, field, steps)
    out["steps_3d"] = np.int32(STEPS_3D)
    out["steps_2d"] = np.int32(STEPS_2D)
    # j3d_gold (3d7pt_gold.cpp:28-45): four interior-only steps, zero boundary in the input
    for dt, tag in ((np.float64, "f64"), (np.float32, "f32")):
        f = O.libc_rand(10 * 9 * 11, 5, dt).reshape(10, 9, 11)
        f[0], f[-1], f[:, 0], f[:, -1], f[:, :, 0], f[:, :, -1] = 0, 0, 0, 0, 0, 0
        out[f"j3d_gold4_{tag}_in"] = f
        out[f"j3d_gold4_{tag}_out"] = O.ref_j3d_gold_4step(f)
    # genTridiag (conjugateGradient/matrix/tridiag.cpp:11-47), float and double,
    # from the C library's rand() seeded as in a fresh process (srand(1)).
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    tri = ctypes.CDLL(str(O.REF_DIR / "libref_tridiag.so"))
    n_tri = 50
    nz = (n_tri - 2) * 3 + 4
    for tag, dt, sym in (("f32", np.float32, "_Z10genTridiagIfiEvPT0_S1_PT_S0_S0_"),
                         ("f64", np.float64, "_Z10genTridiagIdiEvPT0_S1_PT_S0_S0_")):
        rowptr = np.zeros(n_tri + 1, np.int32)
        col = np.zeros(nz, np.int32)
        val = np.zeros(nz, dt)
        libc.srand(1)
        fn = getattr(tri, sym)
        fn.restype = None
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        fn(rowptr.ctypes.data, col.ctypes.data, val.ctypes.data, n_tri, nz)
        out[f"tridiag_{tag}_rowptr"], out[f"tridiag_{tag}_col"], out[f"tridiag_{tag}_val"] = rowptr, col, val
    np.savez(HERE / "ref_gold.npz", **out)
    print(f"wrote {HERE / 'ref_gold.npz'}: {len(out)} arrays")


if __name__ == "__main__":
    main()
