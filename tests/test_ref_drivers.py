"""The drop-in boundary, shown with the reference's own drivers: the PERKS
tree's stencil/3dstencil/j3d.driver.cpp and stencil/2dstencil/jacobi.driver.cpp
compile unmodified against include/perks_amd/jacobi_cuda.hpp (the reference's
declaration header kept out by its include guard PERKS_CUDA_HEADER) and link
with libperks_amd.so (oracle/Makefile, target `drivers`).  Needs the PERKS
source tree, so it runs where that tree is present; the built executables are
run on the GPU by tests/test_ref_drivers_gpu.py."""
import pathlib
import subprocess

import pytest

from perks_amd import _lib as L

ROOT = pathlib.Path(__file__).resolve().parent.parent
REF = pathlib.Path("/root/reference/stencil")
OUT = ROOT / "oracle" / "_ref"

pytestmark = pytest.mark.skipif(not (REF / "3dstencil" / "j3d.driver.cpp").exists(),
                                reason="the PERKS source tree is not present")


def _undefined_symbols(exe):
    out = subprocess.run(["nm", "-D", "--undefined-only", str(exe)], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


@pytest.fixture(scope="module")
def drivers():
    if not L.LIB_PATH.exists():
        pytest.skip("libperks_amd.so is not built")
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "drivers"], check=True)
    return OUT / "j3d_driver_3d7pt", OUT / "jacobi_driver_2d5pt"


def test_reference_3d_driver_builds_and_binds_the_abi(drivers):
    j3d, _ = drivers
    assert j3d.exists()
    used = {s for s in _undefined_symbols(j3d) if s.startswith("perks_")}
    # j3d_iterative<float/double>, getminHeight (the --small loop) and getMinWidthY (--checkmindomain)
    assert {"perks_j3d_iterative_f32", "perks_j3d_iterative_f64", "perks_j3d_min_height",
            "perks_get_min_width_y"} <= used, used
    assert "perks_get_min_width_y_2d" not in used


def test_reference_2d_driver_builds_and_binds_the_abi(drivers):
    _, j2d = drivers
    assert j2d.exists()
    used = {s for s in _undefined_symbols(j2d) if s.startswith("perks_")}
    # jacobi_iterative<float/double>; the 6- and 3-argument getMinWidthY calls
    # (jacobi.driver.cpp:95-142) bind the 2D overloads, not the 3D one
    assert {"perks_jacobi2d_iterative_f32", "perks_jacobi2d_iterative_f64", "perks_get_min_width_y_2d"} <= used, used
    assert "perks_get_min_width_y" not in used


def test_dropin_header_has_every_reference_signature(tmp_path):
    """Takes the address of each reference template with its exact reference
    signature (3D: common/jacobi_cuda.cuh:7,12; 2D: common/jacobi_cuda.cuh:6,14,16)."""
    src = tmp_path / "sig.cpp"
    src.write_text('''#include "perks_amd/jacobi_cuda.hpp"
int (*j3f)(float*, int, int, int, float*, int, int, int, bool, bool, int, bool, bool) = &j3d_iterative<float>;
int (*j3d)(double*, int, int, int, double*, int, int, int, bool, bool, int, bool, bool) = &j3d_iterative<double>;
int (*mw3)(int, int, int, bool) = &getMinWidthY<float>;
int (*j2f)(float*, int, int, float*, int, int, int, bool, bool, bool, int, bool) = &jacobi_iterative<float>;
int (*j2d)(double*, int, int, double*, int, int, int, bool, bool, bool, int, bool) = &jacobi_iterative<double>;
int (*mw2a)(int, int, int, bool, int, bool) = &getMinWidthY<double>;
int (*mw2b)(int, int, bool) = &getMinWidthY<float>;
''')
    subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-I", str(ROOT / "include"), str(src)], check=True)
