"""y tiling of the persistent 7-point kernel.

A workgroup's R rows store R - 2(L-1) rows; at heights of at least
2 * (R - (L-1)) the two edge tiles start at the y faces and store R - (L-1)
rows each (face rows take the boundary rule by selects at every level).
Heights around that threshold and around the next tile-count steps, at
several pass depths and both boundaries, must give the CPU oracle's bits
(FP64) and the per-step kernel's (FP32, BF16, the two-column variant).
"""
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P

pytestmark = pytest.mark.gpu

CONFIG_3D = [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4]
# rows per workgroup of each 7-point persistent variant (waves x rows per wave)
ROWS_F64 = {0: 36, 2: 16, 3: 40, 4: 32}
ROWS_F32 = {0: 64, 2: 32, 4: 32}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def heights(rows, depth):
    eh, vh = rows - (depth - 1), rows - 2 * (depth - 1)
    t = 2 * eh
    return sorted({t - 1, t, t + 1, t + vh - 1, t + vh, t + vh + 1})


@pytest.mark.parametrize("boundary", ["zero", "clamp"])
@pytest.mark.parametrize("variant", sorted(ROWS_F64))
def test_fp64_persistent_matches_oracle_across_heights(gpu, variant, boundary):
    si = P.shape_info("3d_star_r1")
    st = O.stencil_from_shape(si, boundary, coefs=CONFIG_3D)
    for depth in (2, 4):
        for ny in heights(ROWS_F64[variant], depth):
            dims, steps = (7, ny, 70), 2 * depth + 1
            say(f"[ytiles] f64 variant={variant} {boundary} depth={depth} dims={dims}")
            field = O.uniform(int(np.prod(dims)), ny * 13 + depth, 0.0, 1.0).reshape(dims)
            got, rep = P.j3d_run(field, steps, method="persistent", boundary=boundary, coefs=CONFIG_3D,
                                 tune=[0, 0, depth, 0, variant])
            assert rep.reserved[0] == depth
            want = O.iterate(field, st, steps)
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), f"ny={ny} depth={depth}"


@pytest.mark.parametrize("variant", sorted(ROWS_F32))
@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_fp32_arithmetic_equals_per_step_across_heights(gpu, dtype, variant):
    torch = gpu
    tdt = torch.float32 if dtype == "f32" else torch.int16
    nz, nx, steps = 9, 130, 9
    for boundary in ("zero", "clamp"):
        for ny in heights(ROWS_F32[variant], 4):
            say(f"[ytiles] {dtype} variant={variant} {boundary} ny={ny}")
            init = torch.empty(nz * ny * nx, dtype=tdt, device="cuda")
            P.fill_uniform(init, init.numel(), ny, -1.0, 1.0, dtype=dtype)
            outs = []
            for method, tune in (("plain", None), ("persistent", [0, 0, 4, 0, variant])):
                a = init.clone()
                b = torch.empty_like(a)
                rep = P.j3d_device(a, b, nz, ny, nx, steps, dtype=dtype, method=method, boundary=boundary,
                                   coefs=CONFIG_3D, tune=tune)
                torch.cuda.synchronize()
                outs.append(b if rep.result_in_b else a)
            assert torch.equal(outs[1], outs[0]), f"ny={ny} {boundary}"


def test_tiles_cross_many_slabs(gpu):
    """A 96 x 200 x 250 field: several z-slabs per tile column, x and y edge
    tiles at both faces, a short last pass."""
    si = P.shape_info("3d_star_r1")
    dims, steps = (96, 200, 250), 10
    field = O.uniform(int(np.prod(dims)), 77, 0.0, 1.0).reshape(dims)
    want = O.iterate(field, O.stencil_from_shape(si, "zero", coefs=CONFIG_3D), steps)
    for variant in (0, 4):
        say(f"[ytiles] f64 dims={dims} variant={variant}")
        got, _ = P.j3d_run(field, steps, method="persistent", boundary="zero", coefs=CONFIG_3D,
                           tune=[0, 0, 4, 0, variant])
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), f"variant={variant}"
