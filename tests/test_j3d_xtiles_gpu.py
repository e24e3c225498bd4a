"""x tiling of the persistent 7-point kernel under the zero boundary.

At widths of at least 2 * (64 - L) the two edge tiles start at the x faces
and store 64 - L columns (the shifts bring the boundary's zeros into lanes 0
and 63), interior tiles 64 - 2L.  Widths around that threshold and around
each tile-count step, at every pass depth, every variant of one column per
lane, and FP64 / FP32 / BF16 storage, must give the per-step kernel's bits
and (FP64) the CPU oracle's.
"""
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P

pytestmark = pytest.mark.gpu

CONFIG_3D = [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4]
WIDTHS = [119, 120, 121, 122, 126, 127, 128, 175, 176, 177, 178, 231, 232, 233, 300]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("depth", [1, 2, 3, 4])
@pytest.mark.parametrize("nx", WIDTHS)
def test_fp64_persistent_matches_oracle_across_widths(gpu, nx, depth):
    dims, steps = (9, 21, nx), 9
    say(f"[xtiles] f64 dims={dims} depth={depth}")
    si = P.shape_info("3d_star_r1")
    field = O.uniform(int(np.prod(dims)), nx * 7 + depth, 0.0, 1.0).reshape(dims)
    got, rep = P.j3d_run(field, steps, method="persistent", boundary="zero", coefs=CONFIG_3D, tune=[0, 0, depth])
    assert rep.reserved[0] == depth
    want = O.iterate(field, O.stencil_from_shape(si, "zero", coefs=CONFIG_3D), steps)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("dtype", ["f32", "bf16"])
def test_variants_equal_per_step_across_widths(gpu, dtype, variant):
    torch = gpu
    tdt = torch.float32 if dtype == "f32" else torch.int16
    # the BF16 per-step kernel needs even widths
    for nx in ((120, 121, 177, 232, 300) if dtype == "f32" else (120, 122, 176, 232, 300)):
        nz, ny, steps = 10, 37, 11
        say(f"[xtiles] {dtype} variant={variant} nx={nx}")
        init = torch.empty(nz * ny * nx, dtype=tdt, device="cuda")
        P.fill_uniform(init, init.numel(), nx, -1.0, 1.0, dtype=dtype)
        outs = []
        for method, tune in (("plain", None), ("persistent", [0, 0, 3, 0, variant]),
                             ("persistent", [0, 0, 4, 0, variant])):
            a = init.clone()
            b = torch.empty_like(a)
            rep = P.j3d_device(a, b, nz, ny, nx, steps, dtype=dtype, method=method, boundary="zero",
                               coefs=CONFIG_3D, tune=tune)
            torch.cuda.synchronize()
            outs.append(b if rep.result_in_b else a)
        assert torch.equal(outs[1], outs[0]) and torch.equal(outs[2], outs[0]), f"nx={nx}"
