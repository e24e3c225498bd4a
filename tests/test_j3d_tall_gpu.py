"""The streaming 3D kernel's tall tile (opts.tune[4] = 4: 64 x 16 cells per
wave, FP32 arithmetic), every shape, against the oracle bit for bit: FP32
storage at several grid shapes, both boundaries and both methods, both pass
depths, and FP16/BF16 storage with both roundings; FP64 storage refuses it.
"""
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P
from perks_amd import _lib as L

pytestmark = pytest.mark.gpu

SHAPES = ["3d7pt", "3d13pt", "3d17pt", "3d27pt", "poisson"]
TALL = 4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def tall(shape, depth=0):
    return [0, 0, depth, 0, TALL, 1 if shape == "3d7pt" else 0]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("method", ["plain", "persistent"])
@pytest.mark.parametrize("boundary", ["clamp", "zero"])
def test_tall_tile_matches_oracle(gpu, shape, method, boundary):
    si = P.shape_info(shape)
    st = O.stencil_from_shape(si, boundary)
    for dims, steps in (((19, 37, 70), 7), ((6, 5, 130), 3), ((3, 1, 1), 4), ((40, 33, 61), 5)):
        say(f"[tall] {shape} {method} {boundary} dims={dims} steps={steps}")
        field = O.libc_rand(int(np.prod(dims)), 5, np.float32).reshape(dims)
        got, _ = P.j3d_run(field, steps, shape=shape, method=method, boundary=boundary, tune=tall(shape))
        want = O.iterate(field, st, steps)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), f"dims={dims}"


@pytest.mark.parametrize("shape", ["3d13pt", "3d17pt", "3d27pt", "poisson"])
@pytest.mark.parametrize("depth", [1, 2])
def test_tall_tile_pass_depths(gpu, shape, depth):
    dims, steps = (23, 50, 100), 8
    si = P.shape_info(shape)
    rng = np.random.default_rng(depth)
    c = rng.random(len(si.coef))
    coefs = list(c / c.sum())
    field = O.uniform(int(np.prod(dims)), 9, -1.0, 1.0, dtype=np.float32).reshape(dims)
    say(f"[tall] {shape} depth={depth}")
    got, rep = P.j3d_run(field, steps, shape=shape, method="persistent", boundary="zero", coefs=coefs,
                         tune=tall(shape, depth))
    assert rep.reserved[0] == (depth if si.radius == 1 else 1)
    want = O.iterate(field, O.stencil_from_shape(si, "zero", coefs=coefs, divisor=0.0), steps)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("shape", ["3d27pt", "3d13pt"])
@pytest.mark.parametrize("fmt", ["bf16", "f16"])
@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
def test_tall_tile_16bit_storage(gpu, shape, fmt, rounding):
    torch = gpu
    dims, steps = (11, 40, 90), 5
    si = P.shape_info(shape)
    n = int(np.prod(dims))
    for method in ("plain", "persistent"):
        say(f"[tall] {shape} {fmt} {rounding} {method}")
        a = torch.empty(n, dtype=torch.int16, device="cuda")
        b = torch.empty_like(a)
        P.fill_uniform(a, n, 13, 0.0, 1.0, dtype=fmt)
        init = a.cpu().numpy().view(np.uint16).reshape(dims).copy()
        rep = P.j3d_device(a, b, *dims, steps, dtype=fmt, shape=shape, method=method, boundary="clamp",
                           rounding=rounding, round_seed=21, tune=tall(shape))
        got = (b if rep.result_in_b else a).cpu().numpy().view(np.uint16).reshape(dims)
        st = O.stencil_from_shape(si, "clamp")
        want = O.iterate16_sr(init, st, steps, fmt, 21) if rounding == "stochastic" else O.iterate16(init, st, steps, fmt)
        assert np.array_equal(got, want), method


def test_tall_tile_refused_for_fp64(gpu):
    with pytest.raises(L.PerksError) as e:
        P.j3d_run(np.zeros((4, 16, 16)), 2, shape="3d27pt", method="persistent", boundary="zero", tune=tall("3d27pt"))
    assert e.value.code == L.ERR_UNSUPPORTED
