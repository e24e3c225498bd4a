"""Every 3D shape of the reference (7- and 13-point stars, 17-point type-0,
27-point box, 19-point Poisson) on both shape-generic 3D kernels (register
streaming and LDS planes), bit for bit against the oracle and against the
reference gold's committed vectors."""
import pathlib
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P

pytestmark = pytest.mark.gpu

SHAPES = ["3d7pt", "3d13pt", "3d17pt", "3d27pt", "poisson"]
GOLD = np.load(pathlib.Path(__file__).resolve().parent / "golden" / "ref_gold.npz")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_equal(got, want, what):
    if not np.array_equal(got.view(np.uint8), want.view(np.uint8)):
        bad = np.argwhere(got != want)
        pytest.fail(f"{what}: {len(bad)} cells differ, first {bad[:3].tolist()}")


KERNELS = {"stream": 2, "lds": 3}  # opts.tune[4] on the shape-generic path


def generic(shape, kernel="stream", depth=0):
    # the 7-point star has its own kernels; tune[5] = 1 sends it to the generic ones
    return [0, 0, depth, 0, KERNELS[kernel], 1 if shape == "3d7pt" else 0]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("method", ["plain", "persistent"])
@pytest.mark.parametrize("kernel", ["stream", "lds"])
def test_3d_shapes_match_committed_reference_gold(gpu, shape, tag, method, kernel):
    field = GOLD[f"{shape}_{tag}_in"]
    steps = int(GOLD["steps_3d"])
    say(f"[j3d-shapes] reference gold {shape} {tag} {method} {kernel} dims={field.shape}")
    got, _ = P.j3d_run(field, steps, shape=shape, method=method, boundary="clamp", tune=generic(shape, kernel))
    check_equal(got, GOLD[f"{shape}_{tag}_out"], f"{shape} {tag} {method}")


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("method", ["plain", "persistent"])
@pytest.mark.parametrize("boundary", ["clamp", "zero"])
@pytest.mark.parametrize("dims,steps", [((19, 37, 70), 7), ((6, 5, 130), 3), ((3, 1, 1), 4), ((40, 9, 61), 5)])
@pytest.mark.parametrize("kernel", ["stream", "lds"])
def test_3d_shapes_match_oracle(gpu, shape, dtype, method, boundary, dims, steps, kernel):
    say(f"[j3d-shapes] {shape} {np.dtype(dtype).name} {method} {boundary} {kernel} dims={dims} steps={steps}")
    si = P.shape_info(shape)
    field = O.libc_rand(int(np.prod(dims)), 5, dtype).reshape(dims)
    got, _ = P.j3d_run(field, steps, shape=shape, method=method, boundary=boundary, tune=generic(shape, kernel))
    want = O.iterate(field, O.stencil_from_shape(si, boundary), steps)
    check_equal(got, want, f"{shape} {method} {boundary}")


@pytest.mark.parametrize("shape", ["3d13pt", "3d27pt", "poisson"])
@pytest.mark.parametrize("depth", [1, 2, 3])
@pytest.mark.parametrize("kernel", ["stream", "lds"])
def test_3d_shapes_pass_depths_and_own_weights(gpu, shape, depth, kernel):
    dims, steps = (23, 30, 100), 8
    say(f"[j3d-shapes] {shape} depth={depth} {kernel}")
    si = P.shape_info(shape)
    rng = np.random.default_rng(depth)
    c = rng.random(len(si.coef))
    coefs = list(c / c.sum())
    field = O.uniform(int(np.prod(dims)), 9, -1.0, 1.0).reshape(dims)
    got, rep = P.j3d_run(field, steps, shape=shape, method="persistent", boundary="zero", coefs=coefs,
                         tune=generic(shape, kernel, depth))
    # LDS kernel: FP64 tiles are 64 x 16, a level's ring of 2R+1 planes takes 24 KiB (R=1) or 40 KiB
    # (R=2) of the 64 KiB; streaming kernel: built for 2 levels at radius 1, 1 at radius 2
    assert rep.reserved[0] == (min(depth, 2) if si.radius == 1 else 1)
    want = O.iterate(field, O.stencil_from_shape(si, "zero", coefs=coefs, divisor=0.0), steps)
    check_equal(got, want, f"{shape} depth {depth}")


@pytest.mark.parametrize("shape", ["3d27pt", "3d13pt"])
@pytest.mark.parametrize("fmt", ["bf16", "f16"])
@pytest.mark.parametrize("method", ["plain", "persistent"])
@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
@pytest.mark.parametrize("kernel", ["stream", "lds"])
def test_3d_shapes_16bit_storage(gpu, shape, fmt, method, rounding, kernel):
    torch = gpu
    dims, steps = (11, 20, 90), 5
    say(f"[j3d-shapes] {shape} {fmt} {method} {rounding}")
    si = P.shape_info(shape)
    n = int(np.prod(dims))
    a = torch.empty(n, dtype=torch.int16, device="cuda")
    b = torch.empty_like(a)
    P.fill_uniform(a, n, 13, 0.0, 1.0, dtype=fmt)
    init = a.cpu().numpy().view(np.uint16).reshape(dims).copy()
    rep = P.j3d_device(a, b, *dims, steps, dtype=fmt, shape=shape, method=method, boundary="clamp",
                       rounding=rounding, round_seed=21, tune=generic(shape, kernel))
    got = (b if rep.result_in_b else a).cpu().numpy().view(np.uint16).reshape(dims)
    st = O.stencil_from_shape(si, "clamp")
    want = O.iterate16_sr(init, st, steps, fmt, 21) if rounding == "stochastic" else O.iterate16(init, st, steps, fmt)
    check_equal(got, want, f"{shape} {fmt} {method} {rounding}")


def test_generic_kernel_equals_7pt_kernels(gpu):
    torch = gpu
    dims, steps = (40, 64, 128), 9
    say("[j3d-shapes] generic vs 7-point kernels")
    field = O.uniform(int(np.prod(dims)), 3, 0.0, 1.0).reshape(dims)
    coefs = [0.1] * 6 + [0.4]
    ref, _ = P.j3d_run(field, steps, method="persistent", boundary="zero", coefs=coefs)
    for method in ("plain", "persistent"):
        got, _ = P.j3d_run(field, steps, method=method, boundary="zero", coefs=coefs, tune=[0, 0, 0, 0, 0, 1])
        check_equal(got, ref, f"generic {method}")
