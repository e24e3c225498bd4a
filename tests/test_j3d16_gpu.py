"""3D 7-point Jacobi with FP16 / BF16 storage on the GPU, bit for bit against
the oracle's reduced-precision iteration (FP32 arithmetic in the gold's order,
every step rounded to the storage format by round-to-nearest-even)."""
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P

pytestmark = pytest.mark.gpu

CONFIG_3D = [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run16(torch, dims, steps, fmt, method, boundary, coefs, seed=7, lo=0.0, hi=1.0, tune=None):
    n = int(np.prod(dims))
    a = torch.empty(n, dtype=torch.int16, device="cuda")
    b = torch.empty_like(a)
    P.fill_uniform(a, n, seed, lo, hi, dtype=fmt)
    init = a.cpu().numpy().view(np.uint16).reshape(dims).copy()
    rep = P.j3d_device(a, b, *dims, steps, dtype=fmt, method=method, boundary=boundary, coefs=coefs, tune=tune)
    got = (b if rep.result_in_b else a).cpu().numpy().view(np.uint16).reshape(dims)
    return init, got, rep


@pytest.mark.parametrize("fmt", ["bf16", "f16"])
@pytest.mark.parametrize("method", ["plain", "persistent"])
@pytest.mark.parametrize("boundary", ["zero", "clamp"])
@pytest.mark.parametrize("dims,steps", [((9, 40, 130), 5), ((5, 33, 64), 1), ((20, 70, 258), 9), ((3, 4, 6), 2)])
def test_j3d_16bit_storage_matches_oracle(gpu, fmt, method, boundary, dims, steps):
    torch = gpu
    say(f"[j3d16] {fmt} {method} {boundary} dims={dims} steps={steps}")
    coefs = CONFIG_3D if boundary == "zero" else None
    init, got, _ = run16(torch, dims, steps, fmt, method, boundary, coefs)
    want_init = O.to16(O.uniform(int(np.prod(dims)), 7, 0.0, 1.0, np.float32), fmt).reshape(dims)
    assert np.array_equal(init, want_init), "device 16-bit fill differs from the host sequence"
    st = O.stencil_from_shape(P.shape_info("3d7pt"), boundary, coefs=coefs)
    want = O.iterate16(init, st, steps, fmt)
    if not np.array_equal(got, want):
        bad = np.argwhere(got != want)
        pytest.fail(f"{len(bad)} cells differ, first {bad[:4].tolist()}")


@pytest.mark.parametrize("fmt", ["bf16", "f16"])
def test_j3d_16bit_persistent_equals_plain_and_pass_depths(gpu, fmt):
    torch = gpu
    dims, steps = (24, 50, 200), 11
    say(f"[j3d16] plain vs persistent depths {fmt}")
    _, plain, _ = run16(torch, dims, steps, fmt, "plain", "zero", CONFIG_3D, seed=3, lo=-1.0, hi=1.0)
    for depth in (1, 2, 3, 4):
        _, pers, rep = run16(torch, dims, steps, fmt, "persistent", "zero", CONFIG_3D, seed=3, lo=-1.0, hi=1.0,
                             tune=[0, 0, depth])
        assert rep.reserved[0] == depth
        assert np.array_equal(pers, plain), f"depth {depth}"


def test_j3d_16bit_plain_needs_even_width(gpu):
    torch = gpu
    a = torch.zeros(4 * 5 * 7, dtype=torch.int16, device="cuda")
    b = torch.empty_like(a)
    with pytest.raises(P.PerksError) as e:
        P.j3d_device(a, b, 4, 5, 7, 2, dtype="bf16", method="plain", boundary="zero", coefs=CONFIG_3D)
    assert e.value.code == -2
    rep = P.j3d_device(a, b, 4, 5, 7, 2, dtype="bf16", method="persistent", boundary="zero", coefs=CONFIG_3D)
    assert rep.result_in_b in (0, 1)


def run16_sr(torch, dims, steps, fmt, method, seed, round_seed, tune=None, lo=0.0, hi=1.0):
    n = int(np.prod(dims))
    a = torch.empty(n, dtype=torch.int16, device="cuda")
    b = torch.empty_like(a)
    P.fill_uniform(a, n, seed, lo, hi, dtype=fmt)
    init = a.cpu().numpy().view(np.uint16).reshape(dims).copy()
    rep = P.j3d_device(a, b, *dims, steps, dtype=fmt, method=method, boundary="zero", coefs=CONFIG_3D, tune=tune,
                       rounding="stochastic", round_seed=round_seed)
    return init, (b if rep.result_in_b else a).cpu().numpy().view(np.uint16).reshape(dims), rep


@pytest.mark.parametrize("fmt", ["bf16", "f16"])
@pytest.mark.parametrize("method", ["plain", "persistent"])
@pytest.mark.parametrize("dims,steps", [((9, 40, 130), 5), ((20, 70, 258), 9), ((4, 3, 2), 3)])
def test_j3d_stochastic_rounding_matches_oracle(gpu, fmt, method, dims, steps):
    torch = gpu
    say(f"[j3d16-sr] {fmt} {method} dims={dims} steps={steps}")
    init, got, _ = run16_sr(torch, dims, steps, fmt, method, seed=11, round_seed=0xC0FFEE)
    st = O.stencil_from_shape(P.shape_info("3d7pt"), "zero", coefs=CONFIG_3D)
    want = O.iterate16_sr(init, st, steps, fmt, 0xC0FFEE)
    if not np.array_equal(got, want):
        bad = np.argwhere(got != want)
        pytest.fail(f"{len(bad)} cells differ, first {bad[:4].tolist()}")


@pytest.mark.parametrize("fmt", ["bf16", "f16"])
def test_j3d_stochastic_rounding_plain_equals_persistent(gpu, fmt):
    torch = gpu
    dims, steps = (24, 50, 200), 11
    say(f"[j3d16-sr] plain vs persistent {fmt}")
    _, plain, _ = run16_sr(torch, dims, steps, fmt, "plain", seed=4, round_seed=5, lo=-1.0)
    for depth in (1, 3, 4):
        _, pers, _ = run16_sr(torch, dims, steps, fmt, "persistent", seed=4, round_seed=5, lo=-1.0, tune=[0, 0, depth])
        assert np.array_equal(pers, plain), f"depth {depth}"
    _, other, _ = run16_sr(torch, dims, steps, fmt, "plain", seed=4, round_seed=6, lo=-1.0)
    assert not np.array_equal(other, plain), "the seed changes the rounding"
