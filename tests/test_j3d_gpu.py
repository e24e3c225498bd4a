"""3D Jacobi on the GPU through the C ABI, compared bit for bit with the oracle.

The kernels are built with -ffp-contract=off and sum in the reference gold's
order, so FP64 and FP32 results must be identical to the CPU oracle (which is
itself pinned to the reference gold in test_oracle.py).
"""
import sys

import ctypes as C

import numpy as np
import pytest

import oracle as O
import perks_amd as P

pytestmark = pytest.mark.gpu

CONFIG_3D = [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_case(method, dims, steps, dtype, boundary, coefs=None, seed=1, **kw):
    say(f"[j3d] method={method} dims={dims} steps={steps} dtype={np.dtype(dtype).name} bc={boundary} coefs={'cfg' if coefs else 'ref'} {kw}")
    si = P.shape_info("3d_star_r1")
    n = int(np.prod(dims))
    field = (O.libc_rand(n, seed, dtype) if coefs is None else O.uniform(n, seed, 0.0, 1.0, dtype)).reshape(dims)
    got, rep = P.j3d_run(field, steps, method=method, boundary=boundary, coefs=coefs, **kw)
    want = O.iterate(field, O.stencil_from_shape(si, boundary, coefs=coefs), steps)
    if not np.array_equal(got.view(np.uint8), want.view(np.uint8)):
        rms, mx = O.rms_error(got, want)
        bad = np.argwhere(got != want)
        pytest.fail(f"mismatch: {len(bad)} cells differ, first {bad[:3].tolist()}, rms {rms:.3e} max {mx:.3e}")
    return rep


PLAIN_CASES = [
    ((8, 32, 64), 1),
    ((16, 64, 128), 3),
    ((33, 40, 70), 4),     # ragged x (70), ragged y (40), odd z
    ((5, 7, 9), 2),        # smaller than one tile
    ((64, 96, 192), 5),
    ((1, 33, 65), 2),      # single plane
]


@pytest.mark.parametrize("dims,steps", PLAIN_CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("boundary", ["clamp", "zero"])
def test_plain_matches_oracle(gpu, dims, steps, dtype, boundary):
    run_case("plain", dims, steps, dtype, boundary)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_plain_config_coefs(gpu, dtype):
    run_case("plain", (40, 48, 96), 6, dtype, "zero", coefs=CONFIG_3D)


def test_plain_zero_steps_is_copy(gpu):
    field = O.uniform(4 * 8 * 64, 3, 0, 1).reshape(4, 8, 64)
    got, rep = P.j3d_run(field, 0, method="plain")
    assert np.array_equal(got, field)


def test_fill_uniform_matches_host(gpu):
    torch = gpu
    n = 1_000_003
    for dt, tdt, name in ((np.float64, torch.float64, "f64"), (np.float32, torch.float32, "f32")):
        d = torch.empty(n, dtype=tdt, device="cuda")
        P.fill_uniform(d, n, 1234, -1.0, 1.0, dtype=name)
        torch.cuda.synchronize()
        want = O.uniform(n, 1234, -1.0, 1.0, dt)
        assert np.array_equal(d.cpu().numpy(), want)


def test_device_api_parity_and_result_buffer(gpu):
    torch = gpu
    dims = (24, 64, 128)
    n = int(np.prod(dims))
    si = P.shape_info("3d_star_r1")
    for steps in (1, 2, 3):
        a = torch.empty(n, dtype=torch.float64, device="cuda")
        b = torch.empty(n, dtype=torch.float64, device="cuda")
        P.fill_uniform(a, n, 9, 0.0, 1.0)
        rep = P.j3d_device(a, b, *dims, steps, method="plain", boundary="zero", coefs=CONFIG_3D)
        res = (b if rep.result_in_b else a).cpu().numpy().reshape(dims)
        assert rep.result_in_b == steps % 2
        want = O.iterate(O.uniform(n, 9, 0.0, 1.0).reshape(dims), O.stencil_from_shape(si, "zero", coefs=CONFIG_3D),
                         steps)
        assert np.array_equal(res, want)


PERSIST_CASES = [
    ((16, 64, 128), 1, 0),
    ((16, 64, 128), 2, 0),
    ((33, 40, 70), 5, 0),      # ragged in every axis, steps not a multiple of the pass depth
    ((5, 7, 9), 4, 0),         # smaller than one tile
    ((64, 96, 192), 7, 0),
    ((1, 33, 65), 3, 0),       # single plane
    ((40, 70, 130), 6, 1),     # one step per pass
    ((40, 70, 130), 6, 2),     # two steps per pass
    ((40, 70, 130), 7, 3),     # three steps per pass
    ((96, 100, 120), 9, 3),
]


@pytest.mark.parametrize("dims,steps,depth", PERSIST_CASES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("boundary", ["clamp", "zero"])
def test_persistent_matches_oracle(gpu, dims, steps, depth, dtype, boundary):
    tune = [0, 0, depth] if depth else None
    run_case("persistent", dims, steps, dtype, boundary, tune=tune)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_persistent_config_coefs(gpu, dtype):
    run_case("persistent", (48, 80, 160), 10, dtype, "zero", coefs=CONFIG_3D)


def test_persistent_result_buffer_parity(gpu):
    torch = gpu
    dims = (20, 40, 100)
    n = int(np.prod(dims))
    si = P.shape_info("3d_star_r1")
    for steps in (1, 2, 3, 4, 6):
        a = torch.empty(n, dtype=torch.float64, device="cuda")
        b = torch.empty(n, dtype=torch.float64, device="cuda")
        P.fill_uniform(a, n, 5, 0.0, 1.0)
        rep = P.j3d_device(a, b, *dims, steps, method="persistent", boundary="clamp")
        res = (b if rep.result_in_b else a).cpu().numpy().reshape(dims)
        want = O.iterate(O.uniform(n, 5, 0.0, 1.0).reshape(dims), O.stencil_from_shape(si, "clamp"), steps)
        assert np.array_equal(res, want), steps


GOLD = np.load(__import__("pathlib").Path(__file__).resolve().parent / "golden" / "ref_gold.npz")


@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("method", ["plain", "persistent"])
def test_matches_committed_reference_gold(gpu, tag, method):
    field = GOLD[f"3d7pt_{tag}_in"]
    steps = int(GOLD["steps_3d"])
    say(f"[j3d] reference gold vector 3d7pt {tag} {method} dims={field.shape} steps={steps}")
    got, _ = P.j3d_run(field, steps, method=method, boundary="clamp")
    assert np.array_equal(got.view(np.uint8), GOLD[f"3d7pt_{tag}_out"].view(np.uint8))


@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("method", ["plain", "persistent"])
def test_matches_committed_interior_gold(gpu, tag, method):
    """The reference's interior-only j3d_gold: zero-halo form on the interior."""
    f = GOLD[f"j3d_gold4_{tag}_in"]
    say(f"[j3d] interior gold 3d7pt {tag} {method}")
    got, _ = P.j3d_run(np.ascontiguousarray(f[1:-1, 1:-1, 1:-1]), 4, method=method, boundary="zero")
    assert np.array_equal(got, GOLD[f"j3d_gold4_{tag}_out"][1:-1, 1:-1, 1:-1])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_min_height_is_the_resident_capacity(gpu, dtype):
    """getminHeight / getMinWidthY: the largest height whose domain the
    resident-tile kernel holds (blocks of 62 columns x 8 rows x 6 (FP64) or
    12 (FP32) planes, one per wave); it runs there and refuses one more layer."""
    wy, wx = 512, 512
    h = P.j3d_iterative(np.empty(0, dtype), 0, wy, wx, None, getminHeight=True)
    bz = 6 if dtype == np.float64 else 12
    caps = P.query_device()
    assert h > 0 and h % bz == 0
    assert h // bz * -(-wx // 62) * -(-wy // 8) <= caps.cu_count * 8
    lib = P.load()
    dt = 0 if dtype == np.float64 else 1
    assert lib.perks_get_min_width_y(dt, wx, wy, 256, 0) == h
    assert lib.perks_get_min_width_y(dt, 8192, 8192, 256, 0) == -4  # not one layer of blocks fits
    # the 27-point box keeps the on-chip-capacity figure
    o = P.make_opts(shape="3d27pt")
    onchip = caps.cu_count * (caps.lds_per_cu + caps.vgprs_per_simd * caps.simds_per_cu * caps.warp_size * 4)
    h27 = lib.perks_j3d_min_height(dt, wy, wx, C.byref(o))
    assert h27 * wy * wx * np.dtype(dtype).itemsize <= onchip < (h27 + 1) * wy * wx * np.dtype(dtype).itemsize
