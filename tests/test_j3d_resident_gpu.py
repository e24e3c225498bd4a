"""Resident-tile form of the persistent 7-point method (opts.tune[4] = 5).

The whole domain stays in VGPRs for the whole solve (one wave per block of
62 columns x 8 rows x 6 (FP64) or 12 (FP32) planes); only the blocks' shells
go through memory, one grid barrier per step.  Results must equal the CPU
oracle (and so the reference gold) bit for bit: domains whose blocks overhang
every face, single cells, both boundaries, with and without a divisor, odd
and even step counts; at the largest resident height of a 512 x 512 domain,
the per-step kernel's bits.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P
from perks_amd import _lib as L

pytestmark = pytest.mark.gpu

CONFIG_3D = [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4]
RESIDENT = [0, 0, 0, 0, 5]
DIMS = [(1, 1, 1), (2, 3, 5), (6, 8, 62), (7, 9, 63), (12, 16, 124), (13, 21, 130), (25, 40, 200)]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("boundary", ["zero", "clamp"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_resident_matches_oracle(gpu, dtype, boundary):
    st = O.stencil_from_shape(P.shape_info("3d_star_r1"), boundary, coefs=CONFIG_3D)
    for dims in DIMS:
        for steps in (1, 2, 7):
            say(f"[resident] {np.dtype(dtype).name} {boundary} dims={dims} steps={steps}")
            field = O.uniform(int(np.prod(dims)), dims[2] + steps, 0.0, 1.0, dtype=dtype).reshape(dims)
            got, rep = P.j3d_run(field, steps, method="persistent", boundary=boundary, coefs=CONFIG_3D,
                                 tune=RESIDENT)
            assert rep.reserved[0] == 1 and rep.result_in_b == steps % 2
            want = O.iterate(field, st, steps)
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), f"dims={dims} steps={steps}"


def test_resident_reference_weights_and_divisor(gpu):
    """The reference's own 7-point weights under the clamp rule, and a divisor."""
    si = P.shape_info("3d_star_r1")
    field = O.uniform(11 * 19 * 70, 5, 0.0, 1.0).reshape(11, 19, 70)
    got, _ = P.j3d_run(field, 5, method="persistent", boundary="clamp", tune=RESIDENT)
    assert np.array_equal(got, O.iterate(field, O.stencil_from_shape(si, "clamp"), 5))
    coefs = [1.0, 2.0, 1.0, 2.0, 1.0, 2.0, 3.0]
    got, _ = P.j3d_run(field, 4, method="persistent", boundary="zero", coefs=coefs, divisor=12.0, tune=RESIDENT)
    want = O.iterate(field, O.stencil_from_shape(si, "zero", coefs=coefs, divisor=12.0), 4)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_resident_at_its_capacity_equals_per_step(gpu, dtype):
    """At the height perks_j3d_min_height reports for 512 x 512 the resident
    form runs and equals the per-step kernel after 100 steps; one more layer
    of blocks is refused."""
    torch = gpu
    lib = P.load()
    dt = 0 if dtype == "f64" else 1
    o = P.make_opts(boundary="zero", coefs=CONFIG_3D)
    ny = nx = 512
    h = lib.perks_j3d_min_height(dt, ny, nx, C.byref(o))
    assert h > 0
    tdt = torch.float64 if dtype == "f64" else torch.float32
    init = torch.empty(h * ny * nx, dtype=tdt, device="cuda")
    P.fill_uniform(init, init.numel(), 4, 0.0, 1.0, dtype=dtype)
    outs = []
    for method, tune in (("plain", None), ("persistent", RESIDENT), ("persistent", None)):
        say(f"[resident] {dtype} {h}x{ny}x{nx} {method} {tune}")
        a = init.clone()
        b = torch.empty_like(a)
        rep = P.j3d_device(a, b, h, ny, nx, 100, dtype=dtype, method=method, boundary="zero", coefs=CONFIG_3D,
                           tune=tune)
        torch.cuda.synchronize()
        outs.append(b if rep.result_in_b else a)
    assert torch.equal(outs[1], outs[0]) and torch.equal(outs[2], outs[0])
    bz = 6 if dtype == "f64" else 12
    big = torch.empty((h + bz) * ny * nx, dtype=tdt, device="cuda")
    with pytest.raises(L.PerksError) as e:
        P.j3d_device(big, torch.empty_like(big), h + bz, ny, nx, 2, dtype=dtype, method="persistent",
                     boundary="zero", coefs=CONFIG_3D, tune=RESIDENT)
    assert e.value.code == L.ERR_UNSUPPORTED_SIZE


def test_resident_refusals(gpu):
    torch = gpu
    a = torch.zeros(4 * 8 * 8, dtype=torch.int16, device="cuda")
    with pytest.raises(L.PerksError) as e:  # 16-bit storage
        P.j3d_device(a, torch.empty_like(a), 4, 8, 8, 2, dtype="bf16", method="persistent", boundary="zero",
                     tune=RESIDENT)
    assert e.value.code == L.ERR_UNSUPPORTED
    f = np.zeros((4, 8, 8))
    with pytest.raises(L.PerksError) as e:  # the contracted update is built for the ring kernel only
        P.j3d_run(f, 2, method="persistent", boundary="zero", coefs=CONFIG_3D, tune=[0, 0, 0, 0, 5, 0, 1])
    assert e.value.code == L.ERR_UNSUPPORTED
