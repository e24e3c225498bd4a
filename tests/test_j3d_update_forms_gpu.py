"""Update forms of the persistent 7-point kernel.

With six equal positive neighbour weights the kernel forms each product k*x
once per cell and shares it between the terms that read x; otherwise it
multiplies per term as the gold does.  Both must give the gold's bits, so
each weight set below runs every tile variant at two pass depths under both
boundaries and is compared with the CPU oracle bit for bit -- FP64 and FP32
storage, and FP16/BF16 storage (FP32 arithmetic) with both roundings; the
shared-product kernels must not spill.
"""
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P
from perks_amd import _lib as L

pytestmark = pytest.mark.gpu

WEIGHTS = {
    "equal_positive": ([0.1] * 6 + [0.4], None),          # shared products
    "equal_positive_div": ([1.0] * 6 + [2.0], 8.0),       # shared products, then the division
    "equal_negative": ([-0.1] * 6 + [1.6], None),         # k < 0: the gold form
    "one_differs": ([0.1] * 5 + [0.1000001, 0.4], None),  # the gold form
}


RESIDENT = 5  # opts.tune[4]: the resident-tile variant


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("weights", sorted(WEIGHTS))
@pytest.mark.parametrize("boundary", ["zero", "clamp"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_forms_match_oracle(gpu, dtype, boundary, weights):
    coefs, div = WEIGHTS[weights]
    st = O.stencil_from_shape(P.shape_info("3d_star_r1"), boundary, coefs=coefs, divisor=div)
    dims = (13, 70, 130)
    field = O.uniform(int(np.prod(dims)), 21, 0.0, 1.0, dtype=dtype).reshape(dims)
    want = {}
    for variant in (0, 1, 2, 3, 4):
        for depth in (3, 4):
            steps = 9
            say(f"[forms] {np.dtype(dtype).name} {boundary} {weights} variant={variant} depth={depth}")
            got, _ = P.j3d_run(field, steps, method="persistent", boundary=boundary, coefs=coefs, divisor=div,
                               tune=[0, 0, depth, 0, variant])
            if steps not in want:
                want[steps] = O.iterate(field, st, steps)
            assert np.array_equal(got.view(np.uint8), want[steps].view(np.uint8)), f"variant={variant} depth={depth}"


@pytest.mark.parametrize("weights", ["equal_positive", "one_differs"])
@pytest.mark.parametrize("boundary", ["zero", "clamp"])
@pytest.mark.parametrize("fmt", ["bf16", "f16"])
def test_forms_16bit_storage_match_oracle(gpu, fmt, boundary, weights):
    torch = gpu
    coefs, div = WEIGHTS[weights]
    st = O.stencil_from_shape(P.shape_info("3d_star_r1"), boundary, coefs=coefs, divisor=div)
    dims, steps = (11, 70, 130), 9
    n = int(np.prod(dims))
    a0 = torch.empty(n, dtype=torch.int16, device="cuda")
    P.fill_uniform(a0, n, 5, 0.0, 1.0, dtype=fmt)
    init = a0.cpu().numpy().view(np.uint16).reshape(dims).copy()
    want = {"nearest": O.iterate16(init, st, steps, fmt), "stochastic": O.iterate16_sr(init, st, steps, fmt, 33)}
    for rounding in ("nearest", "stochastic"):
        for variant in (0, 1, 2, 3, 4):
            for depth in (3, 4):
                say(f"[forms] {fmt} {boundary} {weights} {rounding} variant={variant} depth={depth}")
                a = a0.clone()
                b = torch.empty_like(a)
                rep = P.j3d_device(a, b, *dims, steps, dtype=fmt, method="persistent", boundary=boundary, coefs=coefs,
                                   divisor=div, rounding=rounding, round_seed=33, tune=[0, 0, depth, 0, variant])
                got = (b if rep.result_in_b else a).cpu().numpy().view(np.uint16).reshape(dims)
                assert np.array_equal(got, want[rounding]), f"{rounding} variant={variant} depth={depth}"


@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16", "f16"])
def test_shared_product_kernels_do_not_spill(gpu, dtype):
    for rounding in (("nearest", "stochastic") if dtype in ("bf16", "f16") else ("nearest",)):
        rows = P.j3d_kernel_table(64, 64, 64, dtype=dtype, shape="3d7pt", boundary="zero", rounding=rounding,
                                  coefs=WEIGHTS["equal_positive"][0])
        for r in rows:
            if dtype in ("bf16", "f16") and r["method"] == "persistent" and r["variant"] == RESIDENT:
                assert r["status"] == L.ERR_UNSUPPORTED, r  # the resident-tile kernel: FP64/FP32 storage
                continue
            assert r["spill_bytes"] == 0 and r["status"] == L.PERKS_OK, r
