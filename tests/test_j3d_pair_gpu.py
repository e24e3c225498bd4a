"""The persistent 3D 7-point kernel with two columns per lane (tile variant
4: FP32, FP16 and BF16 storage, even grid widths), bit for bit against the
oracle and against the per-step kernel, at every pass depth, under both
boundary rules, with nearest and stochastic rounding; odd widths are refused."""
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P

pytestmark = pytest.mark.gpu
PAIR = 4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_equal(got, want, what):
    if not np.array_equal(got.view(np.uint8), want.view(np.uint8)):
        bad = np.argwhere(got != want)
        pytest.fail(f"{what}: {len(bad)} cells differ, first {bad[:3].tolist()}")


@pytest.mark.parametrize("dims,steps", [((19, 37, 70), 7), ((6, 5, 130), 9), ((33, 64, 128), 13),
                                        ((9, 70, 250), 5), ((3, 2, 2), 6)])
@pytest.mark.parametrize("boundary", ["clamp", "zero"])
@pytest.mark.parametrize("depth", [1, 2, 3, 4])
def test_pair_f32_matches_oracle(gpu, dims, steps, boundary, depth):
    say(f"[j3d-pair] f32 dims={dims} steps={steps} {boundary} depth={depth}")
    coefs = [0.11, 0.09, 0.1, 0.12, 0.08, 0.1, 0.4]
    field = O.uniform(int(np.prod(dims)), 6, 0.0, 1.0, np.float32).reshape(dims)
    got, rep = P.j3d_run(field, steps, method="persistent", boundary=boundary, coefs=coefs,
                         tune=[0, 0, depth, 0, PAIR])
    assert rep.reserved[0] == depth
    want = O.iterate(field, O.stencil_from_shape(P.shape_info("3d_star_r1"), boundary, coefs=coefs), steps)
    check_equal(got, want, f"dims {dims} {boundary} depth {depth}")


@pytest.mark.parametrize("fmt", ["bf16", "f16"])
@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
@pytest.mark.parametrize("boundary", ["clamp", "zero"])
def test_pair_16bit_matches_oracle_and_per_step(gpu, fmt, rounding, boundary):
    torch = gpu
    dims, steps = (21, 40, 132), 11
    say(f"[j3d-pair] {fmt} {rounding} {boundary}")
    n = int(np.prod(dims))
    coefs = [0.1] * 6 + [0.4]
    a = torch.empty(n, dtype=torch.int16, device="cuda")
    P.fill_uniform(a, n, 17, 0.0, 1.0, dtype=fmt)
    init = a.clone()
    host = init.cpu().numpy().view(np.uint16).reshape(dims).copy()

    def run(method, tune):
        x = init.clone()
        y = torch.empty_like(x)
        rep = P.j3d_device(x, y, *dims, steps, dtype=fmt, method=method, boundary=boundary, coefs=coefs,
                           rounding=rounding, round_seed=9, tune=tune)
        return (y if rep.result_in_b else x).cpu().numpy().view(np.uint16).reshape(dims)

    got = run("persistent", [0, 0, 4, 0, PAIR])
    st = O.stencil_from_shape(P.shape_info("3d_star_r1"), boundary, coefs=coefs)
    want = O.iterate16_sr(host, st, steps, fmt, 9) if rounding == "stochastic" else O.iterate16(host, st, steps, fmt)
    check_equal(got, want, f"{fmt} {rounding} {boundary} vs oracle")
    check_equal(got, run("plain", None), f"{fmt} {rounding} {boundary} vs per-step")


def test_pair_needs_an_even_width(gpu):
    field = O.uniform(4 * 6 * 9, 3, 0.0, 1.0, np.float32).reshape(4, 6, 9)
    out = np.empty_like(field)
    rc = P.j3d_iterative(field, 4, 6, 9, out, iteration=3, tune=[0, 0, 0, 0, PAIR])
    assert rc == -13  # PERKS_ERR_UNSUPPORTED
