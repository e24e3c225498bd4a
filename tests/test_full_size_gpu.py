"""Parity at the benchmark sizes, through properties that do not need the
CPU oracle at that size: every kernel and method that computes the same
thing must give the same bits.

* 3D 7-point at 512^3: the persistent kernel (every tile variant, 4 steps
  per pass) against the per-step kernel, FP64 and FP32, both boundaries.
* 3D 27- and 13-point at 512^3: register streaming (per step and persistent)
  against the LDS-plane kernel.
* 2D boxes at 8192^2: row streaming (per step and persistent) against the
  LDS-tile kernel, FP32 and BF16 with stochastic rounding.
* 512 uneven grids: one captured plan against each grid solved alone.

The small-size tests pin each kernel to the oracle bit for bit; these carry
that parity to the full sizes.  The 2D 5-point FP64 case at 2048^2 x 100
steps (the size the reference runs its CPU gold at) is checked against the
CPU oracle directly.
"""
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run3(torch, init, n, steps, dtype, **kw):
    a = init.clone()
    b = torch.empty_like(a)
    rep = P.j3d_device(a, b, n, n, n, steps, dtype=dtype, **kw)
    torch.cuda.synchronize()
    return (b if rep.result_in_b else a), rep


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("boundary", ["zero", "clamp"])
def test_3d7pt_512_persistent_variants_equal_per_step(gpu, dtype, boundary):
    torch = gpu
    n, steps = 512, 9
    say(f"[full] 3d7pt {n}^3 {dtype} {boundary} steps={steps}")
    tdt = torch.float64 if dtype == "f64" else torch.float32
    init = torch.empty(n ** 3, dtype=tdt, device="cuda")
    P.fill_uniform(init, n ** 3, 1, 0.0, 1.0, dtype=dtype)
    ref, _ = run3(torch, init, n, steps, dtype, method="plain", boundary=boundary)
    for var in range(5):
        got, rep = run3(torch, init, n, steps, dtype, method="persistent", boundary=boundary,
                        tune=[0, 0, 4, 0, var])
        assert rep.reserved[0] == 4
        assert torch.equal(got, ref), f"variant {var} differs from the per-step kernel"
        del got


@pytest.mark.parametrize("shape", ["3d27pt", "3d13pt"])
def test_3d_shapes_512_streaming_equals_lds(gpu, shape):
    torch = gpu
    n, steps = 512, 4
    say(f"[full] {shape} {n}^3 f64 steps={steps}")
    init = torch.empty(n ** 3, dtype=torch.float64, device="cuda")
    P.fill_uniform(init, n ** 3, 2, 0.0, 1.0, dtype="f64")
    ref, _ = run3(torch, init, n, steps, "f64", shape=shape, method="plain", boundary="clamp", tune=[0, 0, 0, 0, 3])
    for method in ("plain", "persistent"):
        got, _ = run3(torch, init, n, steps, "f64", shape=shape, method=method, boundary="clamp",
                      tune=[0, 0, 0, 0, 2])
        assert torch.equal(got, ref), f"{shape} streaming {method} differs from the LDS kernel"
        del got


@pytest.mark.parametrize("shape", ["2d9pt", "2d25pt"])
@pytest.mark.parametrize("dtype,rounding", [("f32", "nearest"), ("bf16", "stochastic")])
def test_2d_8192_streaming_equals_lds(gpu, shape, dtype, rounding):
    torch = gpu
    n, steps = 8192, 12
    say(f"[full] {shape} {n}^2 {dtype} {rounding} steps={steps}")
    si = P.shape_info(shape)
    rng = np.random.default_rng(1)
    c = rng.random(len(si.coef))
    coefs = list(c / c.sum())
    tdt = torch.float32 if dtype == "f32" else torch.int16
    init = torch.empty(n * n, dtype=tdt, device="cuda")
    P.fill_uniform(init, n * n, 1, -1.0, 1.0, dtype=dtype)

    def run(method, tile):
        a = init.clone()
        b = torch.empty_like(a)
        rep = P.jacobi2d_device(a, b, n, n, steps, dtype=dtype, shape=shape, method=method, boundary="clamp",
                                coefs=coefs, rounding=rounding, round_seed=5, tune=[0, 0, 0, 0, tile])
        torch.cuda.synchronize()
        return b if rep.result_in_b else a

    ref = run("plain", 3)
    for method in ("plain", "persistent"):
        assert torch.equal(run(method, 2), ref), f"{shape} streaming {method} differs from the LDS kernel"


def test_batch_512_grids_plan_equals_each_grid_alone(gpu):
    torch = gpu
    rng = np.random.default_rng(1)
    specs = []
    for _ in range(512):
        wy, wx = (int(round(float(np.exp(rng.uniform(np.log(64), np.log(1024)))))) for _ in range(2))
        specs.append((wy, wx, int(rng.integers(5, 40))))
    say(f"[full] batch of {len(specs)} grids, {sum(wy * wx for wy, wx, _ in specs)} cells")
    coefs = [0.2] * 5
    bufs = []
    for i, (wy, wx, _) in enumerate(specs):
        a = torch.empty(wy * wx, dtype=torch.float32, device="cuda")
        P.fill_uniform(a, wy * wx, 100 + i, 0.0, 1.0, dtype="f32")
        bufs.append((a, torch.empty_like(a)))
    inits = [a.clone() for a, _ in bufs]
    plan = P.BatchPlan([(a, b, wy, wx, st) for (a, b), (wy, wx, st) in zip(bufs, specs)], dtype="f32",
                       autotune=True, shape="2d5pt", method="persistent", boundary="zero", coefs=coefs)
    say(f"[full] plan config {plan.config}")
    plan.run()
    torch.cuda.synchronize()
    for i, ((a, b), (wy, wx, st)) in enumerate(zip(bufs, specs)):
        got = b if plan.result_in_b[i] else a
        x = inits[i].clone()
        y = torch.empty_like(x)
        rep = P.jacobi2d_device(x, y, wy, wx, st, dtype="f32", shape="2d5pt", method="plain", boundary="zero",
                                coefs=coefs, tune=[0, 0, 0, 0, 3])
        want = y if rep.result_in_b else x
        assert torch.equal(got, want), f"grid {i} {wy}x{wx} steps {st} differs from its own per-step solve"
    plan.close()


@pytest.mark.parametrize("method", ["plain", "persistent"])
@pytest.mark.parametrize("tile", [2, 3])
def test_2d5pt_2048_fp64_100_steps_matches_oracle(gpu, method, tile):
    n, steps = 2048, 100
    say(f"[full] 2d5pt {n}^2 f64 {steps} steps {method} tile={tile}")
    coefs = [0.2] * 5
    field = O.uniform(n * n, 1, 0.0, 1.0).reshape(n, n)
    got, _ = P.jacobi2d_run(field, steps, shape="2d5pt", method=method, boundary="zero", coefs=coefs,
                            tune=[0, 0, 0, 0, tile])
    want = O.iterate(field, O.stencil_from_shape(P.shape_info("2d5pt"), "zero", coefs=coefs), steps, threads=4)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))
