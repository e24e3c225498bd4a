"""Parity at the benchmark configurations' full scope (sizes and step counts),
through equalities between independent kernels and methods; each of them is
pinned to the CPU oracle bit for bit at small sizes by the other tests.

* 3D 7-point 512^3, 1000 time steps (the headline configuration), FP64 and
  FP32: the persistent kernel (4 steps per pass, 250 grid barriers) against
  the per-step kernel (1000 launches).
* 2D box 8192^2 with FP16 storage, 500 steps, 2d9pt and 2d25pt: row
  streaming per step and persistent against the LDS-tile kernel per step.
* 512 uneven grids with 50-500 steps each, 2d5pt and 2d9pt: one autotuned,
  graph-captured persistent plan against each grid's own per-step solve.
"""
import sys

import numpy as np
import pytest

import perks_amd as P

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_3d7pt_512_1000_steps_persistent_equals_per_step(gpu, dtype):
    torch = gpu
    n, steps = 512, 1000
    say(f"[scope] 3d7pt {n}^3 {dtype} {steps} steps")
    tdt = torch.float64 if dtype == "f64" else torch.float32
    init = torch.empty(n ** 3, dtype=tdt, device="cuda")
    P.fill_uniform(init, n ** 3, 1, 0.0, 1.0, dtype=dtype)
    out = {}
    for method in ("persistent", "plain"):
        a = init.clone()
        b = torch.empty_like(a)
        rep = P.j3d_device(a, b, n, n, n, steps, dtype=dtype, method=method, boundary="zero",
                           coefs=[0.1] * 6 + [0.4])
        torch.cuda.synchronize()
        out[method] = b if rep.result_in_b else a
        say(f"[scope]   {method}: {rep.elapsed_ms:.1f} ms, {rep.gcells_per_s:.0f} GCells/s")
        del a, b
    assert torch.equal(out["persistent"], out["plain"])


@pytest.mark.parametrize("shape", ["2d9pt", "2d25pt"])
def test_2d_8192_fp16_500_steps_streaming_equals_lds(gpu, shape):
    torch = gpu
    n, steps = 8192, 500
    say(f"[scope] {shape} {n}^2 f16 {steps} steps")
    si = P.shape_info(shape)
    c = np.random.default_rng(1).random(len(si.coef))
    coefs = list(c / c.sum())
    init = torch.empty(n * n, dtype=torch.int16, device="cuda")
    P.fill_uniform(init, n * n, 1, -1.0, 1.0, dtype="f16")

    def run(method, tile):
        a = init.clone()
        b = torch.empty_like(a)
        rep = P.jacobi2d_device(a, b, n, n, steps, dtype="f16", shape=shape, method=method, boundary="clamp",
                                coefs=coefs, tune=[0, 0, 0, 0, tile])
        torch.cuda.synchronize()
        say(f"[scope]   {method} tile {tile}: {rep.elapsed_ms:.1f} ms")
        return b if rep.result_in_b else a

    ref = run("plain", 3)
    for method, tile in (("plain", 2), ("persistent", 2), ("persistent", 4)):
        assert torch.equal(run(method, tile), ref), f"{shape} {method} tile {tile} differs from the LDS kernel"


@pytest.mark.parametrize("shape", ["2d5pt", "2d9pt"])
def test_batch_512_grids_50_to_500_steps(gpu, shape):
    torch = gpu
    rng = np.random.default_rng(7)
    specs = []
    for _ in range(512):
        wy, wx = (int(round(float(np.exp(rng.uniform(np.log(64), np.log(1024)))))) for _ in range(2))
        specs.append((wy, wx, int(rng.integers(50, 501))))
    say(f"[scope] batch {shape}: {len(specs)} grids, {sum(wy * wx * st for wy, wx, st in specs)} cell-steps")
    si = P.shape_info(shape)
    coefs = [1.0 / len(si.coef)] * len(si.coef)
    bufs = []
    for i, (wy, wx, _) in enumerate(specs):
        a = torch.empty(wy * wx, dtype=torch.float32, device="cuda")
        P.fill_uniform(a, wy * wx, 300 + i, 0.0, 1.0, dtype="f32")
        bufs.append((a, torch.empty_like(a)))
    inits = [a.clone() for a, _ in bufs]
    plan = P.BatchPlan([(a, b, wy, wx, st) for (a, b), (wy, wx, st) in zip(bufs, specs)], dtype="f32",
                       autotune=True, shape=shape, method="persistent", boundary="zero", coefs=coefs)
    say(f"[scope]   plan config {plan.config}")
    rep = plan.run()
    torch.cuda.synchronize()
    say(f"[scope]   batch {rep.elapsed_ms:.1f} ms, {rep.gcells_per_s:.0f} GCells/s")
    for i, ((a, b), (wy, wx, st)) in enumerate(zip(bufs, specs)):
        got = b if plan.result_in_b[i] else a
        x = inits[i].clone()
        y = torch.empty_like(x)
        r = P.jacobi2d_device(x, y, wy, wx, st, dtype="f32", shape=shape, method="plain", boundary="zero",
                              coefs=coefs)
        want = y if r.result_in_b else x
        assert torch.equal(got, want), f"grid {i} {wy}x{wx} steps {st} differs from its own per-step solve"
    plan.close()
