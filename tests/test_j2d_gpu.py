"""2D Jacobi on the GPU through the C ABI, bit for bit against the oracle:
the LDS-tile kernels (tile variants 0 and 1; the row-streaming default is
covered in test_j2d_stream_gpu.py, and the batch plan autotuner here may pick
any variant).

Every 2D shape of the reference (stars r1..r6, boxes r1 and r2) with its
reference weights -- including the FP32 box shapes whose centre weight the
reference writes as a double literal -- under both boundary rules, by the
per-step and the persistent method; FP16/BF16 storage against the oracle's
reduced-precision path; and batches of independent grids.
"""
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P

pytestmark = pytest.mark.gpu

SHAPES = ["2d5pt", "2ds9pt", "2d13pt", "2d17pt", "2d21pt", "2ds25pt", "2d9pt", "2d25pt"]
LDS = [0, 0, 0, 0, 3]    # opts.tune: LDS tile 128x32 (64x32 FP64)
LDS1 = [0, 0, 0, 0, 1]   # LDS tile 128x64 (64x64 FP64)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_equal(got, want, what):
    if not np.array_equal(got.view(np.uint8), want.view(np.uint8)):
        bad = np.argwhere(got != want)
        rms, mx = O.rms_error(got.astype(np.float64), want.astype(np.float64))
        pytest.fail(f"{what}: {len(bad)} cells differ, first {bad[:3].tolist()}, rms {rms:.3e} max {mx:.3e}")


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("method", ["plain", "persistent"])
@pytest.mark.parametrize("boundary", ["clamp", "zero"])
@pytest.mark.parametrize("tile", [LDS, LDS1])
def test_2d_shapes_match_oracle(gpu, shape, dtype, method, boundary, tile):
    dims, steps = (70, 150), 11
    say(f"[j2d] {shape} {np.dtype(dtype).name} {method} {boundary} dims={dims} steps={steps} tile={tile[4]}")
    si = P.shape_info(shape)
    field = O.libc_rand(dims[0] * dims[1], 3, dtype).reshape(dims)
    got, rep = P.jacobi2d_run(field, steps, shape=shape, method=method, boundary=boundary, tune=tile)
    want = O.iterate(field, O.stencil_from_shape(si, boundary), steps)
    check_equal(got, want, f"{shape} {method} {boundary}")


@pytest.mark.parametrize("dims,steps,depth", [((1, 1), 3, 0), ((5, 300), 9, 0), ((257, 33), 17, 0),
                                               ((200, 200), 20, 3), ((131, 262), 25, 8), ((64, 64), 1, 0)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_2d_sizes_and_pass_depths(gpu, dims, steps, depth, dtype):
    say(f"[j2d] sizes dims={dims} steps={steps} depth={depth} {np.dtype(dtype).name}")
    coefs = [0.2] * 5
    si = P.shape_info("2d5pt")
    field = O.uniform(dims[0] * dims[1], 4, 0.0, 1.0, dtype).reshape(dims)
    got, _ = P.jacobi2d_run(field, steps, shape="2d5pt", method="persistent", boundary="zero", coefs=coefs,
                            tune=[0, 0, depth, 0, 3])
    want = O.iterate(field, O.stencil_from_shape(si, "zero", coefs=coefs), steps)
    check_equal(got, want, f"dims {dims}")


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
@pytest.mark.parametrize("shape", ["2d9pt", "2d25pt"])
@pytest.mark.parametrize("method", ["plain", "persistent"])
def test_2d_reduced_precision_storage(gpu, fmt, shape, method):
    torch = gpu
    dims, steps = (90, 200), 13
    say(f"[j2d] storage {fmt} {shape} {method}")
    si = P.shape_info(shape)
    rng = np.random.default_rng(5)
    c = rng.random(len(si.coef))
    coefs = list(c / c.sum())
    n = dims[0] * dims[1]
    a = torch.empty(n, dtype=torch.int16, device="cuda")
    b = torch.empty(n, dtype=torch.int16, device="cuda")
    P.fill_uniform(a, n, 21, -1.0, 1.0, dtype=fmt)
    init = a.cpu().numpy().view(np.uint16).reshape(dims).copy()
    want_init = O.to16(O.uniform(n, 21, -1.0, 1.0, np.float32), fmt).reshape(dims)
    assert np.array_equal(init, want_init), "device 16-bit fill differs from the host sequence"
    rep = P.jacobi2d_device(a, b, dims[0], dims[1], steps, dtype=fmt, shape=shape, method=method, boundary="clamp",
                            coefs=coefs, tune=LDS)
    got = (b if rep.result_in_b else a).cpu().numpy().view(np.uint16).reshape(dims)
    want = O.iterate16(init, O.stencil_from_shape(si, "clamp", coefs=coefs), steps, fmt)
    check_equal(got, want, f"{fmt} {shape}")


@pytest.mark.parametrize("dtype,tdt", [("f32", "float32"), ("f64", "float64")])
def test_2d_batch_matches_per_grid_oracle(gpu, dtype, tdt):
    torch = gpu
    rng = np.random.default_rng(11)
    specs = []
    for i in range(12):
        wy, wx = int(rng.integers(20, 180)), int(rng.integers(20, 260))
        steps = int(rng.integers(1, 30))
        specs.append((wy, wx, steps))
    say(f"[j2d] batch {dtype} {specs}")
    si = P.shape_info("2d9pt")
    coefs = [0.1, 0.1, 0.1, 0.1, 0.2, 0.1, 0.1, 0.1, 0.1]
    bufs, fields = [], []
    for i, (wy, wx, st) in enumerate(specs):
        f = O.uniform(wy * wx, 100 + i, 0.0, 1.0, np.float64 if dtype == "f64" else np.float32).reshape(wy, wx)
        a = torch.from_numpy(f.copy()).cuda()
        b = torch.empty_like(a)
        bufs.append((a, b))
        fields.append(f)
    rep, flags = P.jacobi2d_batch([(a, b, wy, wx, st) for (a, b), (wy, wx, st) in zip(bufs, specs)], dtype=dtype,
                                  shape="2d9pt", method="persistent", boundary="zero", coefs=coefs, tune=LDS)
    for i, ((a, b), (wy, wx, st)) in enumerate(zip(bufs, specs)):
        got = (b if flags[i] else a).cpu().numpy()
        want = O.iterate(fields[i], O.stencil_from_shape(si, "zero", coefs=coefs), st)
        check_equal(got, want, f"grid {i} {wy}x{wx} steps {st}")


def test_2d_reference_signature(gpu):
    field = O.libc_rand(64 * 128, 1, np.float64).reshape(64, 128)
    out = np.empty_like(field)
    rc = P.jacobi_iterative(field, 64, 128, out, 256, 0, 5, False, False, False, -1, False)
    assert rc == 0
    want = O.iterate(field, O.stencil_from_shape(P.shape_info("2d5pt"), "clamp"), 5)
    check_equal(out, want, "jacobi_iterative")


GOLD = np.load(__import__("pathlib").Path(__file__).resolve().parent / "golden" / "ref_gold.npz")


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("method", ["plain", "persistent"])
def test_2d_matches_committed_reference_gold(gpu, shape, tag, method):
    field = GOLD[f"{shape}_{tag}_in"]
    steps = int(GOLD["steps_2d"])
    say(f"[j2d] reference gold vector {shape} {tag} {method} dims={field.shape} steps={steps}")
    got, _ = P.jacobi2d_run(field, steps, shape=shape, method=method, boundary="clamp", tune=LDS)
    check_equal(got, GOLD[f"{shape}_{tag}_out"], f"{shape} {tag} {method} vs reference gold")


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
@pytest.mark.parametrize("shape", ["2d5pt", "2d25pt"])
@pytest.mark.parametrize("method", ["plain", "persistent"])
def test_2d_stochastic_rounding_matches_oracle(gpu, fmt, shape, method):
    torch = gpu
    dims, steps = (90, 200), 13
    say(f"[j2d] stochastic rounding {fmt} {shape} {method}")
    si = P.shape_info(shape)
    rng = np.random.default_rng(8)
    c = rng.random(len(si.coef))
    coefs = list(c / c.sum())
    n = dims[0] * dims[1]
    a = torch.empty(n, dtype=torch.int16, device="cuda")
    b = torch.empty(n, dtype=torch.int16, device="cuda")
    P.fill_uniform(a, n, 22, -1.0, 1.0, dtype=fmt)
    init = a.cpu().numpy().view(np.uint16).reshape(dims).copy()
    rep = P.jacobi2d_device(a, b, dims[0], dims[1], steps, dtype=fmt, shape=shape, method=method, boundary="clamp",
                            coefs=coefs, rounding="stochastic", round_seed=77, tune=LDS)
    got = (b if rep.result_in_b else a).cpu().numpy().view(np.uint16).reshape(dims)
    want = O.iterate16_sr(init, O.stencil_from_shape(si, "clamp", coefs=coefs), steps, fmt, 77)
    check_equal(got, want, f"{fmt} {shape} {method} stochastic")


def _batch(torch, dtype, nb, seed):
    rng = np.random.default_rng(seed)
    specs = [(int(rng.integers(20, 180)), int(rng.integers(20, 260)), int(rng.integers(1, 30))) for _ in range(nb)]
    np_dt = np.float64 if dtype == "f64" else np.float32
    fields = [O.uniform(wy * wx, 300 + i, 0.0, 1.0, np_dt).reshape(wy, wx) for i, (wy, wx, _) in enumerate(specs)]
    bufs = [(torch.from_numpy(f.copy()).view(-1).cuda(),
             torch.empty(f.size, dtype=torch.from_numpy(f).dtype, device="cuda")) for f in fields]
    return specs, fields, bufs


@pytest.mark.parametrize("method,autotune", [("persistent", False), ("persistent", True), ("plain", False)])
def test_2d_batch_plan_graph_replays(gpu, method, autotune):
    """The captured launch replays several times with the program's own eager
    work (resetting the inputs) between replays; every replay matches the oracle."""
    torch = gpu
    dtype = "f32"
    specs, fields, bufs = _batch(torch, dtype, 10, 21)
    say(f"[j2d] batch plan {method} autotune={autotune} {specs}")
    si = P.shape_info("2d9pt")
    coefs = [0.1, 0.1, 0.1, 0.1, 0.2, 0.1, 0.1, 0.1, 0.1]
    plan = P.BatchPlan([(a, b, wy, wx, st) for (a, b), (wy, wx, st) in zip(bufs, specs)], dtype=dtype,
                       autotune=autotune, shape="2d9pt", method=method, boundary="zero", coefs=coefs)
    say(f"[j2d] plan config {plan.config}")
    assert plan.config["autotuned"] == int(autotune and method == "persistent")
    wants = [O.iterate(f, O.stencil_from_shape(si, "zero", coefs=coefs), st) for f, (_, _, st) in zip(fields, specs)]
    for replay in range(3):
        for (a, b), f in zip(bufs, fields):
            a.copy_(torch.from_numpy(f.copy()).view(-1).cuda())
            b.fill_(-1.0)
        torch.cuda.synchronize()
        rep = plan.run()
        assert rep.elapsed_ms > 0
        for i, ((a, b), (wy, wx, st)) in enumerate(zip(bufs, specs)):
            got = (b if plan.result_in_b[i] else a).cpu().numpy().reshape(wy, wx)
            check_equal(got, wants[i], f"replay {replay} grid {i} {wy}x{wx} steps {st}")
    plan.close()
