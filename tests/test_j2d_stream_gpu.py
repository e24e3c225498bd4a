"""The row-streaming 2D kernel through the C ABI, bit for bit against the
oracle, with one column per lane (tile variant 2) and with two (variant 4):
every 2D shape under both boundary rules by both methods, sizes that leave
partial strips and segments (including grids narrower than a strip and a
single cell), every pass depth it is built for, FP16/BF16 storage with
nearest and stochastic rounding, the reference's gold vectors, and batches
(direct and through a captured plan).
"""
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P

pytestmark = pytest.mark.gpu

SHAPES = ["2d5pt", "2ds9pt", "2d13pt", "2d17pt", "2d21pt", "2ds25pt", "2d9pt", "2d25pt"]
STREAM = 2  # tile variant of the streaming kernel (opts.tune[4])


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def tune(depth=0):
    return [0, 0, depth, 0, STREAM]


def check_equal(got, want, what):
    if not np.array_equal(got.view(np.uint8), want.view(np.uint8)):
        bad = np.argwhere(got != want)
        rms, mx = O.rms_error(got.astype(np.float64), want.astype(np.float64))
        pytest.fail(f"{what}: {len(bad)} cells differ, first {bad[:3].tolist()}, rms {rms:.3e} max {mx:.3e}")


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("method", ["plain", "persistent"])
@pytest.mark.parametrize("boundary", ["clamp", "zero"])
def test_stream_shapes_match_oracle(gpu, shape, dtype, method, boundary):
    dims, steps = (70, 150), 11
    say(f"[j2s] {shape} {np.dtype(dtype).name} {method} {boundary} dims={dims} steps={steps}")
    si = P.shape_info(shape)
    field = O.libc_rand(dims[0] * dims[1], 3, dtype).reshape(dims)
    got, rep = P.jacobi2d_run(field, steps, shape=shape, method=method, boundary=boundary, tune=tune())
    assert rep.reserved[2] == STREAM
    want = O.iterate(field, O.stencil_from_shape(si, boundary), steps)
    check_equal(got, want, f"{shape} {method} {boundary}")


@pytest.mark.parametrize("dims,steps,depth", [((1, 1), 3, 0), ((5, 300), 9, 0), ((257, 33), 17, 0),
                                               ((200, 200), 20, 3), ((131, 262), 25, 4), ((64, 64), 1, 0),
                                               ((3, 70), 7, 2), ((300, 2), 6, 4), ((97, 56), 8, 4),
                                               ((97, 57), 9, 4)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("boundary", ["clamp", "zero"])
def test_stream_sizes_and_pass_depths(gpu, dims, steps, depth, dtype, boundary):
    say(f"[j2s] sizes dims={dims} steps={steps} depth={depth} {np.dtype(dtype).name} {boundary}")
    for shape in ("2d9pt", "2d5pt"):
        si = P.shape_info(shape)
        field = O.uniform(dims[0] * dims[1], 4, 0.0, 1.0, dtype).reshape(dims)
        got, _ = P.jacobi2d_run(field, steps, shape=shape, method="persistent", boundary=boundary,
                                tune=tune(depth))
        want = O.iterate(field, O.stencil_from_shape(si, boundary), steps)
        check_equal(got, want, f"{shape} dims {dims} depth {depth}")


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
@pytest.mark.parametrize("shape", ["2d5pt", "2d9pt", "2d25pt"])
@pytest.mark.parametrize("method", ["plain", "persistent"])
@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
def test_stream_16bit_storage(gpu, fmt, shape, method, rounding):
    torch = gpu
    dims, steps = (90, 200), 13
    say(f"[j2s] storage {fmt} {shape} {method} {rounding}")
    si = P.shape_info(shape)
    rng = np.random.default_rng(5)
    c = rng.random(len(si.coef))
    coefs = list(c / c.sum())
    n = dims[0] * dims[1]
    a = torch.empty(n, dtype=torch.int16, device="cuda")
    b = torch.empty(n, dtype=torch.int16, device="cuda")
    P.fill_uniform(a, n, 21, -1.0, 1.0, dtype=fmt)
    init = a.cpu().numpy().view(np.uint16).reshape(dims).copy()
    rep = P.jacobi2d_device(a, b, dims[0], dims[1], steps, dtype=fmt, shape=shape, method=method, boundary="clamp",
                            coefs=coefs, rounding=rounding, round_seed=77, tune=tune())
    got = (b if rep.result_in_b else a).cpu().numpy().view(np.uint16).reshape(dims)
    st = O.stencil_from_shape(si, "clamp", coefs=coefs)
    want = (O.iterate16_sr(init, st, steps, fmt, 77) if rounding == "stochastic"
            else O.iterate16(init, st, steps, fmt))
    check_equal(got, want, f"{fmt} {shape} {method} {rounding}")


GOLD = np.load(__import__("pathlib").Path(__file__).resolve().parent / "golden" / "ref_gold.npz")


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("tag", ["f64", "f32"])
@pytest.mark.parametrize("method", ["plain", "persistent"])
def test_stream_matches_committed_reference_gold(gpu, shape, tag, method):
    field = GOLD[f"{shape}_{tag}_in"]
    steps = int(GOLD["steps_2d"])
    say(f"[j2s] reference gold vector {shape} {tag} {method} dims={field.shape} steps={steps}")
    got, _ = P.jacobi2d_run(field, steps, shape=shape, method=method, boundary="clamp", tune=tune())
    check_equal(got, GOLD[f"{shape}_{tag}_out"], f"{shape} {tag} {method} vs reference gold")


def _batch(torch, dtype, nb, seed):
    rng = np.random.default_rng(seed)
    specs = [(int(rng.integers(1, 180)), int(rng.integers(1, 260)), int(rng.integers(1, 30))) for _ in range(nb)]
    np_dt = np.float64 if dtype == "f64" else np.float32
    fields = [O.uniform(wy * wx, 300 + i, 0.0, 1.0, np_dt).reshape(wy, wx) for i, (wy, wx, _) in enumerate(specs)]
    bufs = [(torch.from_numpy(f.copy()).view(-1).cuda(),
             torch.empty(f.size, dtype=torch.from_numpy(f).dtype, device="cuda")) for f in fields]
    return specs, fields, bufs


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("method", ["plain", "persistent"])
def test_stream_batch_matches_per_grid_oracle(gpu, dtype, method):
    torch = gpu
    specs, fields, bufs = _batch(torch, dtype, 14, 31)
    say(f"[j2s] batch {dtype} {method} {specs}")
    si = P.shape_info("2d9pt")
    coefs = [0.1, 0.1, 0.1, 0.1, 0.2, 0.1, 0.1, 0.1, 0.1]
    rep, flags = P.jacobi2d_batch([(a, b, wy, wx, st) for (a, b), (wy, wx, st) in zip(bufs, specs)], dtype=dtype,
                                  shape="2d9pt", method=method, boundary="zero", coefs=coefs, tune=tune())
    for i, ((a, b), (wy, wx, st)) in enumerate(zip(bufs, specs)):
        got = (b if flags[i] else a).cpu().numpy().reshape(wy, wx)
        want = O.iterate(fields[i], O.stencil_from_shape(si, "zero", coefs=coefs), st)
        check_equal(got, want, f"grid {i} {wy}x{wx} steps {st}")


def test_stream_batch_plan_graph_replays(gpu):
    torch = gpu
    specs, fields, bufs = _batch(torch, "f32", 10, 21)
    say(f"[j2s] batch plan {specs}")
    si = P.shape_info("2d5pt")
    coefs = [0.2] * 5
    plan = P.BatchPlan([(a, b, wy, wx, st) for (a, b), (wy, wx, st) in zip(bufs, specs)], dtype="f32",
                       autotune=False, shape="2d5pt", method="persistent", boundary="clamp", coefs=coefs,
                       tune=tune())
    say(f"[j2s] plan config {plan.config}")
    assert plan.config["tile"] == STREAM
    wants = [O.iterate(f, O.stencil_from_shape(si, "clamp", coefs=coefs), st) for f, (_, _, st) in zip(fields, specs)]
    for replay in range(3):
        for (a, b), f in zip(bufs, fields):
            a.copy_(torch.from_numpy(f.copy()).view(-1).cuda())
            b.fill_(-1.0)
        torch.cuda.synchronize()
        plan.run()
        for i, ((a, b), (wy, wx, st)) in enumerate(zip(bufs, specs)):
            got = (b if plan.result_in_b[i] else a).cpu().numpy().reshape(wy, wx)
            check_equal(got, wants[i], f"replay {replay} grid {i} {wy}x{wx} steps {st}")
    plan.close()


def test_stream_batch_plan_survives_release(gpu):
    """A plan owns its barrier block: perks_release() between planning and
    running frees the library's pooled scratch but leaves the plan valid."""
    torch = gpu
    specs, fields, bufs = _batch(torch, "f32", 6, 23)
    say(f"[j2s] plan -> release -> run {specs}")
    si = P.shape_info("2d5pt")
    coefs = [0.2] * 5
    plan = P.BatchPlan([(a, b, wy, wx, st) for (a, b), (wy, wx, st) in zip(bufs, specs)], dtype="f32",
                       autotune=False, shape="2d5pt", method="persistent", boundary="zero", coefs=coefs,
                       tune=tune())
    # a direct persistent run leases a pooled block, which release then frees
    f0 = fields[0]
    P.jacobi2d_run(f0, 3, shape="2d5pt", method="persistent", boundary="zero", coefs=coefs)
    P.release()
    wants = [O.iterate(f, O.stencil_from_shape(si, "zero", coefs=coefs), st) for f, (_, _, st) in zip(fields, specs)]
    for replay in range(2):
        for (a, b), f in zip(bufs, fields):
            a.copy_(torch.from_numpy(f.copy()).view(-1).cuda())
            b.fill_(-1.0)
        torch.cuda.synchronize()
        plan.run()
        for i, ((a, b), (wy, wx, st)) in enumerate(zip(bufs, specs)):
            got = (b if plan.result_in_b[i] else a).cpu().numpy().reshape(wy, wx)
            check_equal(got, wants[i], f"after release, replay {replay} grid {i} {wy}x{wx} steps {st}")
        P.release()
    plan.close()


# ---- two columns per lane (tile variant 4; even widths, FP32/FP16/BF16) -------------
PAIR = 4


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("method", ["plain", "persistent"])
@pytest.mark.parametrize("boundary", ["clamp", "zero"])
def test_pair_shapes_match_oracle(gpu, shape, method, boundary):
    dims, steps = (70, 150), 11
    say(f"[j2s] pairs {shape} f32 {method} {boundary} dims={dims} steps={steps}")
    si = P.shape_info(shape)
    field = O.libc_rand(dims[0] * dims[1], 3, np.float32).reshape(dims)
    got, rep = P.jacobi2d_run(field, steps, shape=shape, method=method, boundary=boundary,
                              tune=[0, 0, 0, 0, PAIR])
    assert rep.reserved[2] == PAIR
    want = O.iterate(field, O.stencil_from_shape(si, boundary), steps)
    check_equal(got, want, f"{shape} {method} {boundary}")


@pytest.mark.parametrize("dims,steps,depth", [((2, 2), 3, 0), ((5, 300), 9, 0), ((257, 34), 17, 0),
                                               ((200, 200), 20, 3), ((131, 262), 25, 4), ((3, 70), 7, 2),
                                               ((300, 2), 6, 4), ((97, 120), 8, 4), ((97, 122), 9, 1),
                                               ((41, 130), 13, 3)])
@pytest.mark.parametrize("boundary", ["clamp", "zero"])
def test_pair_sizes_and_pass_depths(gpu, dims, steps, depth, boundary):
    say(f"[j2s] pairs sizes dims={dims} steps={steps} depth={depth} {boundary}")
    for shape in ("2d9pt", "2d5pt", "2d25pt"):
        si = P.shape_info(shape)
        field = O.uniform(dims[0] * dims[1], 4, 0.0, 1.0, np.float32).reshape(dims)
        got, _ = P.jacobi2d_run(field, steps, shape=shape, method="persistent", boundary=boundary,
                                tune=[0, 0, depth, 0, PAIR])
        want = O.iterate(field, O.stencil_from_shape(si, boundary), steps)
        check_equal(got, want, f"{shape} dims {dims} depth {depth}")


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
@pytest.mark.parametrize("shape", ["2d5pt", "2d9pt", "2d25pt", "2d21pt"])
@pytest.mark.parametrize("method", ["plain", "persistent"])
@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
def test_pair_16bit_storage(gpu, fmt, shape, method, rounding):
    torch = gpu
    dims, steps = (90, 200), 13
    say(f"[j2s] pairs storage {fmt} {shape} {method} {rounding}")
    si = P.shape_info(shape)
    rng = np.random.default_rng(5)
    c = rng.random(len(si.coef))
    coefs = list(c / c.sum())
    n = dims[0] * dims[1]
    a = torch.empty(n, dtype=torch.int16, device="cuda")
    b = torch.empty(n, dtype=torch.int16, device="cuda")
    P.fill_uniform(a, n, 21, -1.0, 1.0, dtype=fmt)
    init = a.cpu().numpy().view(np.uint16).reshape(dims).copy()
    rep = P.jacobi2d_device(a, b, dims[0], dims[1], steps, dtype=fmt, shape=shape, method=method, boundary="clamp",
                            coefs=coefs, rounding=rounding, round_seed=77, tune=[0, 0, 0, 0, PAIR])
    got = (b if rep.result_in_b else a).cpu().numpy().view(np.uint16).reshape(dims)
    st = O.stencil_from_shape(si, "clamp", coefs=coefs)
    want = (O.iterate16_sr(init, st, steps, fmt, 77) if rounding == "stochastic"
            else O.iterate16(init, st, steps, fmt))
    check_equal(got, want, f"{fmt} {shape} {method} {rounding}")


def test_pair_needs_even_widths_and_the_default_falls_back(gpu):
    field = O.uniform(40 * 51, 4, 0.0, 1.0, np.float32).reshape(40, 51)
    out = np.empty_like(field)
    rc = P.jacobi_iterative(field, 40, 51, out, iteration=3, shape="2d5pt", tune=[0, 0, 0, 0, PAIR])
    assert rc == -13  # PERKS_ERR_UNSUPPORTED: odd width
    got, rep = P.jacobi2d_run(field, 3, shape="2d5pt")
    assert rep.reserved[2] == STREAM
    check_equal(got, O.iterate(field, O.stencil_from_shape(P.shape_info("2d5pt"), "clamp"), 3), "default, odd width")
    even = O.uniform(40 * 52, 4, 0.0, 1.0, np.float32).reshape(40, 52)
    _, rep = P.jacobi2d_run(even, 3, shape="2d5pt")
    assert rep.reserved[2] == PAIR
    f64 = even.astype(np.float64)
    _, rep = P.jacobi2d_run(f64, 3, shape="2d5pt")
    assert rep.reserved[2] == STREAM


def test_pair_batch_plan_matches_oracle(gpu):
    torch = gpu
    rng = np.random.default_rng(41)
    specs = [(int(rng.integers(1, 180)), 2 * int(rng.integers(1, 130)), int(rng.integers(1, 30))) for _ in range(12)]
    fields = [O.uniform(wy * wx, 300 + i, 0.0, 1.0, np.float32).reshape(wy, wx) for i, (wy, wx, _) in enumerate(specs)]
    bufs = [(torch.from_numpy(f.copy()).view(-1).cuda(), torch.empty(f.size, dtype=torch.float32, device="cuda"))
            for f in fields]
    say(f"[j2s] pair batch plan {specs}")
    si = P.shape_info("2d9pt")
    coefs = [0.1, 0.1, 0.1, 0.1, 0.2, 0.1, 0.1, 0.1, 0.1]
    plan = P.BatchPlan([(a, b, wy, wx, st) for (a, b), (wy, wx, st) in zip(bufs, specs)], dtype="f32",
                       autotune=False, shape="2d9pt", method="persistent", boundary="zero", coefs=coefs)
    assert plan.config["tile"] == PAIR
    plan.run()
    for i, ((a, b), (wy, wx, st)) in enumerate(zip(bufs, specs)):
        got = (b if plan.result_in_b[i] else a).cpu().numpy().reshape(wy, wx)
        want = O.iterate(fields[i], O.stencil_from_shape(si, "zero", coefs=coefs), st)
        check_equal(got, want, f"grid {i} {wy}x{wx} steps {st}")
    plan.close()
