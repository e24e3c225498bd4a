"""The 3D configuration table and autotuner.

The table lists every compiled configuration for a problem with what its code
object reports; no configuration of the library may spill registers (the
condition the reference's register-folder search enforces,
ProcessRst2cuh.py:5-51).  The autotuner's choice must compute the same bits
as the per-step kernel and leave the caller's field unchanged.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import perks_amd as P
from perks_amd import _lib as L

pytestmark = pytest.mark.gpu

SHAPES3 = ["3d7pt", "3d13pt", "3d17pt", "3d27pt", "poisson"]
RESIDENT = 5  # opts.tune[4]: the resident-tile form of the 7-point persistent method
TALL = 4      # opts.tune[4], other 3D shapes: the streaming kernel's 64 x 16 tile
DTYPES = ["f64", "f32", "f16", "bf16"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", SHAPES3)
def test_table_lists_every_configuration_and_none_spills(gpu, shape, dtype):
    say(f"[tune] table {shape} {dtype}")
    rows = P.j3d_kernel_table(64, 64, 64, dtype=dtype, shape=shape, boundary="clamp")
    assert rows
    resident = [r for r in rows if r["method"] == "persistent" and r["variant"] == RESIDENT]
    if shape == "3d7pt" and dtype in ("f16", "bf16"):  # resident tiles: FP64/FP32 storage only
        assert [r["status"] for r in resident] == [L.ERR_UNSUPPORTED]
        rows = [r for r in rows if r not in resident]
    tall = [r for r in rows if shape != "3d7pt" and r["variant"] == TALL]
    if dtype == "f64" and shape != "3d7pt":  # the tall streaming tile: FP32 arithmetic only
        assert tall and all(r["status"] == L.ERR_UNSUPPORTED for r in tall)
        rows = [r for r in rows if r not in tall]
    for r in rows:
        assert r["vgprs"] > 0, r
        assert r["spill_bytes"] == 0, r
        assert r["status"] == L.PERKS_OK, r
        assert r["occupancy"] > 0, r
        assert r["block"] % 64 == 0, r
    methods = {r["method"] for r in rows}
    assert methods == {"plain", "persistent"}
    if shape == "3d7pt":
        assert len(rows) == 1 + 4 * 5 + (1 if dtype in ("f64", "f32") else 0)
        assert sorted({r["levels"] for r in rows if r["method"] == "persistent"}) == [1, 2, 3, 4]
    else:
        assert {r["variant"] for r in rows} == ({2, 3} if dtype == "f64" else {2, 3, TALL})


@pytest.mark.parametrize("dtype", ["f16", "bf16"])
def test_stochastic_rounding_table_has_no_spills(gpu, dtype):
    """Stochastic rounding makes the largest kernel bodies (the two-column
    7-point variant at four levels once kept its whole march object in scratch
    when the compiler declined to inline it): every entry must be spill-free."""
    rows = P.j3d_kernel_table(64, 64, 64, dtype=dtype, shape="3d7pt", boundary="zero", rounding="stochastic")
    rows = [r for r in rows if r["variant"] != RESIDENT]  # FP64/FP32 storage only
    for r in rows:
        assert r["spill_bytes"] == 0 and r["status"] == L.PERKS_OK, r


def test_two_column_variant_refuses_odd_widths(gpu):
    rows = P.j3d_kernel_table(32, 40, 67, dtype="f32", shape="3d7pt", boundary="zero")
    pair = [r for r in rows if r["method"] == "persistent" and r["variant"] == 4]
    assert pair and all(r["status"] == L.ERR_UNSUPPORTED for r in pair)
    assert all(r["status"] == L.PERKS_OK for r in rows if r["variant"] != 4)  # resident tiles take odd widths
    rows16 = P.j3d_kernel_table(32, 40, 67, dtype="bf16", shape="3d7pt", boundary="zero")
    plain = [r for r in rows16 if r["method"] == "plain"]
    assert plain[0]["status"] == L.ERR_UNSUPPORTED_SIZE


def per_step(torch, init, nz, ny, nx, steps, dtype, **kw):
    a = init.clone()
    b = torch.empty_like(a)
    rep = P.j3d_device(a, b, nz, ny, nx, steps, dtype=dtype, method="plain", **kw)
    torch.cuda.synchronize()
    return b if rep.result_in_b else a


CASES = [("3d7pt", "f64", "zero", "nearest"), ("3d7pt", "f32", "clamp", "nearest"),
         ("3d7pt", "bf16", "clamp", "stochastic"), ("3d27pt", "f32", "clamp", "nearest"),
         ("3d13pt", "f16", "zero", "stochastic"), ("poisson", "f64", "clamp", "nearest")]


@pytest.mark.parametrize("shape,dtype,boundary,rounding", CASES)
def test_autotuned_choice_is_bit_identical_to_per_step(gpu, shape, dtype, boundary, rounding):
    torch = gpu
    nz, ny, nx, steps = 40, 72, 130, 7
    say(f"[tune] autotune {shape} {dtype} {boundary} {rounding} {nz}x{ny}x{nx}")
    tdt = {"f64": torch.float64, "f32": torch.float32}.get(dtype, torch.int16)
    init = torch.empty(nz * ny * nx, dtype=tdt, device="cuda")
    P.fill_uniform(init, init.numel(), 3, 0.0, 1.0, dtype=dtype)
    keep = init.clone()
    kw = dict(shape=shape, boundary=boundary, rounding=rounding, round_seed=9)
    opts, rows = P.j3d_autotune(init, nz, ny, nx, dtype=dtype, sample_steps=8, **kw)
    torch.cuda.synchronize()
    assert torch.equal(init, keep), "the autotuner changed the caller's field"
    timed = [r for r in rows if r["status"] == L.PERKS_OK]
    assert timed and all(r["gcells_per_s"] > 0 for r in timed)
    best = max(timed, key=lambda r: r["gcells_per_s"])
    assert opts.method == (L.METHOD_PLAIN if best["method"] == "plain" else L.METHOD_PERSISTENT)
    assert opts.tune[4] == best["variant"]
    say(f"[tune] chose {best['method']} variant {best['variant']} levels {best['levels']}")
    ref = per_step(torch, init, nz, ny, nx, steps, dtype, **kw)
    a = init.clone()
    b = torch.empty_like(a)
    rep = P.j3d_device(a, b, nz, ny, nx, steps, dtype=dtype, opts=opts)
    torch.cuda.synchronize()
    got = b if rep.result_in_b else a
    assert torch.equal(got, ref)


def test_autotune_on_host_field_matches_oracle(gpu):
    """The tuned options drive the host-buffer entry point to the oracle's bits."""
    import oracle as O
    torch = gpu
    nz, ny, nx, steps = 24, 33, 70, 6
    field = O.uniform(nz * ny * nx, 5, 0.0, 1.0)
    d = torch.from_numpy(field).cuda()
    opts, _ = P.j3d_autotune(d, nz, ny, nx, dtype="f64", shape="3d7pt", boundary="clamp")
    got = np.empty_like(field)
    rep = L.RunReport()
    L.check(P.load().perks_j3d_iterative_f64(field.ctypes.data, nz, ny, nx, got.ctypes.data, steps, C.byref(opts),
                                              C.byref(rep)), "j3d_iterative_f64")
    want = O.iterate(field.reshape(nz, ny, nx), O.stencil_from_shape(P.shape_info("3d7pt"), "clamp"), steps)
    assert np.array_equal(got.reshape(nz, ny, nx).view(np.uint8), want.view(np.uint8))
