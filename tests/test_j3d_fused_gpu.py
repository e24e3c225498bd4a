"""The contracted 7-point update of the persistent kernel (opts.tune[6] = 1).

Each product after the first is fused into the running sum by an FMA, in the
gold's term order.  The result must equal the oracle's contracted restatement
(po_stencil_iterate_fused_*) bit for bit, for every tile variant, pass depth
and boundary, FP64 and FP32; the option is refused where it is not built.  At
the benchmark's size it stays within a stated distance of the uncontracted
(bit-identical-to-gold) result.
"""
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P
from perks_amd import _lib as L

pytestmark = pytest.mark.gpu

CONFIG_3D = [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4]
FUSED = 1
RESIDENT = 5  # opts.tune[4]: the resident-tile variant


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def tune(depth=0, variant=0):
    return [0, 0, depth, 0, variant, 0, FUSED]


@pytest.mark.parametrize("boundary", ["zero", "clamp"])
@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fused_matches_oracle(gpu, dtype, variant, boundary):
    si = P.shape_info("3d_star_r1")
    st = O.stencil_from_shape(si, boundary, coefs=CONFIG_3D)
    for depth in (1, 2, 3, 4):
        for dims in ((11, 37, 70), (9, 61, 130)):
            steps = 2 * depth + 1
            say(f"[fused] {np.dtype(dtype).name} variant={variant} {boundary} depth={depth} dims={dims}")
            field = O.uniform(int(np.prod(dims)), depth * 31 + dims[1], 0.0, 1.0, dtype=dtype).reshape(dims)
            got, rep = P.j3d_run(field, steps, method="persistent", boundary=boundary, coefs=CONFIG_3D,
                                 tune=tune(depth, variant))
            assert rep.reserved[0] == depth
            want = O.iterate_fused(field, st, steps)
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), f"depth={depth} dims={dims}"
            # and it is the contracted form, not the gold's bits
            assert not np.array_equal(got, O.iterate(field, st, steps))


def test_fused_table_has_no_spills(gpu):
    for dtype in ("f64", "f32"):
        rows = P.j3d_kernel_table(64, 64, 64, dtype=dtype, shape="3d7pt", boundary="zero", tune=tune())
        persistent = [r for r in rows if r["method"] == "persistent" and r["variant"] != RESIDENT]
        assert len(persistent) == 4 * 5
        for r in persistent:
            assert r["spill_bytes"] == 0 and r["status"] == L.PERKS_OK, r
        # the per-step kernel and the resident-tile one are built without it
        assert all(r["status"] == L.ERR_UNSUPPORTED for r in rows if r not in persistent)


def test_fused_refused_where_not_built(gpu):
    torch = gpu
    f = np.zeros((4, 8, 8))
    for kw in (dict(method="plain"), dict(divisor=2.0), dict(shape="3d27pt")):
        args = dict(method="persistent", boundary="zero", coefs=CONFIG_3D, tune=tune())
        if "shape" in kw:
            args.pop("coefs")
        args.update(kw)
        with pytest.raises(L.PerksError) as e:
            P.j3d_run(f, 3, **args)
        assert e.value.code == L.ERR_UNSUPPORTED, kw
    a = torch.zeros(4 * 8 * 8, dtype=torch.int16, device="cuda")
    b = torch.empty_like(a)
    with pytest.raises(L.PerksError) as e:
        P.j3d_device(a, b, 4, 8, 8, 3, dtype="bf16", method="persistent", boundary="zero", tune=tune())
    assert e.value.code == L.ERR_UNSUPPORTED
    with pytest.raises(L.PerksError) as e:
        P.j3d_run(f, 3, method="persistent", boundary="zero", coefs=CONFIG_3D, tune=[0, 0, 0, 0, 0, 0, 2])
    assert e.value.code == L.ERR_INVALID_ARG


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_fused_full_size_stays_near_gold(gpu, dtype):
    """512^3, 1000 steps, the benchmark's inputs: the contracted result against
    the uncontracted one (which the other tests pin to the gold bit for bit).
    Measured: FP64 1.3e-14, FP32 9.4e-6 of the field's largest value; the
    bounds leave margin."""
    torch = gpu
    n, T = 512, 1000
    tdt = torch.float64 if dtype == "f64" else torch.float32
    init = torch.empty(n ** 3, dtype=tdt, device="cuda")
    P.fill_uniform(init, init.numel(), 1, 0.0, 1.0, dtype=dtype)
    outs = []
    for t in ([0, 0, 0, 0, 0, 0, 0], tune()):
        say(f"[fused] {dtype} 512^3 x {T} tune={t}")
        a = init.clone()
        b = torch.empty_like(a)
        rep = P.j3d_device(a, b, n, n, n, T, dtype=dtype, method="persistent", boundary="zero", coefs=CONFIG_3D,
                           tune=t)
        torch.cuda.synchronize()
        outs.append((b if rep.result_in_b else a).double())
        del a, b
    scale = outs[0].abs().max().item()
    dev = (outs[1] - outs[0]).abs().max().item() / scale
    say(f"[fused] {dtype} max deviation {dev:.3e} of the largest value {scale:.3e}")
    assert 0 < dev < (1e-12 if dtype == "f64" else 1e-3)
