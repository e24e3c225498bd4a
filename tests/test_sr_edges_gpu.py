"""Stochastic rounding across the 16-bit formats' special ranges, on the GPU
against the oracle bit for bit: fields whose magnitudes run from below the
FP16 subnormal range to the FP16 maximum, grown past it by weights that sum
to more than 1, so results are rounded as subnormals, normals and overflow
to inf in the same run (all positive, so no NaN is formed).  3D: the
per-step 16-bit kernel and the persistent 7-point kernel (every tile
variant); 2D: the row-streaming kernel, both methods."""
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P

pytestmark = pytest.mark.gpu

SEED = 91


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def wide_field(shape, fmt, seed, lo, hi):
    """Positive values log-uniform in [10^lo, 10^hi], rounded to the format."""
    rng = np.random.default_rng(seed)
    mag = 10.0 ** rng.uniform(lo, hi, int(np.prod(shape)))
    if fmt == "f16":
        bits = mag.astype(np.float16).view(np.uint16)
    else:
        f = mag.astype(np.float32).view(np.uint32)
        bits = ((f + 0x7FFF + ((f >> 16) & 1)) >> 16).astype(np.uint16)  # nearest-even bfloat16
    return bits.reshape(shape)


def as_float(bits, fmt):
    return bits.view(np.float16).astype(np.float64) if fmt == "f16" else (bits.astype(np.uint32) << 16).view(np.float32)


# (weights of the 3D 7-point star, range of the initial magnitudes, steps): growth
# into overflow, decay through the FP16 subnormal range to zero
CASES_3D = [([0.25] * 6 + [0.5], (-9.0, 4.8), 1), ([0.25] * 6 + [0.5], (-9.0, 4.8), 4),
            ([0.05] * 6 + [0.1], (-9.0, -2.0), 1), ([0.05] * 6 + [0.1], (-9.0, -2.0), 5)]
CASES_2D = [([0.5] * 5, (-9.0, 4.8), 1), ([0.5] * 5, (-9.0, 4.8), 5),
            ([0.08] * 5, (-9.0, -2.0), 1), ([0.08] * 5, (-9.0, -2.0), 6)]


def check_ranges(wants, fmt):
    if fmt != "f16":
        return
    w = np.concatenate([as_float(x, fmt).ravel() for x in wants])
    assert np.isinf(w).any(), "no overflow reached"
    assert ((w > 0) & (w < 6.1e-5)).any(), "no FP16 subnormal result"
    assert (w == 0).any() and ((w >= 6.1e-5) & np.isfinite(w)).any()


def run_device(torch, init, fn, *args, **kw):
    a = torch.from_numpy(init.reshape(-1).view(np.int16).copy()).cuda()
    b = torch.empty_like(a)
    rep = fn(a, b, *args, **kw)
    torch.cuda.synchronize()
    return (b if rep.result_in_b else a).cpu().numpy().view(np.uint16).reshape(init.shape)


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
def test_sr_special_ranges_3d(gpu, fmt):
    torch = gpu
    dims = (7, 37, 130)
    runs = [("plain", None)] + [("persistent", [0, 0, 0, 0, v]) for v in (0, 1, 2, 3, 4)]
    wants = []
    for i, (coefs, (lo, hi), steps) in enumerate(CASES_3D):
        init = wide_field(dims, fmt, i, lo, hi)
        st = O.stencil_from_shape(P.shape_info("3d_star_r1"), "zero", coefs=coefs)
        want = O.iterate16_sr(init, st, steps, fmt, SEED)
        wants.append(want)
        for method, tune in runs:
            say(f"[sr-edges] 3d {fmt} case={i} {method} tune={tune}")
            got = run_device(torch, init, P.j3d_device, *dims, steps, dtype=fmt, method=method, boundary="zero",
                             coefs=coefs, rounding="stochastic", round_seed=SEED, tune=tune)
            assert np.array_equal(got, want), f"case {i} {method} {tune}"
    check_ranges(wants, fmt)


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
@pytest.mark.parametrize("method", ["plain", "persistent"])
def test_sr_special_ranges_2d(gpu, fmt, method):
    torch = gpu
    dims = (90, 260)
    wants = []
    for i, (coefs, (lo, hi), steps) in enumerate(CASES_2D):
        init = wide_field(dims, fmt, 10 + i, lo, hi)
        st = O.stencil_from_shape(P.shape_info("2d5pt"), "clamp", coefs=coefs)
        want = O.iterate16_sr(init, st, steps, fmt, SEED)
        wants.append(want)
        for tune in (None, [0, 0, 0, 0, 2], [0, 0, 0, 0, 1]):
            say(f"[sr-edges] 2d {fmt} case={i} {method} tune={tune}")
            got = run_device(torch, init, P.jacobi2d_device, *dims, steps, dtype=fmt, shape="2d5pt", method=method,
                             boundary="clamp", coefs=coefs, rounding="stochastic", round_seed=SEED, tune=tune)
            assert np.array_equal(got, want), f"case {i} {method} {tune}"
    check_ranges(wants, fmt)
