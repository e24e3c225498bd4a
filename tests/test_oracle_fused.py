"""The oracle's contracted stencil (po_stencil_iterate_fused_*), pinned on the CPU.

The contracted form is this library's option (opts.tune[6] = 1), with no
reference counterpart: the gold's term order, the first product rounded on
its own and each later one fused into the running sum with a single rounding.
Pinned here against exact rational arithmetic (a fused multiply-add is the
exact a*b + c rounded once to the format, ties to even), which an unfused
restatement, a wrong term order or a wrong first term fails; then its distance
from the uncontracted gold is measured in ulp.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
import perks_amd as P

CONFIG_3D = [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4]
# the 7-point star's neighbours in the kernels' / gold's term order: E W N S T B C
ORDER = [(0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0), (0, 0, 0)]


def round_f64(q: Fraction) -> float:
    return float(q)  # int / int true division: correctly rounded, ties to even


def round_f32(q: Fraction) -> float:
    if q == 0:
        return 0.0
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    e = max(e, -126)
    quantum = Fraction(2) ** (e - 23)
    m = round(a / quantum)  # Fraction rounding: half to even
    v = float(m * quantum)  # exact in double
    return -v if q < 0 else v


def exact_fused_step(field, coefs, boundary, rnd, divisor=0.0):
    nz, ny, nx = field.shape
    out = np.empty_like(field)
    k = [Fraction(rnd(Fraction(c))) for c in coefs]

    def at(z, y, x):
        if boundary == "clamp":
            z, y, x = min(max(z, 0), nz - 1), min(max(y, 0), ny - 1), min(max(x, 0), nx - 1)
        elif not (0 <= z < nz and 0 <= y < ny and 0 <= x < nx):
            return Fraction(0)
        return Fraction(float(field[z, y, x]))

    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                acc = None
                for g, (dz, dy, dx) in enumerate(ORDER):
                    v = at(z + dz, y + dy, x + dx)
                    acc = Fraction(rnd(k[g] * v)) if g == 0 else Fraction(rnd(k[g] * v + acc))
                if divisor:
                    acc = Fraction(rnd(acc / Fraction(rnd(Fraction(divisor)))))
                out[z, y, x] = float(acc)
    return out


def star(boundary, divisor=None):
    return O.stencil_from_shape(P.shape_info("3d_star_r1"), boundary, coefs=CONFIG_3D, divisor=divisor)


@pytest.mark.parametrize("boundary", ["clamp", "zero"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fused_step_is_exact_fma_chain(dtype, boundary):
    rnd = round_f64 if dtype == np.float64 else round_f32
    field = O.uniform(3 * 5 * 6, 41, -1.0, 1.0, dtype=dtype).reshape(3, 5, 6)
    st = star(boundary)
    one = O.iterate_fused(field, st, 1)
    want = exact_fused_step(field, CONFIG_3D, boundary, rnd)
    assert np.array_equal(one.view(np.uint8), want.astype(dtype).view(np.uint8))
    # two steps: the ping-pong feeds the first step's result
    two = O.iterate_fused(field, st, 2)
    want2 = exact_fused_step(want.astype(dtype), CONFIG_3D, boundary, rnd)
    assert np.array_equal(two.view(np.uint8), want2.astype(dtype).view(np.uint8))
    # a chain of separately rounded products and sums differs somewhere
    assert not np.array_equal(one, O.iterate(field, st, 1))


def test_fused_divisor_rounds_after_the_chain():
    field = O.uniform(2 * 4 * 5, 9, 0.0, 1.0).reshape(2, 4, 5)
    got = O.iterate_fused(field, star("clamp", divisor=3.0), 1)
    want = exact_fused_step(field, CONFIG_3D, "clamp", round_f64, divisor=3.0)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_fused_zero_steps_copies():
    field = O.uniform(60, 3, 0.0, 1.0).reshape(3, 4, 5)
    assert np.array_equal(O.iterate_fused(field, star("zero"), 0), field)


def ulps(a, b):
    """Distance in units in the last place of the larger magnitude."""
    sp = np.spacing(np.maximum(np.abs(a), np.abs(b))).astype(np.float64)
    return np.abs(a.astype(np.float64) - b.astype(np.float64)) / sp


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fused_within_few_ulp_of_gold_per_step(dtype):
    """Per step, on the benchmark's inputs (uniform [0, 1), coefficients
    0.4 / 6 x 0.1, zero halo): the contracted result is within 4 ulp of the
    uncontracted gold's at every cell (3 measured at 40^3, FP64 and FP32; about
    30% of the cells differ)."""
    field = O.uniform(40 * 40 * 40, 5, 0.0, 1.0, dtype=dtype).reshape(40, 40, 40)
    st = star("zero")
    d = ulps(O.iterate_fused(field, st, 1), O.iterate(field, st, 1))
    assert d.max() <= 4.0
    assert 0.1 < (d > 0).mean() < 0.6


def test_fp32_double_formed_groups_refused():
    si = P.shape_info("3d_star_r1")
    st = O.stencil([(c, si.offsets[g], 1) for g, c in enumerate(CONFIG_3D)], "zero", 0.0)
    with pytest.raises(ValueError):
        O.iterate_fused(np.zeros((2, 2, 2), np.float32), st, 1)
