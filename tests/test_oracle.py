"""Pins the CPU oracle before it is trusted as the checker.

* against the reference's own CPU gold, compiled from its sources into
  oracle/_ref (every stencil shape, FP64 and FP32, the reference's rand() input);
* against the reference's interior-only 4-step j3d_gold for the zero-halo form;
* against closed forms (a separable sine mode is an eigenvector of the
  symmetric star stencils with zero halo; a constant field under clamp);
* against an independent numpy restatement (padding instead of index clamps)
  on small grids, bit for bit;
* against the same gold's outputs committed as fixtures (tests/golden), which
  hold where oracle/_ref cannot be built;
* the counter RNG against SplitMix64's published first outputs;
* the FP16/BF16 rounding against numpy's float16 and torch's bfloat16.
"""
import math
import pathlib

import numpy as np
import pytest

import oracle as O
import perks_amd as P

REF_3D = {"3d7pt": "3d7pt_gold_gen", "3d13pt": "3d13pt_gold_gen", "3d17pt": "3d17pt_gold_gen",
          "3d27pt": "3d27pt_gold_gen", "poisson": "poisson_gold_gen"}
REF_2D = {"2d5pt": "2d5pt_gold_gen", "2ds9pt": "2ds9pt_gold_gen", "2d13pt": "2d13pt_gold_gen",
          "2d17pt": "2d17pt_gold_gen", "2d21pt": "2d21pt_gold_gen", "2ds25pt": "2ds25pt_gold_gen",
          "2d9pt": "2d9pt_gold_gen", "2d25pt": "2d25pt_gold_gen"}


def need_ref(name):
    if not O.ref_available(name):
        pytest.skip(f"oracle/_ref/libref_{name}.so not built (reference sources absent)")


# ---------------------------------------------------------------- RNG
def test_splitmix64_known_answers():
    # SplitMix64 from state 0: published first outputs (Vigna's splitmix64.c).
    assert O.lib().po_splitmix64(0, 0) == 0xE220A8397B1DCDAF
    assert O.lib().po_splitmix64(0, 1) == 0x6E789E6AA1B965F4
    assert O.lib().po_splitmix64(0, 2) == 0x06C45D188009454F


def test_uniform_range_and_determinism():
    a = O.uniform(100000, 7, 0.0, 1.0)
    b = O.uniform(100000, 7, 0.0, 1.0)
    assert np.array_equal(a, b)
    assert a.min() >= 0.0 and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.01
    c = O.uniform(1000, 8, -1.0, 1.0)
    assert c.min() >= -1.0 and c.max() < 1.0
    assert not np.array_equal(a[:1000], O.uniform(1000, 8, 0.0, 1.0))


# ---------------------------------------------------------------- reference gold
@pytest.mark.parametrize("ref", sorted(REF_3D))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims,steps", [((5, 6, 7), 1), ((9, 8, 12), 4), ((4, 5, 3), 3)])
def test_oracle_matches_reference_gold_3d(ref, dtype, dims, steps):
    need_ref(REF_3D[ref])
    si = P.shape_info(ref)
    field = O.libc_rand(int(np.prod(dims)), 1, dtype).reshape(dims)
    want = O.ref_gold_iterative(REF_3D[ref], field, steps)
    got = O.iterate(field, O.stencil_from_shape(si, "clamp"), steps)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), ref


@pytest.mark.parametrize("ref", sorted(REF_2D))
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims,steps", [((13, 17), 1), ((20, 16), 5), ((7, 9), 2)])
def test_oracle_matches_reference_gold_2d(ref, dtype, dims, steps):
    need_ref(REF_2D[ref])
    si = P.shape_info(ref)
    field = O.libc_rand(int(np.prod(dims)), 3, dtype).reshape(dims)
    want = O.ref_gold_iterative(REF_2D[ref], field, steps)
    got = O.iterate(field, O.stencil_from_shape(si, "clamp"), steps)
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), ref


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_zero_halo_matches_reference_interior_gold(dtype):
    """3d7pt_gold.cpp's j3d_gold updates only interior cells and, after its
    first step, sees a zero boundary; with a zero boundary in the input it is
    four zero-halo steps on the interior."""
    need_ref("3d7pt_gold")
    n = (10, 9, 11)
    field = O.libc_rand(int(np.prod(n)), 5, dtype).reshape(n)
    field[0, :, :] = field[-1, :, :] = 0
    field[:, 0, :] = field[:, -1, :] = 0
    field[:, :, 0] = field[:, :, -1] = 0
    want = O.ref_j3d_gold_4step(field)
    si = P.shape_info("3d7pt")
    got = O.iterate(np.ascontiguousarray(field[1:-1, 1:-1, 1:-1]), O.stencil_from_shape(si, "zero"), 4)
    assert np.array_equal(got, want[1:-1, 1:-1, 1:-1])


# ---------------------------------------------------------------- independent numpy restatement
def numpy_step(u, groups, boundary, divisor=0.0):
    """Pads by the radius (edge replicate or zeros) and slices; same term order."""
    r = max(max(abs(c) for pt in pts for c in pt) for _, pts, _ in groups)
    mode = "edge" if boundary == "clamp" else "constant"
    up = np.pad(u, r, mode=mode)
    nz, ny, nx = u.shape

    def view(dz, dy, dx):
        return up[r + dz:r + dz + nz, r + dy:r + dy + ny, r + dx:r + dx + nx]

    acc = None
    for coef, pts, _ in groups:
        s = view(*pts[0])
        for pt in pts[1:]:
            s = s + view(*pt)
        prod = u.dtype.type(coef) * s
        acc = prod if acc is None else acc + prod
    if divisor:
        acc = acc / u.dtype.type(divisor)
    return acc


CONFIG_3D = [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4]  # E W N S T B C: neighbours 0.1, centre 0.4


@pytest.mark.parametrize("boundary", ["clamp", "zero"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_oracle_matches_numpy_restatement(boundary, dtype):
    si = P.shape_info("3d_star_r1")
    groups = [(CONFIG_3D[g], si.offsets[g], 0) for g in range(7)]
    u = O.uniform(6 * 7 * 9, 11, 0.0, 1.0, dtype).reshape(6, 7, 9)
    want = u
    for _ in range(3):
        want = numpy_step(want, groups, boundary)
    got = O.iterate(u, O.stencil(groups, boundary), 3)
    assert np.array_equal(got, want)


def test_oracle_box_matches_numpy_restatement_f64():
    si = P.shape_info("3d_box_r1")
    groups = [(si.coef[g], si.offsets[g], 1) for g in range(len(si.coef))]
    u = O.uniform(5 * 6 * 7, 12, -1.0, 1.0).reshape(5, 6, 7)
    want = numpy_step(numpy_step(u, groups, "zero", si.divisor), groups, "zero", si.divisor)
    got = O.iterate(u, O.stencil(groups, "zero", si.divisor), 2)
    assert np.array_equal(got, want)


# ---------------------------------------------------------------- closed forms
@pytest.mark.parametrize("n", [7, 16])
def test_sine_mode_is_eigenvector_3d(n):
    th = math.pi / (n + 1)
    s = np.sin(np.arange(1, n + 1) * th)
    u0 = np.einsum("k,j,i->kji", s, s, s).copy()
    lam = 0.4 + 0.6 * math.cos(th)
    steps = 10
    got = O.iterate(u0, O.stencil([(c, [o], 0) for c, o in zip(CONFIG_3D, [
        (0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0), (0, 0, 0)])], "zero"), steps)
    np.testing.assert_allclose(got, lam ** steps * u0, rtol=0, atol=1e-13)


def test_sine_mode_is_eigenvector_2d():
    n = 31
    th = math.pi / (n + 1)
    s = np.sin(np.arange(1, n + 1) * th)
    u0 = np.outer(s, s)
    si = P.shape_info("2d5pt")
    st = O.stencil_from_shape(si, "zero", coefs=[0.2] * 5)
    lam = 0.2 + 0.8 * math.cos(th)
    got = O.iterate(u0, st, 25)
    np.testing.assert_allclose(got, lam ** 25 * u0, rtol=0, atol=1e-13)


def test_constant_field_under_clamp():
    si = P.shape_info("3d7pt")
    c = 0.75
    u = np.full((4, 5, 6), c)
    got = O.iterate(u, O.stencil_from_shape(si, "clamp"), 3)
    ssum = sum(si.coef)
    np.testing.assert_allclose(got, c * ssum ** 3, rtol=1e-14)


def test_zero_steps_copies_and_parity():
    si = P.shape_info("3d7pt")
    u = O.uniform(3 * 4 * 5, 2, 0, 1).reshape(3, 4, 5)
    assert np.array_equal(O.iterate(u, O.stencil_from_shape(si), 0), u)
    one = O.iterate(u, O.stencil_from_shape(si), 1)
    two = O.iterate(one, O.stencil_from_shape(si), 1)
    assert np.array_equal(O.iterate(u, O.stencil_from_shape(si), 2), two)


def test_zero_field_stays_zero():
    si = P.shape_info("3d_box_r1")
    u = np.zeros((4, 4, 4))
    assert not O.iterate(u, O.stencil_from_shape(si, "zero"), 3).any()


def test_omp_is_bitwise_equal():
    si = P.shape_info("3d7pt")
    u = O.uniform(24 * 20 * 33, 4, 0, 1).reshape(24, 20, 33)
    st = O.stencil_from_shape(si, "zero", coefs=CONFIG_3D)
    assert np.array_equal(O.iterate(u, st, 3), O.iterate(u, st, 3, threads=4))


def test_rms_error_metric():
    a = np.zeros((2, 3, 4))
    b = np.zeros((2, 3, 4))
    b[1, 2, 3] = 3.0
    rms, mx = O.rms_error(a, b)
    assert mx == 3.0
    assert abs(rms - math.sqrt(9.0 / 24)) < 1e-15


# ---------------------------------------------------------------- 16-bit storage formats
def _f32_samples():
    rng = np.random.default_rng(17)
    bits = rng.integers(0, 2 ** 32, 60000, dtype=np.uint64).astype(np.uint32)
    # exact ties and neighbours of ties for binary16 (13 dropped bits) and bfloat16 (16 dropped bits),
    # the binary16 subnormal range, the overflow threshold and the specials
    base = rng.integers(0, 2 ** 32, 4000, dtype=np.uint64).astype(np.uint32)
    ties = np.concatenate([(base & ~np.uint32(0x1FFF)) | np.uint32(0x1000), (base & ~np.uint32(0xFFFF)) | np.uint32(0x8000)])
    ties = np.concatenate([ties, ties + np.uint32(1), ties - np.uint32(1)])
    sub = np.linspace(0, 2.0 ** -14, 5001, dtype=np.float32).view(np.uint32)
    special = np.array([0x00000000, 0x80000000, 0x7F800000, 0xFF800000, 0x477FE000, 0x477FEFFF, 0x477FF000,
                        0x477FF001, 0x33000000, 0x33000001, 0x32FFFFFF, 0x38800000, 0x387FFFFF, 0x7F7FFFFF],
                       dtype=np.uint32)
    return np.concatenate([bits, ties, sub, sub | np.uint32(0x80000000), special]).view(np.float32)


def _oracle_round(x, fmt):
    f = O.lib().po_f32_to_f16 if fmt == "f16" else O.lib().po_f32_to_bf16
    return np.array([f(float(v)) for v in x], dtype=np.uint16)


def test_f16_rounding_matches_numpy():
    x = _f32_samples()
    finite = ~np.isnan(x)
    got = _oracle_round(x[finite], "f16")
    with np.errstate(over="ignore"):
        want = x[finite].astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
    nan_bits = _oracle_round(x[np.isnan(x)], "f16")
    assert nan_bits.size > 0 and np.all(np.isnan(nan_bits.view(np.float16)))


def test_f16_widening_is_exact_for_every_pattern():
    h = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    got = np.array([O.lib().po_f16_to_f32(int(v)) for v in h], dtype=np.float32)
    want = h.view(np.float16).astype(np.float32)
    ok = ~np.isnan(want)
    assert np.array_equal(got[ok].view(np.uint32), want[ok].view(np.uint32))
    assert np.all(np.isnan(got[~ok]))


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    x = _f32_samples()
    finite = ~np.isnan(x)
    got = _oracle_round(x[finite], "bf16")
    want = torch.from_numpy(x[finite].copy()).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(got, want)
    back = np.array([O.lib().po_bf16_to_f32(int(v)) for v in got[:2000]], dtype=np.float32)
    assert np.array_equal(back, torch.from_numpy(got[:2000].view(np.int16).copy()).view(torch.bfloat16).float().numpy())


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
def test_reduced_storage_iteration_matches_numpy_restatement(fmt):
    """Each step in float32 on the widened values, then rounded to the storage
    format (numpy float16 / torch bfloat16 do the rounding here)."""
    torch = pytest.importorskip("torch")
    si = P.shape_info("2d5pt")
    coefs = [0.15, 0.2, 0.3, 0.2, 0.15]
    groups = [(coefs[g], [(0,) + tuple(o[1:]) for o in si.offsets[g]], 0) for g in range(5)]
    u = O.uniform(23 * 37, 9, -1.0, 1.0, np.float32).reshape(1, 23, 37)
    bits = O.to16(u, fmt)

    def rnd(a):
        if fmt == "f16":
            return a.astype(np.float16).astype(np.float32)
        return torch.from_numpy(a.copy()).to(torch.bfloat16).float().numpy()

    want = O.from16(bits, fmt)
    for _ in range(6):
        want = rnd(numpy_step(want, groups, "zero"))
    got = O.from16(O.iterate16(bits.reshape(23, 37), O.stencil(groups, "zero"), 6, fmt), fmt)
    assert np.array_equal(got.reshape(23, 37), want.reshape(23, 37))


# ---------------------------------------------------------------- committed reference golden vectors
GOLD = np.load(pathlib.Path(__file__).resolve().parent / "golden" / "ref_gold.npz")


@pytest.mark.parametrize("shape", sorted(REF_3D) + sorted(REF_2D))
@pytest.mark.parametrize("tag", ["f64", "f32"])
def test_oracle_matches_committed_reference_vectors(shape, tag):
    """Outputs of the reference's gold on glibc-rand() inputs, written by
    tests/golden/make_golden.py; holds without oracle/_ref."""
    field = GOLD[f"{shape}_{tag}_in"]
    steps = int(GOLD["steps_3d"] if field.ndim == 3 else GOLD["steps_2d"])
    got = O.iterate(field, O.stencil_from_shape(P.shape_info(shape), "clamp"), steps)
    assert np.array_equal(got.view(np.uint8), GOLD[f"{shape}_{tag}_out"].view(np.uint8))


@pytest.mark.parametrize("tag", ["f64", "f32"])
def test_oracle_matches_committed_interior_gold(tag):
    f = GOLD[f"j3d_gold4_{tag}_in"]
    got = O.iterate(np.ascontiguousarray(f[1:-1, 1:-1, 1:-1]), O.stencil_from_shape(P.shape_info("3d7pt"), "zero"), 4)
    assert np.array_equal(got, GOLD[f"j3d_gold4_{tag}_out"][1:-1, 1:-1, 1:-1])


# ---------------------------------------------------------------- stochastic rounding of 16-bit storage
def test_lowbias32_is_a_bijection_with_full_avalanche():
    f = O.lib().po_lowbias32
    xs = np.arange(1 << 16, dtype=np.uint32) * np.uint32(40503)
    hs = np.array([f(int(x)) for x in xs], dtype=np.uint32)
    assert len(np.unique(hs)) == len(hs)
    assert f(0) == 0
    flips = [bin(f(int(x)) ^ f(int(x) ^ (1 << b))).count("1") for x in xs[:200] for b in range(32)]
    assert 15.0 < np.mean(flips) < 17.0
    # the per-step key changes every bit position about half the time
    keys = np.array([O.lib().po_sr_bits(0, s, 7) for s in range(4096)], dtype=np.uint32)
    bits = (keys[:, None] >> np.arange(32, dtype=np.uint32)) & 1
    assert np.all(np.abs(bits.mean(axis=0) - 0.5) < 0.05)


def _sr_mean(v, fmt):
    """Mean of the stochastic rounding of v over every value of the random bits it uses."""
    f = O.lib().po_f32_to_bf16_sr if fmt == "bf16" else O.lib().po_f32_to_f16_sr
    nr = 1 << (16 if fmt == "bf16" else 13)
    q = np.array([f(float(v), r) for r in range(nr)], dtype=np.uint16)
    return O.from16(q, fmt).astype(np.float64), q


@pytest.mark.parametrize("fmt", ["bf16", "f16"])
def test_stochastic_rounding_is_unbiased_and_brackets(fmt):
    rng = np.random.default_rng(3)
    vals = np.concatenate([rng.uniform(-2, 2, 12), rng.uniform(1e-3, 1e-2, 4), [0.1, -0.3, 1.0 / 3.0]]).astype(np.float32)
    for v in vals:
        w, q = _sr_mean(v, fmt)
        # exactly unbiased: the mean over all random values is v itself
        assert abs(w.mean() - float(v)) <= 1e-12 * max(1.0, abs(float(v))), (fmt, v)
        # only the two neighbours of v in the format are produced
        lo, hi = w.min(), w.max()
        assert lo <= float(v) <= hi
        assert len(np.unique(q)) <= 2
    # representable values never move
    for v in (0.5, -0.75, 1.0, 0.0, 2.0 ** -10):
        w, _ = _sr_mean(np.float32(v), fmt)
        assert np.all(w == v)


def test_stochastic_rounding_specials():
    f16, bf = O.lib().po_f32_to_f16_sr, O.lib().po_f32_to_bf16_sr
    for r in (0, 1, 0x1FFF, 0xFFFF, 0xFFFFFFFF):
        assert O.from16(np.array([bf(float("inf"), r)], np.uint16), "bf16")[0] == np.inf
        assert np.isnan(O.from16(np.array([bf(float("nan"), r)], np.uint16), "bf16")[0])
        assert O.from16(np.array([f16(float("-inf"), r)], np.uint16), "f16")[0] == -np.inf
        # FP16 subnormal results round to nearest even whatever the bits
        v = np.float32(3.3e-6)
        assert f16(float(v), r) == O.lib().po_f32_to_f16(float(v))
    # just above the largest FP16 value: 65504 or inf, in proportion
    v = np.float32(65504.0 + 8.0)
    w, _ = _sr_mean(v, "f16")
    assert set(np.unique(w)) <= {65504.0, np.inf}
    assert abs(np.mean(w == np.inf) - 0.25) < 1e-3


def _numpy_lowbias32(x):
    x = x.astype(np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7feb352d) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846ca68b) & 0xFFFFFFFF
    x ^= x >> 16
    return x.astype(np.uint32)


@pytest.mark.parametrize("fmt", ["bf16", "f16"])
def test_stochastic_iteration_matches_numpy_restatement(fmt):
    """An independent numpy statement of the documented rounding rule (array
    bit operations, its own hash) against the oracle's C loop."""
    torch = pytest.importorskip("torch")
    si = P.shape_info("3d7pt")
    groups = [(CONFIG_3D[g], si.offsets[g], 0) for g in range(7)]
    dims = (5, 6, 9)
    seed = 1234
    u = O.uniform(int(np.prod(dims)), 5, 0.0, 1.0, np.float32).reshape(dims)
    bits = O.to16(u, fmt)
    cur = O.from16(bits, fmt).reshape(dims)
    idx = np.arange(int(np.prod(dims)), dtype=np.uint32).reshape(dims)
    for s in range(4):
        v = numpy_step(cur, groups, "zero").astype(np.float32)
        key = _numpy_lowbias32(np.array([seed ^ ((s * 0x9E3779B9) & 0xFFFFFFFF)], dtype=np.uint64))[0]
        r = _numpy_lowbias32(idx ^ key)
        ub = v.view(np.uint32).astype(np.uint64)
        if fmt == "bf16":
            q = ((ub + (r & 0xFFFF)) >> 16).astype(np.uint16)
            cur = O.from16(q, "bf16").reshape(dims)
        else:
            t = ((ub + (r & 0x1FFF)) & 0xFFFFE000).astype(np.uint32).view(np.float32)
            e = (ub >> 23) & 0xFF
            normal = (e >= 113) & (e < 255)
            with np.errstate(over="ignore"):
                rounded = np.where(normal, t, v).astype(np.float16)
            cur = rounded.astype(np.float32)
    got = O.from16(O.iterate16_sr(bits, O.stencil(groups, "zero"), 4, fmt, seed), fmt).reshape(dims)
    assert np.array_equal(got, cur)


def test_stochastic_rounding_removes_the_stall():
    """Nearest rounding loses every change below half an ulp, so a decaying
    BF16 field stalls; stochastic rounding follows the FP64 solution."""
    si = P.shape_info("3d7pt")
    dims, steps = (16, 16, 16), 300
    st = O.stencil_from_shape(si, "zero", coefs=CONFIG_3D)
    u = O.uniform(int(np.prod(dims)), 2, 0.0, 1.0).reshape(dims)
    ref = O.iterate(u, st, steps)
    bits = O.to16(u.astype(np.float32), "bf16")
    rn = O.from16(O.iterate16(bits, st, steps, "bf16"), "bf16").reshape(dims)
    sr = O.from16(O.iterate16_sr(bits, st, steps, "bf16", 99), "bf16").reshape(dims)
    rel = lambda x: np.sqrt(np.mean((x - ref) ** 2)) / np.sqrt(np.mean(ref ** 2))
    assert rel(sr) < 1e-2 < rel(rn)
