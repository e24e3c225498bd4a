"""The oracle's 7-point loop (po_star7_iterate_omp_f64, the CPU baseline's
all-cores figure) against the oracle's shape restatement, which the other
oracle tests pin to the reference gold: bit for bit, both boundaries, edge
widths 1-3, any thread count, odd and even step counts."""
import numpy as np
import pytest

import oracle as O
import perks_amd as P

CONFIG_3D = [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4]


@pytest.mark.parametrize("boundary", ["zero", "clamp"])
@pytest.mark.parametrize("dims", [(5, 7, 9), (1, 1, 1), (3, 2, 2), (4, 3, 3), (6, 5, 1), (2, 1, 17)])
def test_star7_loop_equals_restatement(boundary, dims):
    rng = np.random.default_rng(sum(dims))
    field = rng.random(dims)
    for coefs in (CONFIG_3D, list(rng.uniform(-1.0, 1.0, 7))):
        st = O.stencil_from_shape(P.shape_info("3d_star_r1"), boundary, coefs=coefs, divisor=0.0)
        for steps in (0, 1, 2, 5):
            want = O.iterate(field, st, steps)
            for threads in (1, 3):
                got = O.star7_iterate(field, coefs, boundary, steps, threads)
                assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (coefs, steps, threads)


def test_star7_loop_is_not_the_contracted_form():
    """A planted contraction (the fused restatement) differs from the loop, so
    the bitwise test above can tell the two apart."""
    field = np.random.default_rng(3).random((6, 9, 11))
    st = O.stencil_from_shape(P.shape_info("3d_star_r1"), "zero", coefs=CONFIG_3D, divisor=0.0)
    assert not np.array_equal(O.star7_iterate(field, CONFIG_3D, "zero", 4), O.iterate_fused(field, st, 4))
