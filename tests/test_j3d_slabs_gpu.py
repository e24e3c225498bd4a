"""z-slabs per tile of the persistent 7-point kernel (opts.tune[7]): any slab
count, up to one slab per plane, gives the planner's bits and the oracle's;
a negative count is rejected."""
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P
from perks_amd import _lib as L

pytestmark = pytest.mark.gpu

CONFIG_3D = [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("boundary", ["zero", "clamp"])
def test_slab_counts_match_oracle(gpu, dtype, boundary):
    dims, steps = (23, 45, 70), 9
    field = O.uniform(int(np.prod(dims)), 17, 0.0, 1.0, dtype=dtype).reshape(dims)
    want = O.iterate(field, O.stencil_from_shape(P.shape_info("3d_star_r1"), boundary, coefs=CONFIG_3D), steps)
    for nzs in (0, 1, 2, 3, 5, 8, 22, 23, 40):
        for variant in (0, 2):
            say(f"[slabs] {np.dtype(dtype).name} {boundary} nzs={nzs} variant={variant}")
            got, _ = P.j3d_run(field, steps, method="persistent", boundary=boundary, coefs=CONFIG_3D,
                               tune=[0, 0, 0, 0, variant, 0, 0, nzs])
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), f"nzs={nzs} variant={variant}"


def test_negative_slab_count_rejected(gpu):
    with pytest.raises(L.PerksError) as e:
        P.j3d_run(np.zeros((4, 8, 8)), 3, method="persistent", boundary="zero", coefs=CONFIG_3D,
                  tune=[0, 0, 0, 0, 0, 0, 0, -1])
    assert e.value.code == L.ERR_INVALID_ARG
