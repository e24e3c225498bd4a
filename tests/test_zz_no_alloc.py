"""perks_jacobi2d_batch_noalloc through the C ABI: the batch entry point whose
scratch (barrier block, work list, per-pass counts) lives in a fixed device
array instead of library allocations.  Its results equal perks_jacobi2d_batch's
bit for bit, on two calls in a row (the second smaller, so the array holds the
first call's contents when it starts), by both methods; sizes beyond the
array's capacity are refused before any GPU work; repeated calls leave the
device's free memory unchanged."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O
import perks_amd as P
from perks_amd import _lib as L

pytestmark = pytest.mark.gpu

MAX_GRIDS, MAX_PASSES = 4096, 16384  # PERKS_NOALLOC_MAX_GRIDS / _PASSES in perks_amd.h
F32 = 1  # PERKS_F32


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def grids_array(bufs, specs):
    arr = (L.Grid2d * len(specs))()
    for i, ((a, b), (wy, wx, st)) in enumerate(zip(bufs, specs)):
        arr[i].a, arr[i].b = a.data_ptr(), b.data_ptr()
        arr[i].width_y, arr[i].width_x, arr[i].steps = wy, wx, st
    return arr


def run(torch, fn, fields, specs, opts):
    bufs = [(torch.from_numpy(f.copy()).view(-1).cuda(), torch.full((f.size,), -1.0, dtype=torch.float32,
                                                                      device="cuda")) for f in fields]
    torch.cuda.synchronize()
    arr = grids_array(bufs, specs)
    rep = L.RunReport()
    rc = getattr(L.load(), fn)(F32, arr, len(specs), C.byref(opts), None, C.byref(rep))
    assert rc == 0, f"{fn} returned {rc}"
    out = [(b if arr[i].result_in_b else a).cpu().numpy().reshape(wy, wx) for i, ((a, b), (wy, wx, _))
           in enumerate(zip(bufs, specs))]
    return out, [arr[i].result_in_b for i in range(len(specs))]


def batch(seed, nb, max_side, max_steps):
    rng = np.random.default_rng(seed)
    specs = [(int(rng.integers(1, max_side)), int(rng.integers(1, max_side)), int(rng.integers(0, max_steps)))
             for _ in range(nb)]
    fields = [O.uniform(wy * wx, 900 + seed + i, 0.0, 1.0, np.float32).reshape(wy, wx)
              for i, (wy, wx, _) in enumerate(specs)]
    return specs, fields


@pytest.mark.parametrize("method", ["persistent", "plain"])
def test_noalloc_equals_batch_on_two_calls(gpu, method):
    torch = gpu
    opts = P.make_opts("2d9pt", "zero", method, [0.1, 0.1, 0.1, 0.1, 0.2, 0.1, 0.1, 0.1, 0.1])
    for call, (nb, side, steps) in enumerate([(24, 300, 40), (5, 90, 13)]):
        specs, fields = batch(17 * call + (method == "plain"), nb, side, steps)
        say(f"[noalloc] {method} call {call}: {len(specs)} grids, max steps {max(s for *_, s in specs)}")
        want, fw = run(torch, "perks_jacobi2d_batch", fields, specs, opts)
        got, fg = run(torch, "perks_jacobi2d_batch_noalloc", fields, specs, opts)
        assert fw == fg
        for i, (g, w) in enumerate(zip(got, want)):
            assert np.array_equal(g.view(np.uint32), w.view(np.uint32)), f"call {call} grid {i} {specs[i]}"


def test_noalloc_refuses_sizes_beyond_its_scratch(gpu):
    torch = gpu
    a = torch.zeros(16, dtype=torch.float32, device="cuda")
    b = torch.zeros(16, dtype=torch.float32, device="cuda")
    lib = L.load()
    opts = P.make_opts("2d5pt", "zero", "plain", [0.2] * 5)
    too_many = [((a, b), (4, 4, 1))] * (MAX_GRIDS + 1)
    arr = grids_array([x for x, _ in too_many], [s for _, s in too_many])
    rc = lib.perks_jacobi2d_batch_noalloc(F32, arr, MAX_GRIDS + 1, C.byref(opts), None, C.byref(L.RunReport()))
    assert rc == L.ERR_UNSUPPORTED_SIZE
    arr = grids_array([(a, b)], [(4, 4, MAX_PASSES + 1)])  # per-step method: one pass per time step
    rc = lib.perks_jacobi2d_batch_noalloc(F32, arr, 1, C.byref(opts), None, C.byref(L.RunReport()))
    assert rc == L.ERR_UNSUPPORTED_SIZE


def test_noalloc_calls_leave_free_memory_unchanged(gpu):
    torch = gpu
    opts = P.make_opts("2d5pt", "clamp", "persistent", [0.2] * 5)
    specs, fields = batch(5, 12, 200, 30)
    bufs = [(torch.from_numpy(f.copy()).view(-1).cuda(), torch.empty(f.size, dtype=torch.float32, device="cuda"))
            for f in fields]
    arr = grids_array(bufs, specs)
    lib = L.load()
    assert lib.perks_jacobi2d_batch_noalloc(F32, arr, len(specs), C.byref(opts), None, C.byref(L.RunReport())) == 0
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(3):
        assert lib.perks_jacobi2d_batch_noalloc(F32, arr, len(specs), C.byref(opts), None, C.byref(L.RunReport())) == 0
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
