"""Host-side mirror of the reference's Jacobi entry points over the C ABI.

``j3d_iterative`` keeps the argument names, meaning and return codes of the
reference's ``template<class REAL> int j3d_iterative(REAL* h_input, int height,
int width_y, int width_x, REAL* h_output, int bdimx, int blkpsm, int iteration,
bool useSM, bool usewarmup, int warmupiteration, bool isDoubleTile, bool
getminHeight=false)`` (stencil/3dstencil/common/jacobi_cuda.cuh:7-8,
stencil/3dstencil/j3d.cu:129-140): host arrays in and out, 0 on success, the
planner's negative codes on refusal.  The stencil shape, boundary rule, method
and coefficients, which the reference fixes at compile time, are keyword
arguments here.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L

BOUNDARIES = {"clamp": L.BC_CLAMP, "zero": L.BC_ZERO}
METHODS = {"plain": L.METHOD_PLAIN, "persistent": L.METHOD_PERSISTENT}


@dataclass
class ShapeInfo:
    name: str
    id: int
    dims: int
    radius: int
    coef: list
    divisor: float
    npts: list
    prod_f64: list
    offsets: list  # per group: list of (dz, dy, dx)


def shape_id(shape) -> int:
    if isinstance(shape, int):
        return shape
    if shape in L.SHAPES:
        return L.SHAPES[shape]
    if shape in L.REF_DIR_SHAPE:
        return L.SHAPES[L.REF_DIR_SHAPE[shape]]
    raise ValueError(f"unknown shape {shape!r}")


def shape_info(shape) -> ShapeInfo:
    lib = L.load()
    sid = shape_id(shape)
    dims, radius, ng = C.c_int32(), C.c_int32(), C.c_int32()
    coef = (C.c_double * L.MAX_COEFS)()
    div = C.c_double()
    npts = (C.c_int32 * L.MAX_COEFS)()
    pf = (C.c_int32 * L.MAX_COEFS)()
    offs = (C.c_int8 * (L.MAX_COEFS * L.MAX_PTS * 3))()
    L.check(lib.perks_shape_info(sid, C.byref(dims), C.byref(radius), C.byref(ng), coef, C.byref(div), npts, pf,
                                 offs), "perks_shape_info")
    n = ng.value
    offsets = []
    for g in range(n):
        pts = []
        for p in range(npts[g]):
            b = (g * L.MAX_PTS + p) * 3
            pts.append((offs[b], offs[b + 1], offs[b + 2]))
        offsets.append(pts)
    name = next((k for k, v in L.SHAPES.items() if v == sid), str(sid))
    return ShapeInfo(name, sid, dims.value, radius.value, list(coef[:n]), div.value, list(npts[:n]), list(pf[:n]),
                     offsets)


ROUNDING = {"nearest": 0, "stochastic": 1}


def make_opts(shape="3d_star_r1", boundary="clamp", method="persistent", coefs=None, divisor=None,
              bdimx=0, blkpsm=0, use_lds_cache=0, reg_cache_planes=-1, double_tile=0, warmup=0,
              warmup_steps=0, tune=None, rounding="nearest", round_seed=0) -> L.StencilOpts:
    lib = L.load()
    o = L.StencilOpts()
    lib.perks_stencil_opts_init(C.byref(o), shape_id(shape))
    o.boundary = BOUNDARIES[boundary] if isinstance(boundary, str) else int(boundary)
    o.method = METHODS[method] if isinstance(method, str) else int(method)
    if coefs is not None:
        o.use_ref_coefs = 0
        for i, c in enumerate(coefs):
            o.coef[i] = float(c)
        o.divisor = 0.0 if divisor is None else float(divisor)
    o.bdimx, o.blkpsm, o.use_lds_cache = int(bdimx), int(blkpsm), int(use_lds_cache)
    o.reg_cache_planes, o.double_tile = int(reg_cache_planes), int(double_tile)
    o.warmup, o.warmup_steps = int(warmup), int(warmup_steps)
    if tune:
        for i, v in enumerate(tune):
            o.tune[i] = int(v)
    o.rounding = ROUNDING[rounding] if isinstance(rounding, str) else int(rounding)
    o.round_seed = int(round_seed) & 0xFFFFFFFF
    return o


def _dtype_fn(arr: np.ndarray, kind: str):
    lib = L.load()
    if arr.dtype == np.float64:
        return getattr(lib, f"perks_{kind}_f64")
    if arr.dtype == np.float32:
        return getattr(lib, f"perks_{kind}_f32")
    raise TypeError(f"unsupported dtype {arr.dtype}")


def j3d_iterative(h_input, height, width_y, width_x, h_output, bdimx=256, blkpsm=0, iteration=3, useSM=False,
                  usewarmup=False, warmupiteration=-1, isDoubleTile=False, getminHeight=False, *,
                  shape="3d_star_r1", boundary="clamp", method="persistent", coefs=None, divisor=None,
                  report: L.RunReport | None = None, **tuning) -> int:
    """Runs `iteration` Jacobi steps of h_input[height][width_y][width_x] into
    h_output.  Returns 0, a negative planner code, or (getminHeight) the
    minimum height that caches fully -- as the reference does."""
    opts = make_opts(shape, boundary, method, coefs, divisor, bdimx=bdimx, blkpsm=blkpsm,
                     use_lds_cache=1 if useSM else -1, double_tile=int(bool(isDoubleTile)),
                     warmup=int(bool(usewarmup)), warmup_steps=max(0, int(warmupiteration)), **tuning)
    lib = L.load()
    if getminHeight:
        dt = L.F64 if np.dtype(h_input.dtype if h_input is not None else np.float64) == np.float64 else L.F32
        return lib.perks_j3d_min_height(dt, width_y, width_x, C.byref(opts))
    for a in (h_input, h_output):
        if not (isinstance(a, np.ndarray) and a.flags.c_contiguous and a.size == height * width_y * width_x):
            raise ValueError("h_input/h_output must be C-contiguous arrays of height*width_y*width_x elements")
    if h_input.dtype != h_output.dtype:
        raise TypeError("h_input and h_output must share a dtype")
    rep = report if report is not None else L.RunReport()
    fn = _dtype_fn(h_input, "j3d_iterative")
    return fn(h_input.ctypes.data, height, width_y, width_x, h_output.ctypes.data, int(iteration), C.byref(opts),
              C.byref(rep))


def j3d_run(field: np.ndarray, steps: int, **kw) -> tuple[np.ndarray, L.RunReport]:
    """Convenience form: returns (result, report); raises on any error code."""
    field = np.ascontiguousarray(field)
    nz, ny, nx = field.shape
    out = np.empty_like(field)
    rep = L.RunReport()
    rc = j3d_iterative(field, nz, ny, nx, out, iteration=steps, report=rep, **kw)
    L.check(rc, "j3d_iterative")
    return out, rep


def _ptr(t) -> int:
    if isinstance(t, int):
        return t
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    raise TypeError("expected a device pointer (int) or a torch tensor")


def j3d_device(d_a, d_b, height, width_y, width_x, steps, dtype="f64", stream=None, opts=None,
               **kw) -> L.RunReport:
    """Runs on device-resident buffers (torch tensors or raw pointers) of
    f64 / f32 / f16 / bf16 storage (16-bit: raw bit patterns, FP32
    arithmetic).  The result is in d_b when report.result_in_b, else in d_a."""
    lib = L.load()
    o = opts if opts is not None else make_opts(**kw)
    rep = L.RunReport()
    s = None if stream is None else _ptr_stream(stream)
    L.check(lib.perks_j3d_device(_dt(dtype), _ptr(d_a), _ptr(d_b), height, width_y, width_x, int(steps), C.byref(o),
                                 s, C.byref(rep)), "j3d_device")
    return rep


def _config_rows(table, n) -> list[dict]:
    rows = []
    for i in range(n):
        t = table[i]
        rows.append({"method": "plain" if t.method == L.METHOD_PLAIN else "persistent", "variant": t.variant,
                     "levels": t.levels, "block": t.block, "vgprs": t.vgprs, "spill_bytes": t.spill_bytes,
                     "lds_bytes": t.lds_bytes, "occupancy": t.occupancy, "tile": (t.tile_x, t.tile_y),
                     "status": t.status, "gcells_per_s": t.gcells_per_s})
    return rows


def j3d_kernel_table(height, width_y, width_x, dtype="f64", opts=None, **kw) -> list[dict]:
    """Every launch configuration of the 3D path for this problem, with what
    its code object reports (VGPRs, spill bytes, LDS, occupancy) and its
    status (0, or the error code saying why it cannot run)."""
    lib = L.load()
    o = opts if opts is not None else make_opts(**kw)
    n = C.c_int(0)
    L.check(lib.perks_j3d_kernel_table(_dt(dtype), height, width_y, width_x, C.byref(o), None, 0, C.byref(n)),
            "j3d_kernel_table")
    table = (L.KernelConfig * max(1, n.value))()
    L.check(lib.perks_j3d_kernel_table(_dt(dtype), height, width_y, width_x, C.byref(o), table, n.value,
                                       C.byref(n)), "j3d_kernel_table")
    return _config_rows(table, n.value)


def j3d_autotune(d_in, height, width_y, width_x, dtype="f64", sample_steps=0, stream=None, opts=None,
                 **kw) -> tuple[L.StencilOpts, list[dict]]:
    """Times every runnable configuration on the field d_in (left unchanged)
    and returns the options with the fastest one selected (method, tune[2],
    tune[4]) and the table with the measured rates."""
    lib = L.load()
    o = opts if opts is not None else make_opts(**kw)
    n = C.c_int(0)
    L.check(lib.perks_j3d_kernel_table(_dt(dtype), height, width_y, width_x, C.byref(o), None, 0, C.byref(n)),
            "j3d_kernel_table")
    table = (L.KernelConfig * max(1, n.value))()
    s = None if stream is None else _ptr_stream(stream)
    L.check(lib.perks_j3d_autotune(_dt(dtype), _ptr(d_in), height, width_y, width_x, int(sample_steps),
                                   C.byref(o), table, n.value, C.byref(n), s), "j3d_autotune")
    return o, _config_rows(table, n.value)


def _ptr_stream(stream) -> int:
    if isinstance(stream, int):
        return stream
    if hasattr(stream, "cuda_stream"):
        return stream.cuda_stream
    raise TypeError("expected a hipStream_t (int) or a torch stream")


DTYPES = {"f64": L.F64, "f32": L.F32, "f16": L.F16, "bf16": L.BF16}


def _dt(dtype) -> int:
    return DTYPES[dtype] if isinstance(dtype, str) else int(dtype)


def fill_uniform(d, n, seed, lo, hi, dtype="f64", stream=None) -> None:
    lib = L.load()
    fn = {L.F64: lib.perks_fill_uniform_f64, L.F32: lib.perks_fill_uniform_f32, L.F16: lib.perks_fill_uniform_f16,
          L.BF16: lib.perks_fill_uniform_bf16}[_dt(dtype)]
    s = None if stream is None else _ptr_stream(stream)
    L.check(fn(_ptr(d), int(n), int(seed), float(lo), float(hi), s), "fill_uniform")


def query_device(device: int = 0) -> L.DeviceCaps:
    caps = L.DeviceCaps()
    L.check(L.load().perks_query_device(device, C.byref(caps)), "query_device")
    return caps


def plan(height, width_y, width_x, dtype="f64", caps: L.DeviceCaps | None = None, occupancy=0,
         **kw) -> tuple[int, L.PlanInfo]:
    """Pure host planning when `caps` is given.  Returns (code, plan)."""
    o = make_opts(**kw)
    info = L.PlanInfo()
    dt = _dt(dtype)
    rc = L.load().perks_j3d_plan(dt, height, width_y, width_x, C.byref(o), C.byref(caps) if caps else None,
                                 int(occupancy), C.byref(info))
    return rc, info


# ---- 2D ---------------------------------------------------------------------------
def jacobi_iterative(h_input, width_y, width_x, h_output, bdimx=256, blkpsm=0, iteration=3, async_=False,
                     useSM=False, usewarmup=False, warmupiteration=-1, isDoubleTile=False, *, shape="2d_star_r1",
                     boundary="clamp", method="persistent", coefs=None, divisor=None,
                     report: L.RunReport | None = None, **tuning) -> int:
    """The reference's jacobi_iterative<REAL> (stencil/2dstencil/common/jacobi_cuda.cuh:6-10,
    jacobi-2d.cu:369-374): h_input[width_y][width_x] -> h_output after
    `iteration` steps; returns 0 or a negative planner code."""
    opts = make_opts(shape, boundary, method, coefs, divisor, bdimx=bdimx, blkpsm=blkpsm,
                     use_lds_cache=1 if useSM else -1, double_tile=int(bool(isDoubleTile)),
                     warmup=int(bool(usewarmup)), warmup_steps=max(0, int(warmupiteration)), **tuning)
    for a in (h_input, h_output):
        if not (isinstance(a, np.ndarray) and a.flags.c_contiguous and a.size == width_y * width_x):
            raise ValueError("h_input/h_output must be C-contiguous arrays of width_y*width_x elements")
    if h_input.dtype != h_output.dtype:
        raise TypeError("h_input and h_output must share a dtype")
    rep = report if report is not None else L.RunReport()
    fn = _dtype_fn(h_input, "jacobi2d_iterative")
    return fn(h_input.ctypes.data, width_y, width_x, h_output.ctypes.data, int(iteration), C.byref(opts),
              C.byref(rep))


def jacobi2d_run(field: np.ndarray, steps: int, **kw) -> tuple[np.ndarray, L.RunReport]:
    field = np.ascontiguousarray(field)
    ny, nx = field.shape
    out = np.empty_like(field)
    rep = L.RunReport()
    L.check(jacobi_iterative(field, ny, nx, out, iteration=steps, report=rep, **kw), "jacobi_iterative")
    return out, rep


def jacobi2d_device(d_a, d_b, width_y, width_x, steps, dtype="f64", stream=None, opts=None, **kw) -> L.RunReport:
    """Device buffers (torch tensors or pointers) of f64/f32/f16/bf16 storage;
    the result is in d_b when report.result_in_b, else in d_a."""
    o = opts if opts is not None else make_opts(**kw)
    rep = L.RunReport()
    s = None if stream is None else _ptr_stream(stream)
    L.check(L.load().perks_jacobi2d_device(_dt(dtype), _ptr(d_a), _ptr(d_b), width_y, width_x, int(steps), C.byref(o),
                                           s, C.byref(rep)), "jacobi2d_device")
    return rep


class BatchPlan:
    """A captured, optionally autotuned launch of a fixed batch of 2D grids
    (perks_jacobi2d_batch_plan).  run() replays it; close() frees it."""

    def __init__(self, grids, dtype="f32", autotune=True, stream=None, opts=None, **kw):
        o = opts if opts is not None else make_opts(**kw)
        arr = (L.Grid2d * len(grids))()
        for i, (a, b, wy, wx, st) in enumerate(grids):
            arr[i].a, arr[i].b = _ptr(a), _ptr(b)
            arr[i].width_y, arr[i].width_x, arr[i].steps = int(wy), int(wx), int(st)
        self._lib = L.load()
        self._h = C.c_void_p()
        self.n = len(grids)
        s = None if stream is None else _ptr_stream(stream)
        L.check(self._lib.perks_jacobi2d_batch_plan(_dt(dtype), arr, len(grids), C.byref(o), int(bool(autotune)), s,
                                                    C.byref(self._h)), "jacobi2d_batch_plan")
        flags = (C.c_int32 * self.n)()
        cfg = (C.c_int32 * 6)()
        L.check(self._lib.perks_jacobi2d_batch_plan_info(self._h, flags, self.n, cfg), "batch_plan_info")
        self.result_in_b = list(flags)
        self.config = dict(zip(("tile", "steps_per_pass", "wg_per_cu", "workgroups", "autotuned", "passes"), cfg))

    def run(self, stream=None) -> L.RunReport:
        rep = L.RunReport()
        s = None if stream is None else _ptr_stream(stream)
        L.check(self._lib.perks_jacobi2d_batch_run(self._h, s, C.byref(rep)), "jacobi2d_batch_run")
        return rep

    def close(self):
        if self._h:
            self._lib.perks_jacobi2d_batch_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def jacobi2d_batch(grids, dtype="f32", stream=None, opts=None, **kw):
    """grids: list of (d_a, d_b, width_y, width_x, steps).  Returns (report,
    list of result_in_b flags)."""
    o = opts if opts is not None else make_opts(**kw)
    arr = (L.Grid2d * len(grids))()
    for i, (a, b, wy, wx, st) in enumerate(grids):
        arr[i].a, arr[i].b = _ptr(a), _ptr(b)
        arr[i].width_y, arr[i].width_x, arr[i].steps = int(wy), int(wx), int(st)
    rep = L.RunReport()
    s = None if stream is None else _ptr_stream(stream)
    L.check(L.load().perks_jacobi2d_batch(_dt(dtype), arr, len(grids), C.byref(o), s, C.byref(rep)),
            "jacobi2d_batch")
    return rep, [arr[i].result_in_b for i in range(len(grids))]
