#!/usr/bin/env python3
"""Benchmark of the persistent 3D 7-point Jacobi solve on one MI355X.

Workload (one bench "step" = one complete solve, inputs already in HBM):
  3D 7-point star Jacobi, FP64, 512^3 interior with a one-cell zero halo,
  centre coefficient 0.4 and six neighbours 0.1, 1000 time steps, input
  uniform[0,1) from a seed.  The persistent method runs all 1000 time steps
  in one cooperative launch; the per-step method (one launch per time step)
  is timed on the same GPU and buffers for comparison, and the two results
  are compared bit for bit.  config.contracted times the same solve with the
  contracted update (opts.tune[6] = 1: products fused into the running sum,
  within 4 ulp of the gold per step, not bit-identical) and reports its
  deviation from the bit-identical result; `value` is the bit-identical form.

Prints one JSON line (see the keys at the bottom).  GCells/s counts
time steps * cells / seconds, the reference's own definition
(stencil/3dstencil/j3d.cu:754).  The roofline prices the algorithmic HBM
bytes of the kernel's pass structure against 8 TB/s: a kernel that runs L
time steps per pass must read and write the field once per pass, so a launch
of T steps moves 2 * sizeof(T) * cells * ceil(T / L) bytes (L = 1 for the
per-step kernel: the reference's single-pass model, j3d.cu:756).  The
single-pass figure is reported beside it as roofline.single_pass_gbs.  The
counter-measured traffic (FETCH_SIZE + WRITE_SIZE) and the VALU figures come
from profiles/<kernel>_traffic.json, used only when the kernel sources it
was recorded from are the ones in this tree (SHA-256 over KERNEL_SOURCES).

    python bench.py --record-profile DIR_FETCH DIR_WRITE [DIR_SQ]
reads the rocprofv3 counter CSVs of the default workload (one directory per
counter pass) and writes that JSON.

Other workloads (--workload; each prints its own JSON line):
  j3d-storage  the same 3D solve with FP64 / FP32 / FP16 / BF16 storage (FP32
               arithmetic for the 16-bit forms), persistent and per-step
               kernels, with max and RMS error against the FP64 result.
  j3d-shapes   every 3D shape (7/13/17/27-point, Poisson) at 512^3, FP64 and
               FP32, persistent and per-step kernels.
  j2d-storage  2D box r1 (2d9pt) and r2 (2d25pt), 8192^2, 500 steps, storage
               FP32 / FP16 / BF16 with FP32 arithmetic, against FP64 on the
               same inputs (max and RMS error at 10 / 100 / 500 steps).
  batch2d      512 independent 2D grids (sides log-uniform in [64,1024],
               steps uniform in [50,500], FP32) in one persistent launch,
               against solving them one after another with the per-step kernel.
  cg           CG on the 7-point Poisson matrix, 256^3, CSR and DIA, FP64 and
               FP32-inner/FP64-refinement, to ||r||/||b|| < 1e-10.
  j3d-small    the reference's --small / --checkmindomain search (j3d.driver.cpp:62-156):
               getminHeight for 512 x 512 and 256 x 256 planes, FP64 and FP32,
               then the 7-point solve at that height, 1000 steps, with the
               resident-tile kernel, the temporal-blocking kernel and the
               per-step kernel, compared bit for bit.
  j2d-domain   the reference's domain selection for the 2D experiments (iteratedomain.sh /
               iteratemindomain.sh / ProcessDomaindata.py:20-63): the per-step 2d5pt kernel
               swept over square domains, FP32 and FP64; the chosen domain is the smallest
               one larger than both getMinWidthY's domain and half the chip's L2 whose
               per-step rate is within 85% of the sweep's peak; there the persistent and
               per-step methods are timed and compared bit for bit.
  j3d-tune     the configuration table of the 3D path at 512^3 (VGPRs, spills,
               LDS, occupancy of every method / tile variant / pass depth) and
               the autotuner's timing of each, for the 7-point star in every
               storage format and the 27-point box in FP64 / FP32.
"""
from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "GCells/s + % of HBM roofline, 3d7pt fp64 512³, persistent vs per-step kernel"
HBM_PEAK_GBS = 8000.0
CONFIG_COEFS = [0.1, 0.1, 0.1, 0.1, 0.1, 0.1, 0.4]  # E W N S T B, centre
TRAFFIC_FILE = ROOT / "profiles" / "j3d_persistent_f64_512_traffic.json"
TRAFFIC_FILE_FUSED = ROOT / "profiles" / "j3d_persistent_fused_f64_512_traffic.json"  # --contracted
PERSISTENT_KERNEL = "j3d_star1_persistent_kernel"
# Files that define the persistent 7-point kernel and its launch: a recorded
# profile applies only while their contents are unchanged.
KERNEL_SOURCES = ["perks_amd/csrc/j3d_persistent_kernel.h", "perks_amd/csrc/j3d_persistent.hip",
                  "perks_amd/csrc/j3d_persistent_f64.hip", "perks_amd/csrc/j3d_persistent.h",
                  "perks_amd/csrc/grid_sync.h", "perks_amd/csrc/perks_device.h", "perks_amd/csrc/j3d_common.h",
                  "perks_amd/csrc/storage16.h", "perks_amd/csrc/storage_fmt.h", "perks_amd/build.py"]
SIMDS, SHADER_ENGINES = 1024, 32  # MI355X: 256 CUs x 4 SIMDs; 8 XCDs x 4 SEs


def kernel_sources_sha() -> str:
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        h.update(rel.encode())
        h.update((ROOT / rel).read_bytes())
    return h.hexdigest()


def pass_roofline(kernel, cells, time_steps, steps_per_pass, esz, seconds, traffic=None):
    """Roofline of one launch: algorithmic bytes = one read and one write of
    the field per pass of `steps_per_pass` time steps.  `cells` and
    `time_steps` may be lists (a batch of grids)."""
    cl = cells if isinstance(cells, list) else [cells]
    ts = time_steps if isinstance(time_steps, list) else [time_steps] * len(cl)
    L = max(1, int(steps_per_pass))
    alg = sum(2 * esz * c * (-(-t // L)) for c, t in zip(cl, ts))
    single = sum(2 * esz * c * t for c, t in zip(cl, ts))
    achieved = alg / seconds / 1e9
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
           "algorithmic_bytes_per_launch": alg, "steps_per_pass": L,
           "algorithmic_model": "one read + one write of the field per pass of steps_per_pass time steps",
           "single_pass_gbs": round(single / seconds / 1e9, 1)}
    if traffic:
        out["traffic_gbs"] = round(traffic / seconds / 1e9, 1)
    return out


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed solves")
    ap.add_argument("--warmup", type=int, default=1, help="untimed solves")
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--tsteps", type=int, default=1000, help="Jacobi time steps per solve")
    ap.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-per-step", action="store_true", help="skip the per-step comparison")
    ap.add_argument("--no-contracted", action="store_true", help="skip the contracted-update (FMA) leg")
    ap.add_argument("--contracted", action="store_true",
                    help="time the contracted update (opts.tune[6] = 1) as the main leg (for its counter profile)")
    ap.add_argument("--cpu-steps", type=int, default=10, help="time steps of the CPU sample")
    ap.add_argument("--method", choices=["persistent", "plain"], default="persistent")
    ap.add_argument("--workload", choices=["j3d", "j3d-storage", "j3d-shapes", "j2d-storage", "batch2d", "cg",
                                                  "cg-counters", "j3d-tune", "j3d-small", "j2d-domain"],
                    default="j3d")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--record-profile", nargs="+", metavar="DIR",
                    help="write the counter profile of the default workload from rocprofv3 CSV directories: "
                         "FETCH_SIZE pass, WRITE_SIZE pass[, SQ pass]")
    return ap.parse_args()


def cpu_baseline(n: int, tsteps: int, dtype: str) -> dict:
    """The reference's CPU gold j3d_gold_iterative, built from its sources
    into oracle/_ref, single-threaded on a bounded sample of the same grid
    (same 7-point update per cell; the gold fixes the clamp boundary and its
    own coefficients).  Falls back to the oracle restatement when the
    reference build is absent."""
    sys.path.insert(0, str(ROOT / "tests"))
    import numpy as np
    import oracle as O

    np_dt = np.float64 if dtype == "f64" else np.float32
    field = O.uniform(n ** 3, 1, 0.0, 1.0, np_dt).reshape(n, n, n)
    if O.ref_available("3d7pt_gold_gen"):
        kind, what = "reference", "j3d_gold_iterative (stencil/3dstencil/3d7pt/3d7pt_gold_gen.cpp), g++ -O3 -ffp-contract=off"
        t0 = time.perf_counter()
        O.ref_gold_iterative("3d7pt_gold_gen", field, tsteps)
        dt = time.perf_counter() - t0
    else:
        import perks_amd as P
        kind, what = "port", "oracle po_stencil_iterate_f64 (restatement of the reference gold)"
        st = O.stencil_from_shape(P.shape_info("3d_star_r1"), "zero", coefs=CONFIG_COEFS)
        t0 = time.perf_counter()
        O.iterate(field, st, tsteps)
        dt = time.perf_counter() - t0
    out = {
        "value": round(tsteps * n ** 3 / dt / 1e9, 4),
        "unit": "GCells/s",
        "cores": 1,
        "kind": kind,
        "sample": f"{what}; {n}^3 {dtype}, {tsteps} time steps, {dt:.1f} s on {os.cpu_count()} host CPUs "
                  f"({_cpu_model()}), one thread used",
    }
    # The same sample on every host core: the gold's 7-point loop restated in
    # the oracle with OpenMP over planes (bit-identical to the gold), capped at
    # the 16 CPUs a one-GPU share of the host provides.
    threads = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()))
    if dtype == "f64":
        t0 = time.perf_counter()
        O.star7_iterate(field, CONFIG_COEFS, "zero", tsteps, threads)
        dtm = time.perf_counter() - t0
        what = "oracle po_star7_iterate_omp_f64 (the gold's loop, g++ -O3 -ffp-contract=off, OpenMP over planes)"
    else:
        import perks_amd as P
        st = O.stencil_from_shape(P.shape_info("3d_star_r1"), "zero", coefs=CONFIG_COEFS)
        t0 = time.perf_counter()
        O.iterate(field, st, tsteps, threads=threads)
        dtm = time.perf_counter() - t0
        what = "oracle po_stencil_iterate_omp_f32"
    out["all_cores"] = {"value": round(tsteps * n ** 3 / dtm / 1e9, 4), "unit": "GCells/s", "cores": threads,
                        "kind": "port", "sample": f"{what}, {threads} threads; {n}^3 {dtype}, "
                                                  f"{tsteps} time steps, {dtm:.2f} s"}
    return out


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def copy_bandwidth(torch, like) -> float:
    """GB/s of a device-to-device copy of a field-sized buffer (read + write
    bytes over the time), timed with HIP events on the current stream."""
    dst = torch.empty_like(like)
    dst.copy_(like)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    reps = 5
    for _ in range(reps):
        dst.copy_(like)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del dst
    return 2 * like.numel() * like.element_size() / (ms * 1e-3) / 1e9


def load_traffic(n, tsteps, dtype, contracted=False):
    """The recorded counter profile of the persistent kernel (of its contracted
    form), or (None, why)."""
    try:
        d = json.loads((TRAFFIC_FILE_FUSED if contracted else TRAFFIC_FILE).read_text())
    except (OSError, ValueError):
        return None, "no recorded profile"
    if d.get("size") != n or d.get("time_steps") != tsteps or d.get("dtype") != dtype:
        return None, "recorded profile is for another workload"
    if d.get("source_sha256") != kernel_sources_sha():
        return None, "recorded profile is for other kernel sources"
    return d, None


def run_j3d(args):
    import torch
    import perks_amd as P

    n, T = args.size, args.tsteps
    cells = n ** 3
    tdt = torch.float64 if args.dtype == "f64" else torch.float32
    esz = 8 if args.dtype == "f64" else 4
    a = torch.empty(cells, dtype=tdt, device="cuda")
    b = torch.empty(cells, dtype=tdt, device="cuda")
    init = torch.empty(cells, dtype=tdt, device="cuda")
    P.fill_uniform(init, cells, 1, 0.0, 1.0, dtype=args.dtype)
    opts = dict(boundary="zero", coefs=CONFIG_COEFS)
    if args.contracted:
        opts["tune"] = [0, 0, 0, 0, 0, 0, 1]

    def solve(method):
        a.copy_(init)
        torch.cuda.synchronize()
        rep = P.j3d_device(a, b, n, n, n, T, dtype=args.dtype, method=method, **opts)
        return rep, (b if rep.result_in_b else a)

    for _ in range(args.warmup):
        solve(args.method)
    torch.cuda.synchronize()
    launch_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rep, res = solve(args.method)
        launch_ms.append(rep.elapsed_ms)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_s = sum(launch_ms) / 1e3
    value = args.steps * T * cells / kern_s / 1e9
    persistent_result = res.clone()
    levels = rep.reserved[0]

    cfg = {
        "workload": f"3d7pt star Jacobi {args.dtype}, {n}^3 interior + zero halo, c=0.4 / 6x0.1, "
                    f"{T} time steps per solve, {args.method} kernel",
        "size": n, "time_steps": T, "method": args.method, "steps_per_pass": levels,
        "grid": rep.grid[0], "block": rep.block,
        "wall_s_timed": round(wall, 3), "parallelism": "single GPU",
    }
    if not args.no_per_step and args.method == "persistent":
        reps = []
        for _ in range(max(1, min(args.steps, 2))):
            prep, pres = solve("plain")
            reps.append(prep.elapsed_ms)
        plain = len(reps) * T * cells / (sum(reps) / 1e3) / 1e9
        cfg["per_step_kernel_gcells_per_s"] = round(plain, 2)
        cfg["speedup_vs_per_step"] = round(value / plain, 3)
        cfg["bitwise_equal_to_per_step"] = bool(torch.equal(pres.view(torch.int64 if esz == 8 else torch.int32),
                                                            persistent_result.view(torch.int64 if esz == 8
                                                                                   else torch.int32)))

    if args.method == "persistent" and not args.no_contracted and not args.contracted:
        # The contracted update (opts.tune[6] = 1): the gold's term order with
        # each product fused into the running sum; not the gold's bits.
        fused = dict(opts, tune=[0, 0, 0, 0, 0, 0, 1])
        ms = []
        for i in range(args.warmup + args.steps):
            a.copy_(init)
            torch.cuda.synchronize()
            frep = P.j3d_device(a, b, n, n, n, T, dtype=args.dtype, method="persistent", **fused)
            if i >= args.warmup:
                ms.append(frep.elapsed_ms)
        fres = (b if frep.result_in_b else a).double()
        ref = persistent_result.double()
        dev = ((fres - ref).abs().max() / ref.abs().max()).item()
        fval = len(ms) * T * cells / (sum(ms) / 1e3) / 1e9
        fr = pass_roofline(PERSISTENT_KERNEL, cells, T, frep.reserved[0], esz, sum(ms) / len(ms) / 1e3)
        cfg["contracted"] = {
            "what": "opts.tune[6] = 1: each product after the first fused into the running sum (FMA), gold term "
                    "order; within 4 ulp of the gold per step, not bit-identical",
            "gcells_per_s": round(fval, 2), "ms_per_solve": round(sum(ms) / len(ms), 3),
            "roofline_frac": fr["frac"], "achieved_gbs": fr["achieved"],
            "max_rel_deviation_vs_uncontracted": float(f"{dev:.3e}"),
        }
        if "per_step_kernel_gcells_per_s" in cfg:
            cfg["contracted"]["speedup_vs_per_step"] = round(fval / cfg["per_step_kernel_gcells_per_s"], 3)
        fprof, fwhy = load_traffic(n, T, args.dtype, contracted=True)
        if fprof:
            cfg["contracted"]["profile"] = {k: fprof[k] for k in ("kernel", "git_commit", "valu_busy",
                                                                  "valu_lane_insts_per_cell") if k in fprof}
            cfg["contracted"]["traffic_gbs"] = round(fprof["hbm_bytes_per_launch"] / (sum(ms) / len(ms) / 1e3) / 1e9,
                                                     1)
        else:
            cfg["contracted"]["profile_note"] = fwhy
        del fres, ref

    avg_launch_s = kern_s / args.steps
    prof, why = (load_traffic(n, T, args.dtype, args.contracted) if args.method == "persistent"
                 else (None, "per-step kernel"))
    roofline = pass_roofline(PERSISTENT_KERNEL if args.method == "persistent" else "j3d_star1_plain_kernel",
                             cells, T, levels if args.method == "persistent" else 1, esz, avg_launch_s,
                             prof["hbm_bytes_per_launch"] if prof else None)
    if prof:
        roofline["profile"] = {k: prof[k] for k in ("kernel", "git_commit", "valu_busy", "valu_lane_insts_per_cell",
                                                     "avg_launch_ms") if k in prof}
        if prof.get("valu_busy") is not None:
            roofline["limiter"] = "valu"
    else:
        roofline["traffic_note"] = why
    roofline["copy_gbs"] = round(copy_bandwidth(torch, a), 1)
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "GCells/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(kern_s / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.dtype,
        "data": "synthetic", "config": cfg, "roofline": roofline,
    }
    del a, b, init
    torch.cuda.empty_cache()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_steps, args.dtype)
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------- 3D storage precision
def run_j3d_storage(args):
    import torch
    import perks_amd as P

    n, T = args.size, args.tsteps
    cells = n ** 3
    storage = {"f64": torch.float64, "f32": torch.float32, "f16": torch.int16, "bf16": torch.int16}
    view = {"f16": torch.float16, "bf16": torch.bfloat16}
    rows, ref = {}, None
    forms = [("f64", "nearest"), ("f32", "nearest"), ("bf16", "nearest"), ("f16", "nearest"), ("bf16", "stochastic"),
             ("f16", "stochastic")]
    for dt, rounding in forms:
        name = dt if rounding == "nearest" else f"{dt}_sr"
        a = torch.empty(cells, dtype=storage[dt], device="cuda")
        b = torch.empty_like(a)
        init = torch.empty_like(a)
        P.fill_uniform(init, cells, args.seed, 0.0, 1.0, dtype=dt)
        row = {}
        for method in ("persistent", "plain"):
            times = []
            for _ in range(max(1, args.steps)):
                a.copy_(init)
                torch.cuda.synchronize()
                rep = P.j3d_device(a, b, n, n, n, T, dtype=dt, method=method, boundary="zero", coefs=CONFIG_COEFS,
                                   rounding=rounding, round_seed=args.seed)
                times.append(rep.elapsed_ms)
            ms = sum(times) / len(times)
            row[method] = {"gcells_per_s": round(T * cells / (ms / 1e3) / 1e9, 2), "ms_per_solve": round(ms, 3)}
            res = b if rep.result_in_b else a
            val = res.view(view[dt]).double() if dt in view else res.double()
            if dt == "f64" and method == "persistent":
                ref = val.clone()
            elif method == "persistent":
                row["error_vs_f64"] = _errors(torch, val, ref)
            if method == "plain":
                prev = row.get("_persistent_result")
                row["plain_equals_persistent"] = bool(torch.equal(prev, res)) if prev is not None else None
            else:
                row["_persistent_result"] = res.clone()
        row.pop("_persistent_result", None)
        rows[name] = row
        del a, b, init
        torch.cuda.empty_cache()
    for name in rows:
        for m in ("persistent", "plain"):
            rows[name][m]["speedup_vs_f32"] = round(rows[name][m]["gcells_per_s"] / rows["f32"][m]["gcells_per_s"], 3)
    # best against best: each format's faster method against FP32's faster method
    f32_best = max(rows["f32"][m]["gcells_per_s"] for m in ("persistent", "plain"))
    for name in rows:
        bm = max(("persistent", "plain"), key=lambda m: rows[name][m]["gcells_per_s"])
        rows[name]["best_method"] = bm
        rows[name]["best_vs_f32_best"] = round(rows[name][bm]["gcells_per_s"] / f32_best, 3)
    v = rows["bf16_sr"]["plain"]
    out = {
        "metric": "GCells/s, 3d7pt 512^3 1000 steps: FP64 / FP32 / FP16 / BF16 storage, persistent and per-step, "
                  "error vs FP64",
        "value": v["gcells_per_s"], "unit": "GCells/s", "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": v["ms_per_solve"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 storage / f32 arithmetic", "data": "synthetic",
        "config": {"workload": f"3d7pt star Jacobi, {n}^3 interior + zero halo, c=0.4 / 6x0.1, input uniform[0,1) "
                               f"(seed {args.seed}), {T} time steps; value = BF16 storage with stochastic "
                               "rounding, per-step kernel",
                   "size": n, "time_steps": T, "results": rows},
        "roofline": pass_roofline("j3d_star1_plain16_kernel<bf16, stochastic> (x 1000 launches)", cells, T, 1, 2,
                                  v["ms_per_solve"] / 1e3),
    }
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------- every 3D shape
def run_j3d_shapes(args):
    import torch
    import perks_amd as P

    n = args.size
    T = 100 if args.tsteps == 1000 else args.tsteps
    cells = n ** 3
    rows = {}
    for dt, tdt in (("f64", torch.float64), ("f32", torch.float32)):
        a = torch.empty(cells, dtype=tdt, device="cuda")
        b = torch.empty_like(a)
        init = torch.empty_like(a)
        P.fill_uniform(init, cells, args.seed, 0.0, 1.0, dtype=dt)
        for shape in ("3d7pt", "3d13pt", "3d17pt", "3d27pt", "poisson"):
            row = {}
            ref = None
            equal = True
            # the streaming kernel's tile heights (opts.tune[4]: 2 = 64 x 8, 4 = 64 x 16, FP32 arithmetic
            # only); the 7-point star runs on its own kernels
            tiles = [("default", None)] if shape == "3d7pt" else (
                [("64x8", [0, 0, 0, 0, 2]), ("64x16", [0, 0, 0, 0, 4])] if dt == "f32" else [("64x8", [0, 0, 0, 0, 2])])
            for method in ("persistent", "plain"):
                best = None
                for tname, tune in tiles:
                    times = []
                    for _ in range(max(1, args.steps)):
                        a.copy_(init)
                        torch.cuda.synchronize()
                        rep = P.j3d_device(a, b, n, n, n, T, dtype=dt, shape=shape, method=method, boundary="clamp",
                                           tune=tune)
                        times.append(rep.elapsed_ms)
                    # bit patterns: with the reference weights the FP32 fields overflow to inf/NaN
                    res = (b if rep.result_in_b else a).view(torch.int64 if dt == "f64" else torch.int32)
                    if ref is None:
                        ref = res.clone()
                    else:
                        equal = equal and bool(torch.equal(res, ref))
                    ms = sum(times) / len(times)
                    entry = {"gcells_per_s": round(T * cells / (ms / 1e3) / 1e9, 2), "ms_per_solve": round(ms, 3),
                             "steps_per_pass": rep.reserved[0] if method == "persistent" else 1, "tile": tname}
                    row[f"{method}_{tname}"] = entry
                    if best is None or entry["gcells_per_s"] > best["gcells_per_s"]:
                        best = entry
                row[method] = best
            row["speedup"] = round(row["persistent"]["gcells_per_s"] / row["plain"]["gcells_per_s"], 3)
            row["all_results_bitwise_equal"] = equal
            rows[f"{shape}_{dt}"] = row
            del ref
        del a, b, init
        torch.cuda.empty_cache()
    v = rows["3d27pt_f64"]["persistent"]
    out = {
        "metric": "GCells/s, every 3D shape at 512^3, persistent vs per-step", "value": v["gcells_per_s"],
        "unit": "GCells/s", "n_gpus": 1, "steps": args.steps, "warmup": 0, "ms_per_step": v["ms_per_solve"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D Jacobi, every reference shape with its reference weights, {n}^3, clamp boundary, "
                               f"{T} time steps; value = 3d27pt FP64 persistent", "size": n, "time_steps": T,
                   "results": rows},
        "roofline": pass_roofline("j3s_kernel<3d27pt, double> (register streaming)", cells, T, v["steps_per_pass"], 8,
                                  v["ms_per_solve"] / 1e3),
    }
    print(json.dumps(out), flush=True)


def run_j3d_small(args):
    import ctypes as C
    import torch
    import perks_amd as P

    lib = P.load()
    T = args.tsteps
    rows = []
    for dtype in ("f64", "f32"):
        dt = 0 if dtype == "f64" else 1
        tdt = torch.float64 if dtype == "f64" else torch.float32
        for n in (512, 256):
            o = P.make_opts(boundary="zero", coefs=CONFIG_COEFS)
            h = lib.perks_j3d_min_height(dt, n, n, C.byref(o))  # --checkmindomain prints "h wy wx"
            row = {"dtype": dtype, "checkmindomain": f"{h} {n} {n}", "cells": h * n * n, "time_steps": T}
            init = torch.empty(h * n * n, dtype=tdt, device="cuda")
            P.fill_uniform(init, init.numel(), args.seed, 0.0, 1.0, dtype=dtype)
            a, b = torch.empty_like(init), torch.empty_like(init)
            results = {}
            for name, method, tune in (("resident", "persistent", [0, 0, 0, 0, 5]),
                                       ("temporal_blocking", "persistent", None), ("per_step", "plain", None)):
                ms = []
                for i in range(1 + args.steps):
                    a.copy_(init)
                    torch.cuda.synchronize()
                    rep = P.j3d_device(a, b, h, n, n, T, dtype=dtype, method=method, boundary="zero",
                                       coefs=CONFIG_COEFS, tune=tune)
                    if i > 0:
                        ms.append(rep.elapsed_ms)
                res = (b if rep.result_in_b else a).clone()
                v = len(ms) * T * h * n * n / (sum(ms) / 1e3) / 1e9
                row[name] = {"gcells_per_s": round(v, 1), "ms_per_solve": round(sum(ms) / len(ms), 3),
                             "us_per_step": round(sum(ms) / len(ms) / T * 1e3, 2), "grid": rep.grid[0],
                             "steps_per_pass": rep.reserved[0]}
                results[name] = res
            row["bitwise_equal"] = all(torch.equal(results[k], results["per_step"]) for k in results)
            best = max(("resident", "temporal_blocking", "per_step"), key=lambda k: row[k]["gcells_per_s"])
            row["fastest"] = best
            rows.append(row)
            del a, b, init, results
            torch.cuda.empty_cache()
    # The resident kernel's step cost against its grid size: G workgroups of
    # eight 62 x 8 x 6 FP64 blocks each (a 6 x 64 x 62G domain), so the work
    # per wave is fixed and the growth is the grid barrier's.
    scaling = []
    for G in (1, 8, 64, 128, 256):
        nz, ny, nx = 6, 64, 62 * G
        a = torch.zeros(nz * ny * nx, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        ms = []
        for i in range(3):
            rep = P.j3d_device(a, b, nz, ny, nx, T, dtype="f64", method="persistent", boundary="zero",
                               coefs=CONFIG_COEFS, tune=[0, 0, 0, 0, 5])
            ms.append(rep.elapsed_ms)
        scaling.append({"workgroups": rep.grid[0], "us_per_step": round(min(ms) / T * 1e3, 2)})
        del a, b
    print(json.dumps({"workload": "j3d-small: the 7-point solve at the getminHeight domain (zero halo, c=0.4 / 6x0.1)",
                      "unit": "GCells/s", "rows": rows, "resident_step_vs_grid": scaling}), flush=True)


L2_BYTES = 8 * 4 * 1024 * 1024  # MI355X: 4 MB of L2 per XCD, 8 XCDs


def run_j2d_domain(args):
    import torch
    import perks_amd as P

    lib = P.load()
    T = 100
    sizes = [256, 512, 768, 1024, 1536, 2048, 3072, 4096, 6144, 8192]
    rows = {}
    for dt in ("f32", "f64"):
        esz = 4 if dt == "f32" else 8
        tdt = torch.float32 if dt == "f32" else torch.float64
        sweep = []
        for n in sizes:
            a = torch.empty(n * n, dtype=tdt, device="cuda")
            P.fill_uniform(a, n * n, args.seed, 0.0, 1.0, dtype=dt)
            b = torch.empty_like(a)
            ms = []
            for _ in range(2):
                rep = P.jacobi2d_device(a, b, n, n, T, dtype=dt, shape="2d5pt", method="plain", boundary="clamp")
                ms.append(rep.elapsed_ms)
            sweep.append({"width": n, "mib": round(n * n * esz / 2 ** 20, 2),
                          "gcells_per_s": round(T * n * n / (min(ms) / 1e3) / 1e9, 2)})
            del a, b
        # getMinWidthY (jacobi-2d.cu:79-364) for this width: rows whose field the chip holds
        peak = max(r["gcells_per_s"] for r in sweep)
        chosen = None
        for r in sweep:
            min_rows = lib.perks_get_min_width_y_2d(0 if dt == "f64" else 1, r["width"], 256, 256, 1, 0, 0)
            floor = max(min_rows * r["width"] * esz, L2_BYTES // 2)
            r["min_width_y"] = min_rows
            if chosen is None and r["width"] * r["width"] * esz > floor and r["gcells_per_s"] >= 0.85 * peak:
                chosen = r["width"]
        n = chosen or sizes[-1]
        init = torch.empty(n * n, dtype=tdt, device="cuda")
        P.fill_uniform(init, n * n, args.seed, 0.0, 1.0, dtype=dt)
        a, b = torch.empty_like(init), torch.empty_like(init)
        res = {}
        timing = {}
        for method in ("persistent", "plain"):
            ms = []
            for _ in range(2):
                a.copy_(init)
                torch.cuda.synchronize()
                rep = P.jacobi2d_device(a, b, n, n, 500, dtype=dt, shape="2d5pt", method=method, boundary="clamp")
                ms.append(rep.elapsed_ms)
            res[method] = (b if rep.result_in_b else a).view(torch.int32 if dt == "f32" else torch.int64).clone()
            timing[method] = {"gcells_per_s": round(500 * n * n / (min(ms) / 1e3) / 1e9, 2),
                              "ms_per_solve": round(min(ms), 3)}
        rows[dt] = {"sweep_per_step_100_steps": sweep, "peak_gcells_per_s": peak, "chosen_width": n,
                    "at_chosen_500_steps": timing,
                    "speedup": round(timing["persistent"]["gcells_per_s"] / timing["plain"]["gcells_per_s"], 3),
                    "bitwise_equal": bool(torch.equal(res["persistent"], res["plain"]))}
        del a, b, init, res
        torch.cuda.empty_cache()
    print(json.dumps({"workload": "j2d-domain: 2d5pt clamp, per-step sweep over square domains, domain choice as "
                                  "ProcessDomaindata.py, then persistent vs per-step there (500 steps)",
                      "l2_bytes": L2_BYTES, "results": rows}), flush=True)


def run_j3d_tune(args):
    import torch
    import perks_amd as P

    n = args.size
    cells = n ** 3
    sample = 48
    tables = {}
    best = {}
    for shape, dtypes in (("3d7pt", ("f64", "f32", "f16", "bf16")), ("3d27pt", ("f64", "f32"))):
        for dt in dtypes:
            tdt = {"f64": torch.float64, "f32": torch.float32}.get(dt, torch.int16)
            init = torch.empty(cells, dtype=tdt, device="cuda")
            P.fill_uniform(init, cells, args.seed, 0.0, 1.0, dtype=dt)
            opts, rows = P.j3d_autotune(init, n, n, n, dtype=dt, sample_steps=sample, shape=shape, boundary="zero")
            torch.cuda.synchronize()
            key = f"{shape}_{dt}"
            tables[key] = rows
            ok = [r for r in rows if r["status"] == 0]
            b = max(ok, key=lambda r: r["gcells_per_s"])
            best[key] = {"method": b["method"], "variant": b["variant"], "levels": b["levels"],
                         "gcells_per_s": round(b["gcells_per_s"], 2), "tune2": opts.tune[2], "tune4": opts.tune[4]}
            for r in rows:
                r["gcells_per_s"] = round(r["gcells_per_s"], 2)
            del init
            torch.cuda.empty_cache()
    v = best["3d7pt_f64"]
    out = {
        "metric": f"autotuned 3D configuration at 512^3 (sample of {sample} steps per configuration)",
        "value": v["gcells_per_s"], "unit": "GCells/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(sample * cells / (v["gcells_per_s"] * 1e9) * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"3D Jacobi {n}^3, zero boundary, reference weights, input uniform[0,1) (seed "
                               f"{args.seed}); every configuration timed on {sample} steps; value = best 3d7pt FP64",
                   "size": n, "sample_steps": sample, "best": best, "tables": tables},
    }
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------- 2D storage precision
def _errors(torch, got, ref):
    """max and RMS of |got - ref|, absolute and relative to max|ref| / RMS(ref)."""
    d = (got - ref).abs()
    mx, rms = d.max().item(), d.square().mean().sqrt().item()
    rmx, rrms = ref.abs().max().item(), ref.square().mean().sqrt().item()
    return {"max_abs": mx, "rms_abs": rms, "max_rel": mx / rmx if rmx else None,
            "rms_rel": rms / rrms if rrms else None}


def run_j2d_storage(args):
    import numpy as np
    import torch
    import perks_amd as P

    n = 8192 if args.size == 512 else args.size
    T = 500 if args.tsteps == 1000 else args.tsteps
    checkpoints = [c for c in (10, 100, 500) if c < T] + [T]
    cells = n * n
    storage = {"f64": torch.float64, "f32": torch.float32, "f16": torch.int16, "bf16": torch.int16}
    view = {"f16": torch.float16, "bf16": torch.bfloat16}
    results, kernel_ms = {}, {}
    for shape in ("2d9pt", "2d25pt"):
        si = P.shape_info(shape)
        c = np.random.default_rng(args.seed).random(len(si.coef))
        coefs = list(c / c.sum())
        ref_at = {}
        per = {}
        for dt, rounding in (("f64", "nearest"), ("f32", "nearest"), ("f16", "nearest"), ("bf16", "nearest"),
                             ("f16", "stochastic"), ("bf16", "stochastic")):
            name = dt if rounding == "nearest" else f"{dt}_sr"
            a = torch.empty(cells, dtype=storage[dt], device="cuda")
            b = torch.empty_like(a)
            init = torch.empty_like(a)
            P.fill_uniform(init, cells, args.seed, -1.0, 1.0, dtype=dt)
            opts = P.make_opts(shape, "clamp", "persistent", coefs, rounding=rounding, round_seed=args.seed)
            row = {}
            for steps in checkpoints:
                a.copy_(init)
                torch.cuda.synchronize()
                rep = P.jacobi2d_device(a, b, n, n, steps, dtype=dt, opts=opts)
                res = b if rep.result_in_b else a
                val = res.view(view[dt]).double() if dt in view else res.double()
                if dt == "f64":
                    ref_at[steps] = val.clone()
                else:
                    row[str(steps)] = _errors(torch, val, ref_at[steps])
                if steps == T:
                    times = []
                    for _ in range(max(1, args.steps)):
                        a.copy_(init)
                        torch.cuda.synchronize()
                        times.append(P.jacobi2d_device(a, b, n, n, T, dtype=dt, opts=opts).elapsed_ms)
                    ms = sum(times) / len(times)
                    kernel_ms[(shape, dt)] = ms
                    row["gcells_per_s"] = round(T * cells / (ms / 1e3) / 1e9, 2)
                    row["ms_per_solve"] = round(ms, 3)
                    row["steps_per_pass"] = rep.reserved[0]
                    plain = P.make_opts(shape, "clamp", "plain", coefs, rounding=rounding, round_seed=args.seed)
                    a.copy_(init)
                    torch.cuda.synchronize()
                    pr = P.jacobi2d_device(a, b, n, n, T, dtype=dt, opts=plain)
                    row["per_step_kernel_gcells_per_s"] = round(T * cells / (pr.elapsed_ms / 1e3) / 1e9, 2)
            per[name] = row
            del a, b, init
        del ref_at
        torch.cuda.empty_cache()
        for name in per:
            per[name]["speedup_vs_f32"] = round(per[name]["gcells_per_s"] / per["f32"]["gcells_per_s"], 3)
        results[shape] = per
    bf = results["2d9pt"]["bf16"]
    ms = kernel_ms[("2d9pt", "bf16")]
    out = {
        "metric": "GCells/s, 2D box r1/r2 8192^2 500 steps, FP32 vs FP16 vs BF16 storage (FP32 arithmetic), "
                  "error vs FP64",
        "value": bf["gcells_per_s"], "unit": "GCells/s", "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 storage / f32 arithmetic", "data": "synthetic",
        "config": {"workload": f"2D box stencils 2d9pt (r1) and 2d25pt (r2), {n}^2, clamp boundary, coefficients "
                               f"uniform[0,1) normalised to sum 1 (seed {args.seed}), input uniform[-1,1), {T} steps, "
                               "persistent kernel; value = 2d9pt BF16", "size": n, "time_steps": T,
                   "results": results},
        "roofline": pass_roofline("j2d_stream_kernel<2d9pt, bf16s>", cells, T, bf["steps_per_pass"], 2, ms / 1e3),
    }
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------- batched 2D grids
def batch_specs(seed, count=512):
    import numpy as np
    rng = np.random.default_rng(seed)
    sides = np.exp(rng.uniform(np.log(64), np.log(1024), size=(count, 2))).astype(int)
    steps = rng.integers(50, 501, size=count)
    return [(int(h), int(w), int(t)) for (h, w), t in zip(sides, steps)]


def run_batch2d(args):
    import torch
    import perks_amd as P

    specs = batch_specs(args.seed)
    cell_steps = sum(h * w * t for h, w, t in specs)
    out_rows = {}
    for shape in ("2d5pt", "2d9pt"):
        si = P.shape_info(shape)
        coefs = [1.0 / len(si.coef)] * len(si.coef)
        bufs = []
        for i, (h, w, _) in enumerate(specs):
            a = torch.empty(h * w, dtype=torch.float32, device="cuda")
            P.fill_uniform(a, h * w, args.seed * 100003 + i, 0.0, 1.0, dtype="f32")
            bufs.append((a.clone(), a, torch.empty_like(a)))
        grids = [(a, b, h, w, t) for (_, a, b), (h, w, t) in zip(bufs, specs)]

        def reset():
            for init, a, _ in bufs:
                a.copy_(init)
            torch.cuda.synchronize()

        opts = P.make_opts(shape, "zero", "persistent", coefs)
        t0 = time.perf_counter()
        plan = P.BatchPlan(grids, dtype="f32", autotune=True, opts=opts)
        plan_s = time.perf_counter() - t0
        flags = plan.result_in_b
        reset()
        plan.run()
        times = []
        for _ in range(max(1, args.steps)):
            reset()
            rep = plan.run()
            times.append(rep.elapsed_ms)
        batched = [(b if f else a).clone() for (_, a, b), f in zip(bufs, flags)]
        ms = sum(times) / len(times)
        # the same batch without the plan (default configuration, no graph)
        reset()
        drep, _ = P.jacobi2d_batch(grids, dtype="f32", opts=opts)
        # One grid after another with the per-step kernel (one launch per time step).
        plain = P.make_opts(shape, "zero", "plain", coefs)
        reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        seq = []
        for (init, a, b), (h, w, t) in zip(bufs, specs):
            r = P.jacobi2d_device(a, b, h, w, t, dtype="f32", opts=plain)
            seq.append(b if r.result_in_b else a)
        torch.cuda.synchronize()
        seq_ms = (time.perf_counter() - t0) * 1e3
        equal = all(torch.equal(x.view(torch.int32), y.view(torch.int32)) for x, y in zip(batched, seq))
        out_rows[shape] = {"ms_per_batch": round(ms, 3), "gcells_per_s": round(cell_steps / (ms / 1e3) / 1e9, 2),
                           "sequential_per_step_ms": round(seq_ms, 3),
                           "sequential_per_step_gcells_per_s": round(cell_steps / (seq_ms / 1e3) / 1e9, 2),
                           "speedup": round(seq_ms / ms, 3), "bitwise_equal": equal,
                           "grid": rep.grid[0], "steps_per_pass": rep.reserved[0], "passes": rep.reserved[1],
                           "autotuned_config": plan.config, "plan_seconds": round(plan_s, 3),
                           "untuned_direct_call_ms": round(drep.elapsed_ms, 3)}
        plan.close()
        del bufs, grids, batched, seq
        torch.cuda.empty_cache()
    r = out_rows["2d5pt"]
    out = {
        "metric": "GCells/s, 512 batched 2D grids (5pt/9pt FP32) in one autotuned, hipGraph-captured persistent launch "
                  "vs sequential per-step",
        "value": r["gcells_per_s"], "unit": "GCells/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": r["ms_per_batch"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"512 grids, sides log-uniform in [64,1024] (non-square), steps uniform in [50,500], "
                               f"FP32, zero boundary, equal weights, seed {args.seed}; value = 2d5pt",
                   "grids": len(specs), "cell_steps": cell_steps, "results": out_rows},
        "roofline": pass_roofline(("j2d_stream_kernel" if r["autotuned_config"]["tile"] in (2, 4) else "j2d_kernel")
                                  + "<2d5pt, float> (batched)", [wy * wx for wy, wx, _ in specs],
                                  [st for _, _, st in specs], r["steps_per_pass"], 4, r["ms_per_batch"] / 1e3),
    }
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------- CG
def cg_bytes_per_iter(fmt, n, nnz, vbytes):
    """Bytes one iteration moves in the reference's model (cg_driver.cu:359-361):
    the matrix once plus six vector streams."""
    if fmt == "csr":
        return nnz * vbytes + (nnz + n + 1) * 4 + 6 * n * vbytes
    return 7 * n * vbytes + 6 * n * vbytes


def cg_skewed(torch, P, args):
    """Load balance on an irregular matrix: CG (FP64, CSR, 200 iterations) on
    a symmetric, diagonally dominant matrix with power-law row lengths and
    three hub rows of 5*10^4 to 1.5*10^5 entries (split across the grid),
    next to a control matrix with the same rows, about the same entries and
    the same random columns but every row alike.  Equal times per iteration
    mean the work split absorbs the skew."""
    rowptr, col, val = P.skewed_spd_csr(1_000_000, args.seed, hubs=(150_000, 100_000, 50_000), max_deg=200)
    out = cg_csr_row(torch, P, args, rowptr, col, val)
    # the control: the same rows and about the same entries, every row alike
    rowptr, col, val = P.skewed_spd_csr(1_000_000, args.seed, degree=max(1, round(len(col) / 2e6)))
    ctl = cg_csr_row(torch, P, args, rowptr, col, val)
    out["uniform_control"] = ctl
    out["time_vs_uniform_control"] = round(out["ms_per_iteration"] / ctl["ms_per_iteration"], 3)
    return out


def cg_csr_row(torch, P, args, rowptr, col, val):
    import numpy as np
    n, nnz = len(rowptr) - 1, len(col)
    A = dict(rowptr=torch.from_numpy(rowptr).cuda(), col=torch.from_numpy(col).cuda(),
             val=torch.from_numpy(val).cuda(), nnz=nnz)
    b = torch.empty(n, dtype=torch.float64, device="cuda")
    P.fill_uniform(b, n, args.seed, -1.0, 1.0, dtype="f64")
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    P.cg_solve("csr", n, b, x, tol=1e-10, max_iter=200, static_iter=True, **A)  # warm-up
    x.zero_()
    torch.cuda.synchronize()
    res = P.cg_solve("csr", n, b, x, tol=1e-10, max_iter=200, static_iter=True, **A)
    lens = np.diff(rowptr)
    per_iter = cg_bytes_per_iter("csr", n, nnz, 8)
    return {"n": n, "nnz": nnz, "max_row": int(lens.max()), "median_row": float(np.median(lens)),
            "rows_over_32": int((lens > 32).sum()), "rows_over_4096": int((lens > 4096).sum()),
            "iterations": res.iterations, "ms_per_iteration": round(res.elapsed_ms / max(1, res.iterations), 4),
            "achieved_gbs": round(per_iter * res.iterations / (res.elapsed_ms / 1e3) / 1e9, 1),
            "long_rows": res.long_rows, "grid": res.grid}


def run_cg_counters(args):
    """The four FP64 solves whose counters profiles/cg_poisson_256.md reports,
    one launch each and in this order: CSR cached, CSR uncached, DIA cached,
    DIA uncached (256^3 Poisson, to ||r||/||b|| < 1e-10)."""
    import torch
    import perks_amd as P

    m = 256 if args.size == 512 else args.size
    n = m ** 3
    nnz = P.poisson7_nnz(m, m, m)
    b = torch.empty(n, dtype=torch.float64, device="cuda")
    P.fill_uniform(b, n, args.seed, -1.0, 1.0, dtype="f64")
    out = []
    for fmt in ("csr", "dia"):
        if fmt == "csr":
            A = dict(rowptr=torch.empty(n + 1, dtype=torch.int32, device="cuda"),
                     col=torch.empty(nnz, dtype=torch.int32, device="cuda"),
                     val=torch.empty(nnz, dtype=torch.float64, device="cuda"), nnz=nnz)
            P.poisson7_csr(m, m, m, A["rowptr"], A["col"], A["val"])
        else:
            val = torch.empty(7 * n, dtype=torch.float64, device="cuda")
            A = dict(val=val, offsets=P.poisson7_dia(m, m, m, val))
        for cache in (True, False):
            x = torch.zeros(n, dtype=torch.float64, device="cuda")
            res = P.cg_solve(fmt, n, b, x, tol=1e-10, x0_is_zero=True, cache_vector=cache, cache_matrix=cache, **A)
            out.append({"fmt": fmt, "cache": cache, "iterations": res.iterations, "ms": round(res.elapsed_ms, 3),
                        "lds_bytes": res.lds_bytes, "cache_units": res.cache_units, "grid": res.grid})
        del A
        torch.cuda.empty_cache()
    print(json.dumps({"workload": "cg-counters", "n": n, "solves": out}), flush=True)


def run_cg(args):
    import torch
    import perks_amd as P

    m = 256 if args.size == 512 else args.size
    n = m ** 3
    nnz = P.poisson7_nnz(m, m, m)
    b = torch.empty(n, dtype=torch.float64, device="cuda")
    P.fill_uniform(b, n, args.seed, -1.0, 1.0, dtype="f64")
    rows = {}
    for fmt in ("csr", "dia"):
        if fmt == "csr":
            A = dict(rowptr=torch.empty(n + 1, dtype=torch.int32, device="cuda"),
                     col=torch.empty(nnz, dtype=torch.int32, device="cuda"),
                     val=torch.empty(nnz, dtype=torch.float64, device="cuda"), nnz=nnz)
            P.poisson7_csr(m, m, m, A["rowptr"], A["col"], A["val"])
        else:
            val = torch.empty(7 * n, dtype=torch.float64, device="cuda")
            A = dict(val=val, offsets=P.poisson7_dia(m, m, m, val))
        for prec, cache in (("f64", True), ("f64", False), ("f32_refine", True)):
            x = torch.zeros(n, dtype=torch.float64, device="cuda")
            kw = dict(tol=1e-10, precision=prec, cache_vector=cache, cache_matrix=cache)
            P.cg_solve(fmt, n, b, x, **kw, **A)  # warm-up
            times = []
            for _ in range(max(1, args.steps)):
                x.zero_()
                torch.cuda.synchronize()
                res = P.cg_solve(fmt, n, b, x, **kw, **A)
                times.append(res.elapsed_ms)
            ms = sum(times) / len(times)
            # independent FP64 residual with torch on the 3D grid view
            xv = x.view(m, m, m)
            ax = 6.0 * xv
            ax[1:] -= xv[:-1]
            ax[:-1] -= xv[1:]
            ax[:, 1:] -= xv[:, :-1]
            ax[:, :-1] -= xv[:, 1:]
            ax[:, :, 1:] -= xv[:, :, :-1]
            ax[:, :, :-1] -= xv[:, :, 1:]
            true_rel = ((b - ax.reshape(-1)).norm() / b.norm()).item()
            vb = 8 if prec == "f64" else 4
            per_iter = cg_bytes_per_iter(fmt, n, nnz, vb)
            rows[f"{fmt}_{prec}" + ("" if cache else "_nocache")] = {
                "code": res.code, "iterations": res.iterations, "outer_iterations": res.outer_iterations,
                "time_to_tolerance_ms": round(ms, 3), "ms_per_iteration": round(ms / max(1, res.iterations), 4),
                "final_rel_residual_fp64": res.relative_residual, "true_rel_residual_fp64": true_rel,
                "grid": res.grid,
                "achieved_gbs": round(per_iter * res.iterations / (ms / 1e3) / 1e9, 1),
                "lds_cache_bytes_per_workgroup": res.lds_bytes, "cached_vector_units": res.cache_units,
                "cached_matrix": res.cache_matrix,
            }
        del A
        torch.cuda.empty_cache()
    rows["skewed_csr_f64"] = cg_skewed(torch, P, args)
    r = rows["csr_f64"]
    achieved = r["achieved_gbs"]
    traffic = None
    try:
        t = json.loads((ROOT / "profiles" / "cg_csr_f64_256_traffic.json").read_text())
        if t.get("n") == n and t.get("iterations") == r["iterations"]:
            traffic = t.get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        pass
    out = {
        "metric": "CG time-to-tolerance, 7pt Poisson 256^3, CSR/DIA, FP64 vs FP32-inner/FP64-refinement",
        "value": r["time_to_tolerance_ms"], "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": r["time_to_tolerance_ms"], "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"CG on the 3D Poisson 7-point matrix (diag 6, off-diag -1), n={m}^3, rhs "
                               f"uniform[-1,1) (seed {args.seed}), x0=0, ||r||/||b|| < 1e-10; value = CSR FP64",
                   "n": n, "nnz": nnz, "results": rows},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "cg_persistent_kernel<double, CSR>",
                     "algorithmic_bytes_per_launch": cg_bytes_per_iter("csr", n, nnz, 8) * r["iterations"]},
    }
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------- recorded counter profile
def _counter_rows(directory):
    """(kernel name, dispatch id, counter, value, grid size) of every
    rocprofv3 counter CSV under `directory`."""
    import csv
    out = []
    files = sorted(pathlib.Path(directory).rglob("*counter_collection.csv"))
    if not files:
        raise SystemExit(f"no *counter_collection.csv under {directory}")
    for f in files:
        with open(f, newline="") as fh:
            for row in csv.DictReader(fh):
                out.append((row["Kernel_Name"], (str(f), row.get("Dispatch_Id", "")), row["Counter_Name"],
                            float(row["Counter_Value"]), int(row.get("Grid_Size") or 0)))
    return out


def _per_dispatch(rows, kernel_substr, counter, largest_grid=False):
    """Average over the dispatches of a kernel of the counter's value (summed
    over the rows of one dispatch); with largest_grid, only the dispatches of
    the kernel's largest grid (the field-sized copies, not small ones)."""
    acc = {}
    sel = [r for r in rows if kernel_substr in r[0] and r[2] == counter]
    if largest_grid and sel:
        g = max(r[4] for r in sel)
        sel = [r for r in sel if r[4] == g]
    for name, disp, cname, v, _ in sel:
        acc[disp] = acc.get(disp, 0.0) + v
    if not acc:
        raise SystemExit(f"no {counter} rows for a kernel matching {kernel_substr!r}")
    return sum(acc.values()) / len(acc), len(acc)


def record_profile(args, out_path=None):
    """Writes TRAFFIC_FILE from the counter passes of
        rocprofv3 --pmc FETCH_SIZE -- python3 bench.py --steps 1 --warmup 0 --no-cpu --no-per-step
    (then WRITE_SIZE, then the SQ counters).  The byte counters are scaled by
    what they read for the run's own device-to-device copies of the field
    (__amd_rocclr_copyBuffer, one read and one write of cells * sizeof(T))."""
    import subprocess
    dirs = args.record_profile
    contracted = bool(getattr(args, "contracted", False))
    n, T = args.size, args.tsteps
    esz = 8 if args.dtype == "f64" else 4
    field = n ** 3 * esz
    fetch_rows, write_rows = _counter_rows(dirs[0]), _counter_rows(dirs[1])
    fetch, nd = _per_dispatch(fetch_rows, PERSISTENT_KERNEL, "FETCH_SIZE")
    write, _ = _per_dispatch(write_rows, PERSISTENT_KERNEL, "WRITE_SIZE")
    cf, _ = _per_dispatch(fetch_rows, "copyBuffer", "FETCH_SIZE", largest_grid=True)
    cw, _ = _per_dispatch(write_rows, "copyBuffer", "WRITE_SIZE", largest_grid=True)
    ff, wf = field / (cf * 1024), field / (cw * 1024)
    traffic = int(round(fetch * 1024 * ff + write * 1024 * wf))
    names = sorted({r[0] for r in fetch_rows if PERSISTENT_KERNEL in r[0]})
    if len(names) > 1:
        raise SystemExit(f"several persistent kernels in the counter run ({names}): profile with --no-contracted")
    try:
        commit = subprocess.run(["git", "-C", str(ROOT), "rev-parse", "--short=12", "HEAD"], capture_output=True,
                                text=True).stdout.strip() or None
    except OSError:
        commit = None
    d = {"kernel": names[0] if names else PERSISTENT_KERNEL, "size": n, "time_steps": T, "dtype": args.dtype,
         "dispatches": nd, "fetch_size_kib": fetch, "write_size_kib": write,
         "fetch_factor": round(ff, 4), "write_factor": round(wf, 4), "hbm_bytes_per_launch": traffic,
         "calibration": f"__amd_rocclr_copyBuffer of the {field}-byte field in the same runs: FETCH_SIZE {cf:.1f} KiB, "
                        f"WRITE_SIZE {cw:.1f} KiB",
         "source_sha256": kernel_sources_sha(), "git_commit": commit,
         "contracted": contracted,
         "command": "rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE, then SQ_INSTS_VALU SQ_ACTIVE_INST_VALU "
                    "SQ_BUSY_CYCLES) -- python3 bench.py --steps 1 --warmup 0 --no-cpu --no-per-step "
                    + ("--contracted" if contracted else "--no-contracted")}
    if len(dirs) > 2:
        sq = _counter_rows(dirs[2])
        insts, _ = _per_dispatch(sq, PERSISTENT_KERNEL, "SQ_INSTS_VALU")
        active, _ = _per_dispatch(sq, PERSISTENT_KERNEL, "SQ_ACTIVE_INST_VALU")
        busy, _ = _per_dispatch(sq, PERSISTENT_KERNEL, "SQ_BUSY_CYCLES")
        # SQ_ACTIVE_INST_VALU counts quad-cycles summed over SIMDs; SQ_BUSY_CYCLES is summed over shader engines.
        d["valu_busy"] = round(4 * active / (SIMDS * busy / SHADER_ENGINES), 4)
        d["valu_lane_insts_per_cell"] = round(insts * 64 / (T * n ** 3), 2)
    (out_path or (TRAFFIC_FILE_FUSED if contracted else TRAFFIC_FILE)).write_text(json.dumps(d, indent=1) + "\n")
    print(json.dumps(d))


def main():
    args = parse()
    if args.record_profile:
        record_profile(args)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the persistent stencil path is single-device "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    {"j3d": run_j3d, "j3d-storage": run_j3d_storage, "j3d-shapes": run_j3d_shapes, "j2d-storage": run_j2d_storage,
     "batch2d": run_batch2d, "cg": run_cg, "cg-counters": run_cg_counters,
     "j3d-tune": run_j3d_tune, "j3d-small": run_j3d_small, "j2d-domain": run_j2d_domain}[args.workload](args)


if __name__ == "__main__":
    main()
