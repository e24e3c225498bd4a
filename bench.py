#!/usr/bin/env python3
"""Benchmark of the persistent 3D 7-point Jacobi solve on one MI355X.

Workload (one bench "step" = one complete solve, inputs already in HBM):
  3D 7-point star Jacobi, FP64, 512^3 interior with a one-cell zero halo,
  centre coefficient 0.4 and six neighbours 0.1, 1000 time steps, input
  uniform[0,1) from a seed.  The persistent method runs all 1000 time steps
  in one cooperative launch; the per-step method (one launch per time step)
  is timed on the same GPU and buffers for comparison, and the two results
  are compared bit for bit.

Prints one JSON line (see the keys at the bottom).  GCells/s counts
time steps * cells / seconds, the reference's own definition
(stencil/3dstencil/j3d.cu:754); the roofline prices one read plus one write
of the field per cell-update (j3d.cu:756) against 8 TB/s of HBM.
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
    ap.add_argument("--cpu-steps", type=int, default=10, help="time steps of the CPU sample")
    ap.add_argument("--method", choices=["persistent", "plain"], default="persistent")
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
        kind, what = "reference", "j3d_gold_iterative (stencil/3dstencil/3d7pt/3d7pt_gold_gen.cpp), g++ -O2"
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
    return {
        "value": round(tsteps * n ** 3 / dt / 1e9, 4),
        "unit": "GCells/s",
        "cores": 1,
        "kind": kind,
        "sample": f"{what}; {n}^3 {dtype}, {tsteps} time steps, {dt:.1f} s on {os.cpu_count()} host CPUs "
                  f"({_cpu_model()}), one thread used",
    }


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def load_traffic(n, tsteps, dtype):
    try:
        d = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    if d.get("size") != n or d.get("time_steps") != tsteps or d.get("dtype") != dtype:
        return None
    return d.get("hbm_bytes_per_launch")


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the persistent stencil path is single-device "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    import perks_amd as P

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    n, T = args.size, args.tsteps
    cells = n ** 3
    tdt = torch.float64 if args.dtype == "f64" else torch.float32
    esz = 8 if args.dtype == "f64" else 4
    a = torch.empty(cells, dtype=tdt, device="cuda")
    b = torch.empty(cells, dtype=tdt, device="cuda")
    init = torch.empty(cells, dtype=tdt, device="cuda")
    P.fill_uniform(init, cells, 1, 0.0, 1.0, dtype=args.dtype)
    opts = dict(boundary="zero", coefs=CONFIG_COEFS)

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

    avg_launch_s = kern_s / args.steps
    alg_bytes = T * cells * 2 * esz
    achieved = alg_bytes / avg_launch_s / 1e9
    traffic = load_traffic(n, T, args.dtype)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "j3d_star1_persistent_kernel" if args.method == "persistent" else
                          "j3d_star1_plain_kernel",
                "algorithmic_bytes_per_launch": alg_bytes}
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


if __name__ == "__main__":
    main()
