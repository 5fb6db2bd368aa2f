"""Benchmark of the explicit step on one MI355X.

Workload (N=1): one 3-D block 256x128x128 with 9 synthetic species, plain
FP64, Van Albada MUSCL + AUSM+ + central viscous flux + Wilke transport,
SSP-RK3 with the global CFL time step (the "plain FP64 drop-in" config).
A step is one RK3 step (3 stages) over the whole block; inputs are resident
in HBM before the timed region.

Prints one JSON line: metric/value in Mcell-stage updates/s (interior cells x
stages / wall time), the roofline object of the dominant kernel (the fused
RHS), and the CPU baseline (the oracle on the host cores, bounded sample).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
  python bench.py --copy-calibration   # plain copy kernel of known size, for
                                       # scaling rocprofv3 FETCH/WRITE_SIZE
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per cell-stage of the FP64 SSP-RK3 stage:
#   B = 8 * [(8/3) n_c + n_g + 2] with n_c = 14 conserved, n_g = 10 geometric
BYTES_PER_CELL_STAGE_FP64 = 8.0 * ((8.0 / 3.0) * 14 + 10 + 2)
# mixed precision keeps Q, T and the reductions FP64 and stores geometry FP32
BYTES_PER_CELL_STAGE_MIXED = 8.0 * ((8.0 / 3.0) * 14 + 2) + 4.0 * 10


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def copy_calibration():
    import torch
    n = 1 << 28  # 2 GiB of float64
    a = torch.empty(n, dtype=torch.float64, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8}))


def cpu_baseline(case, threads):
    from oracle import pyoracle
    pyoracle.build()
    used = pyoracle.set_threads(threads)
    o = pyoracle.Oracle(case)
    o.upload_initial()
    o.init_conservative()
    t0 = time.perf_counter()
    o.step(1)
    dt = time.perf_counter() - t0
    cs = case.n_cells * 3
    o.close()
    return {"value": cs / dt / 1e6, "unit": "Mcell-stage updates/s", "cores": used, "kind": "port",
            "sample": f"1 SSP-RK3 step (3 stages) of the same 256x128x128 9-species block, "
                      f"{cs} cell-stages in {dt:.2f} s, OpenMP on {used} host threads"}


def measure(case, args, name):
    from combustioncodegpu_amd.lib import Solver
    g = Solver(case)
    g.upload_initial()
    g.init_conservative()
    g.sync()
    log(f"[bench] {name}: warmup")
    g.step(args.warmup)
    g.sync()
    rc, nfail = g.check()
    if rc != 0:
        log(f"[bench] {name}: health check failed rc={rc} newton_fail={nfail}")
        return None
    log(f"[bench] {name}: timed region")
    g.sync()
    t0 = time.perf_counter()
    ev_ms = g.time_steps(args.steps)
    g.sync()
    wall = time.perf_counter() - t0
    value = case.n_cells * 3 * args.steps / wall / 1e6
    # per-kernel durations from HIP events on the context's stream
    g.set_profiling(True)
    g.time_steps(3)
    kt = g.kernel_times()
    g.set_profiling(False)
    # one timed RHS event spans the three direction launches of one stage
    rhs_ms, rhs_n = kt["rhs"]
    rhs_avg_s = rhs_ms / max(rhs_n, 1) / 1e3
    bpc = BYTES_PER_CELL_STAGE_FP64 if name == "fp64" else BYTES_PER_CELL_STAGE_MIXED
    achieved = bpc * case.n_cells / rhs_avg_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"rhs_c1_{name}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_stage")
    res = g.residual()
    dt = g.get_dt()
    g.close()
    return {
        "name": name, "value": round(value, 3), "ms_per_step": round(1e3 * wall / args.steps, 4),
        "event_ms_per_step": round(ev_ms / args.steps, 4),
        "kernel_ms_per_step": {k: round(v[0] / 3, 4) for k, v in kt.items()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "kernel": "rhs_pencil_kernel (LDS-staged RHS, one launch per face direction)",
                     "kernel_avg_ms_per_stage": round(rhs_avg_s * 1e3, 4), "launches_per_stage": 3,
                     "bytes_per_cell_stage": round(bpc, 2)},
        "dt": dt, "residual": [float(x) for x in res],
    }


def mixed_errors(case, checkpoints=(1, 10, 100)):
    """Per-field error of the mixed path against the FP64 path after 1, 10
    and 100 RK3 steps (max over cells; rho, p, T relative to their max,
    velocity as a vector relative to max |V|, Y absolute)."""
    import numpy as np
    from combustioncodegpu_amd import abi
    from combustioncodegpu_amd.lib import Solver
    ctx = {}
    for prec in ("fp64", "mixed"):
        case.params["precision"] = abi.MIXED if prec == "mixed" else abi.FP64
        ctx[prec] = Solver(case)
        ctx[prec].upload_initial()
        ctx[prec].init_conservative()
    out = {}
    done = 0
    for n in checkpoints:
        for c in ctx.values():
            c.step(n - done)
        done = n
        b = {k: c.download(0, abi.BV)[:, 2:-2, 2:-2, 2:-2] for k, c in ctx.items()}
        y = {k: c.download(0, abi.YK)[:, 2:-2, 2:-2, 2:-2] for k, c in ctx.items()}
        r, m = b["fp64"], b["mixed"]
        vmax = np.sqrt((r[1:4] ** 2).sum(axis=0)).max()
        out[f"after_{n}_steps"] = {
            "rho": float(np.abs(m[0] - r[0]).max() / np.abs(r[0]).max()),
            "velocity": float(np.sqrt(((m[1:4] - r[1:4]) ** 2).sum(axis=0)).max() / vmax),
            "p": float(np.abs(m[4] - r[4]).max() / np.abs(r[4]).max()),
            "T": float(np.abs(m[5] - r[5]).max() / np.abs(r[5]).max()),
            "Y": float(np.abs(y["mixed"] - y["fp64"]).max()),
            "rho_l2": float(np.sqrt(np.mean((m[0] - r[0]) ** 2)) / np.sqrt(np.mean(r[0] ** 2))),
            "T_l2": float(np.sqrt(np.mean((m[5] - r[5]) ** 2)) / np.sqrt(np.mean(r[5] ** 2))),
        }
    for c in ctx.values():
        c.close()
    case.params["precision"] = abi.FP64
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--size", type=str, default="256,128,128")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--rhs-only", type=int, default=0,
                    help="profiling mode: run this many RHS launches on the workload and exit")
    ap.add_argument("--rhs-variant", type=int, default=-1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="precision of the headline line; the other variant is measured beside it")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU layout is "
              f"design-only for this path", file=sys.stderr)
        return 2
    from combustioncodegpu_amd import abi, cases
    from combustioncodegpu_amd.lib import Solver

    mx, my, mz = (int(v) for v in args.size.split(","))
    log(f"[bench] generating workload {mx}x{my}x{mz}, 9 species")
    t0 = time.perf_counter()
    case = cases.c1(mx, my, mz)
    log(f"[bench] generated in {time.perf_counter() - t0:.1f} s")
    if args.rhs_only:
        if args.precision == "mixed":
            case.params["precision"] = abi.MIXED
        g = Solver(case)
        g.upload_initial()
        g.init_conservative()
        if args.rhs_variant >= 0:
            g.set_rhs_variant(args.rhs_variant)
        for _ in range(args.rhs_only):
            g.rhs(-1)
        g.sync()
        log(f"[bench] {args.rhs_only} RHS launches done ({args.precision})")
        g.close()
        return 0
    variants = {}
    for prec in ("fp64", "mixed"):
        case.params["precision"] = abi.MIXED if prec == "mixed" else abi.FP64
        variants[prec] = measure(case, args, prec)
        if variants[prec] is None:
            return 3
    log("[bench] mixed-precision error against the FP64 path")
    errors = mixed_errors(case)
    head = variants[args.precision]
    other = variants["mixed" if args.precision == "fp64" else "fp64"]
    out = {
        "metric": "Mcell-stage updates/s, FP64 plain vs mixed+batched+surrogate; HBM GB/s",
        "value": head["value"],
        "unit": "Mcell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic (seeded smooth T/p/u fields, smoothed Dirichlet mass fractions, 9 synthetic species)",
        "config": {"workload": "explicit step, 3-D block 256x128x128, 9 species, Van Albada MUSCL, "
                               "AUSM+, viscous + Wilke, SSP-RK3 global CFL 0.5",
                   "cells": case.n_cells, "variant": args.precision,
                   "event_ms_per_step": head["event_ms_per_step"]},
        "roofline": head["roofline"],
        "kernel_ms_per_step": head["kernel_ms_per_step"],
        "variants": {
            other["name"]: {k: other[k] for k in ("value", "ms_per_step", "kernel_ms_per_step")},
            "mixed_speedup_vs_fp64": round(variants["mixed"]["value"] / variants["fp64"]["value"], 3),
            "mixed_error_vs_fp64": errors,
        },
        "dt": head["dt"],
        "residual": head["residual"],
    }
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle)")
        try:
            out["cpu_baseline"] = cpu_baseline(case, args.cpu_threads)
        except Exception as e:  # the baseline is reported, never the measurement
            out["cpu_baseline"] = {"error": str(e)}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
