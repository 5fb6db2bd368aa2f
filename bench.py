"""Benchmark of the explicit step on one MI355X.

Workload (N=1): one 3-D block 256x128x128 with 9 synthetic species, plain
FP64, Van Albada MUSCL + AUSM+ + central viscous flux + Wilke transport,
SSP-RK3 with the global CFL time step (the "plain FP64 drop-in" config).
A step is one RK3 step (3 stages) over the whole block; inputs are resident
in HBM before the timed region.

Prints one JSON line: metric/value in Mcell-stage updates/s (interior cells x
stages / wall time) of the plain FP64 step, the roofline object of the
dominant kernel (the RHS), the CPU baseline (the oracle on the host cores,
bounded sample), and beside them the changed forms: mixed precision,
graph-captured autotuned execution, and the 64-block lattice (about 40M
points) run block by block and batched.

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
# MI355X vector peaks: FP64 78.6 TFLOP/s; FP32 157.3 TFLOP/s (the mixed RHS computes in FP32)
VALU_PEAK_TFLOPS = {"fp64": 78.6, "mixed": 157.3}
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


def cpu_baseline(case, threads, steps=5):
    """The oracle on the host cores: the median of `steps` SSP-RK3 steps of
    the same block (about 3 s each on 16 threads, 15 s in all)."""
    import statistics
    from oracle import pyoracle
    pyoracle.build()
    used = pyoracle.set_threads(threads)
    o = pyoracle.Oracle(case)
    o.upload_initial()
    o.init_conservative()
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        o.step(1)
        times.append(time.perf_counter() - t0)
    dt = statistics.median(times)
    cs = case.n_cells * 3
    o.close()
    return {"value": cs / dt / 1e6, "unit": "Mcell-stage updates/s", "cores": used, "kind": "port",
            "sample": f"median of {steps} SSP-RK3 steps (3 stages each) of the same 256x128x128 9-species block, "
                      f"{cs} cell-stages in {dt:.2f} s per step, OpenMP on {used} host threads (the host share "
                      f"of one GPU)"}


def measure_traffic(precision="fp64", timeout_s=240):
    """HBM bytes of one RHS evaluation (three direction launches, default
    tiles) measured in this run: two rocprofv3 counter passes over a child
    `bench.py --rhs-only 2` (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch),
    each scaled by the factor the same counter reads for a plain device copy
    of known size (profiles/rhs_c1_<precision>_traffic.json: a 2 GiB float64
    copy, measured once with `bench.py --copy-calibration`).  Returns bytes per
    stage, or None when the profiler is not available."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if os.path.exists("/opt/rocm/bin/rocprofv3")
                                         else None)
    tpath = os.path.join(ROOT, "profiles", f"rhs_c1_{precision}_traffic.json")
    if prof is None or not os.path.exists(tpath):
        return None
    with open(tpath) as f:
        cal = json.load(f)["calibration"]
    scale = {"FETCH_SIZE": cal["read_bytes_per_FETCH_SIZE_KiB"], "WRITE_SIZE": cal["write_bytes_per_WRITE_SIZE_KiB"]}
    total = 0.0
    env = dict(os.environ, TMPDIR="/tmp")
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        out = tempfile.mkdtemp(prefix="cfx_traffic_", dir="/tmp")
        cmd = ["timeout", "-s", "KILL", str(timeout_s), prof, "--kernel-trace", "--pmc", counter, "-d", out, "-o", "p",
               "--output-format", "csv", "--", sys.executable, os.path.join(ROOT, "bench.py"), "--rhs-only", "2",
               "--precision", precision]
        r = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True)
        files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
        if r.returncode != 0 and not files:
            shutil.rmtree(out, ignore_errors=True)
            return None
        kib = 0.0
        for fn in files:
            with open(fn) as f:
                for row in csv.DictReader(f):
                    if "rhs_pencil_kernel" in row.get("Kernel_Name", "") and row.get("Counter_Name") == counter:
                        kib += float(row["Counter_Value"])
        shutil.rmtree(out, ignore_errors=True)
        if kib <= 0:
            return None
        total += kib / 2 * scale[counter]  # two RHS evaluations
    return total


def prepare(case, args, name, precision="fp64", batched=True, graph=False, autotune=False, mlp=False, **_):
    """A context of one execution form, ready to step: surrogate trained (with
    mlp), tiles autotuned (with autotune), warmed up and health-checked."""
    from combustioncodegpu_amd import abi
    from combustioncodegpu_amd.lib import Solver
    case.params["precision"] = abi.MIXED if precision == "mixed" else abi.FP64
    t0 = time.perf_counter()
    g = Solver(case)   # cfx_create: ghost grid, device metrics, tables, schedules
    g.sync()
    create_ms = 1e3 * (time.perf_counter() - t0)
    case.params["precision"] = abi.FP64
    info = {"create_ms": round(create_ms, 1)}
    if mlp:
        log(f"[bench] {name}: training the transport surrogate")
        t0 = time.perf_counter()
        info["mlp_report"] = g.mlp_train()
        info["mlp_report"]["wall_s"] = round(time.perf_counter() - t0, 2)
        g.set_transport(abi.TRANSPORT_MLP)
    g.set_execution(batched=batched, graph=graph)
    g.upload_initial()
    g.init_conservative()
    g.sync()
    if autotune:
        log(f"[bench] {name}: autotune")
        var, per, ms = g.autotune(reps=3)
        g.upload_initial()
        g.init_conservative()
        info["tiles"] = {"variant": var, "persist": per, "rhs_ms_per_direction": [round(x, 4) for x in ms]}
    log(f"[bench] {name}: warmup")
    g.step(args.warmup)
    g.sync()
    rc, nfail = g.check()
    if rc != 0:
        log(f"[bench] {name}: health check failed rc={rc} newton_fail={nfail}")
        g.close()
        return None, None
    return g, info


def measure_group(case, args, specs, rounds=5, traffic=None):
    """Throughput of several execution forms of one case.  Every form is set
    up first; the timed region then interleaves them: in each of `rounds`
    rounds every form runs its share of args.steps steps between
    synchronisations (wall clock), so that all forms see the same clock and
    thermal history of the GPU; each form's total is exactly args.steps
    steps.  Per-kernel times come from HIP events in a separate profiled run
    of 3 eager steps.  specs: (name, keyword arguments) per form; mlp trains
    the surrogate on the GPU first (outside the timed region)."""
    ctxs = {}
    for name, kw in specs:
        g, info = prepare(case, args, name, **kw)
        if g is None:
            for c, _ in ctxs.values():
                c.close()
            return None
        ctxs[name] = (g, info)
    n_rounds = max(1, min(rounds, args.steps))
    shares = [args.steps // n_rounds + (1 if r < args.steps % n_rounds else 0) for r in range(n_rounds)]
    wall = {name: 0.0 for name in ctxs}
    ev = {name: 0.0 for name in ctxs}
    log(f"[bench] timed region: {args.steps} steps per form in {n_rounds} interleaved rounds")
    for k in shares:
        for name, (g, _) in ctxs.items():
            g.sync()
            t0 = time.perf_counter()
            ev[name] += g.time_steps(k)
            g.sync()
            wall[name] += time.perf_counter() - t0
    outs = {}
    for (name, kw), (g, info) in zip(specs, ctxs.values()):
        precision = kw.get("precision", "fp64")
        value = case.n_cells * 3 * args.steps / wall[name] / 1e6
        # per-kernel durations from HIP events on the context's stream (eager)
        g.set_profiling(True)
        g.time_steps(3)
        kt = g.kernel_times()
        g.set_profiling(False)
        n_ops, n_lev = g.schedule_info()
        out = {
            "name": name, "value": round(value, 3), "ms_per_step": round(1e3 * wall[name] / args.steps, 4),
            "event_ms_per_step": round(ev[name] / args.steps, 4),
            "kernel_ms_per_step": {k: round(v[0] / 3, 4) for k, v in kt.items()},
            "launches_per_step": {k: round(v[1] / 3, 1) for k, v in kt.items()},
            "execution": {"batched": kw.get("batched", True), "graph": kw.get("graph", False),
                          "blocks": len(case.blocks), "ghost_fill_ops_per_stage": n_ops,
                          "ghost_fill_levels_per_stage": n_lev},
        }
        out["create_ms"] = info["create_ms"]
        if "tiles" in info:
            out["rhs_tiles"] = info["tiles"]
        if "mlp_report" in info:
            out["mlp"] = {"training": info["mlp_report"], "counters": g.mlp_stats()}
        # one timed RHS event per direction launch (batched: all blocks at once)
        rhs_ms, rhs_n = kt["rhs"]
        stages = 3 * 3
        rhs_stage_s = rhs_ms / stages / 1e3
        if kw.get("roofline", True):
            bpc = BYTES_PER_CELL_STAGE_FP64 if precision == "fp64" else BYTES_PER_CELL_STAGE_MIXED
            achieved = bpc * case.n_cells / rhs_stage_s / 1e9
            # HBM bytes per RHS evaluation measured in this run with the
            # default tiles (None for autotuned tiles, whose traffic differs)
            tr = (traffic or {}).get(precision) if "tiles" not in info else None
            out["roofline"] = {
                "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": tr,
                "traffic_bytes_per_cell_stage": round(tr / case.n_cells, 1) if tr else None,
                "kernel": "rhs_pencil_kernel (LDS-staged RHS, one launch per face direction)",
                "kernel_avg_ms_per_stage": round(rhs_stage_s * 1e3, 4),
                "launches_per_stage": round(rhs_n / stages, 2), "bytes_per_cell_stage": round(bpc, 2)}
            # the same kernel against the vector ALU: FP64 / FP32 flops per stage
            # counted with rocprofv3 VALU instruction counters on the default tiles
            fpath = os.path.join(ROOT, "profiles", f"rhs_c1_{precision}_flops.json")
            if os.path.exists(fpath) and "tiles" not in info and case.n_cells == 256 * 128 * 128:
                with open(fpath) as f:
                    flops = json.load(f)["flops_per_stage"]
                peak = VALU_PEAK_TFLOPS[precision]
                tf = flops / rhs_stage_s / 1e12
                out["roofline_valu"] = {
                    "bound": "valu", "achieved": round(tf, 3), "peak": peak, "unit": "TFLOP/s",
                    "frac": round(tf / peak, 5), "flops_per_cell_stage": round(flops / case.n_cells, 1),
                    "source": os.path.relpath(fpath, ROOT)}
        out["dt"] = g.get_dt()
        out["residual"] = [float(x) for x in g.residual()]
        g.close()
        outs[name] = out
    return outs


def lusgs_form(case, args, cfl=5.0, steps=10):
    """The implicit LU-SGS step (local time step, CFL 5) on the same block,
    FP64, batched and graph-captured: time per step (one stage: RHS, local
    dt, coefficient kernel, forward and backward hyperplane sweeps, update,
    ghost fill, properties) and per-kernel times; Mcell-updates/s counts
    interior cells x steps."""
    import copy
    from combustioncodegpu_amd.lib import Solver
    c = copy.copy(case)
    c.params = dict(case.params, rk_stages=1, lusgs=1, lusgs_omega=1.0, cfl=cfl)
    g = Solver(c)
    g.set_execution(batched=True, graph=True)
    g.upload_initial()
    g.init_conservative()
    g.step(2)
    g.sync()
    res0 = g.residual()
    ms = g.time_steps(steps)
    res1 = g.residual()
    g.set_profiling(True)
    g.time_steps(2)
    kt = g.kernel_times()
    rc, nfail = g.check()
    g.close()
    return {"ms_per_step": round(ms / steps, 4), "value": round(case.n_cells * steps / ms / 1e3, 3),
            "unit": "Mcell updates/s (one implicit stage per step)", "cfl": cfl, "health_rc": rc,
            "kernel_ms_per_step": {k: round(v[0] / 2, 4) for k, v in kt.items()},
            "launches_per_step": {k: round(v[1] / 2, 1) for k, v in kt.items()},
            "residual_before_after": [[float(x) for x in res0], [float(x) for x in res1]]}


def mixed_errors(case, checkpoints=(1, 10, 100), mlp=True):
    """Per-field error of the mixed path (and of the mixed path with the
    surrogate transport) against the FP64 path after 1, 10 and 100 RK3 steps
    (max over cells; rho, p, T relative to their max, velocity as a vector
    relative to max |V|, Y absolute; L2 relative for rho and T)."""
    import numpy as np
    from combustioncodegpu_amd import abi
    from combustioncodegpu_amd.lib import Solver
    ctx = {}
    forms = ("fp64", "mixed", "mixed_mlp") if mlp else ("fp64", "mixed")
    for prec in forms:
        case.params["precision"] = abi.FP64 if prec == "fp64" else abi.MIXED
        ctx[prec] = Solver(case)
        if prec == "mixed_mlp":
            ctx[prec].mlp_train()
            ctx[prec].set_transport(abi.TRANSPORT_MLP)
        ctx[prec].upload_initial()
        ctx[prec].init_conservative()
    case.params["precision"] = abi.FP64
    out = {f: {} for f in forms if f != "fp64"}
    done = 0
    for n in checkpoints:
        for c in ctx.values():  # one context at a time: no kernels of two contexts overlap
            c.step(n - done)
            c.sync()
        done = n
        b = {k: c.download(0, abi.BV)[:, 2:-2, 2:-2, 2:-2] for k, c in ctx.items()}
        y = {k: c.download(0, abi.YK)[:, 2:-2, 2:-2, 2:-2] for k, c in ctx.items()}
        r = b["fp64"]
        vmax = np.sqrt((r[1:4] ** 2).sum(axis=0)).max()
        for f in out:
            m = b[f]
            out[f][f"after_{n}_steps"] = {
                "rho": float(np.abs(m[0] - r[0]).max() / np.abs(r[0]).max()),
                "velocity": float(np.sqrt(((m[1:4] - r[1:4]) ** 2).sum(axis=0)).max() / vmax),
                "p": float(np.abs(m[4] - r[4]).max() / np.abs(r[4]).max()),
                "T": float(np.abs(m[5] - r[5]).max() / np.abs(r[5]).max()),
                "Y": float(np.abs(y[f] - y["fp64"]).max()),
                "rho_l2": float(np.sqrt(np.mean((m[0] - r[0]) ** 2)) / np.sqrt(np.mean(r[0] ** 2))),
                "T_l2": float(np.sqrt(np.mean((m[5] - r[5]) ** 2)) / np.sqrt(np.mean(r[5] ** 2))),
            }
    if "mixed_mlp" in ctx:
        out["mixed_mlp"]["surrogate_counters"] = ctx["mixed_mlp"].mlp_stats()
    for c in ctx.values():
        c.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 FETCH/WRITE passes of the RHS")
    ap.add_argument("--no-lusgs", action="store_true", help="skip the implicit LU-SGS form")
    ap.add_argument("--no-c3", action="store_true", help="skip the 64-block batched workload")
    ap.add_argument("--no-c4", action="store_true", help="skip the 384x256x256 surrogate workload")
    ap.add_argument("--no-mlp", action="store_true", help="skip every surrogate-transport form")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--size", type=str, default="256,128,128")
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--rhs-only", type=int, default=0,
                    help="profiling mode: run this many RHS launches on the workload and exit")
    ap.add_argument("--steps-only", type=int, default=0,
                    help="profiling mode: run this many steps of the workload's changed form and exit")
    ap.add_argument("--workload", choices=["c1", "c3"], default="c1")
    ap.add_argument("--rhs-variant", type=int, default=-1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--mlp", action="store_true",
                    help="profiling mode: use the surrogate transport (trained first with the default config)")
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
    if args.rhs_only or args.steps_only:
        log(f"[bench] generating workload {args.workload}")
        case = cases.c1(mx, my, mz) if args.workload == "c1" else cases.c3()
        case.params["precision"] = abi.MIXED if args.precision == "mixed" else abi.FP64
        g = Solver(case)
        if args.mlp:
            g.mlp_train()
            g.set_transport(abi.TRANSPORT_MLP)
        g.upload_initial()
        g.init_conservative()
        if args.rhs_variant >= 0:
            g.set_rhs_variant(args.rhs_variant)
        if args.rhs_only:
            for _ in range(args.rhs_only):
                g.rhs(-1)
        else:
            g.set_execution(batched=True, graph=True)
            g.autotune(reps=2)
            g.step(args.steps_only)
        g.sync()
        log(f"[bench] {args.workload} {args.precision}: profiling run done")
        g.close()
        return 0

    traffic = {}
    if not args.no_traffic and (mx, my, mz) == (256, 128, 128):
        for prec in ("fp64", "mixed"):
            log(f"[bench] HBM traffic of the {prec} RHS (two rocprofv3 counter passes)")
            traffic[prec] = measure_traffic(prec)
    log(f"[bench] generating workload {mx}x{my}x{mz}, 9 species")
    t0 = time.perf_counter()
    case = cases.c1(mx, my, mz)
    log(f"[bench] generated in {time.perf_counter() - t0:.1f} s")
    # the plain FP64 drop-in: eager launches as the reference's Driver issues them
    specs = [("fp64", dict(precision="fp64", batched=False, graph=False)),
             ("mixed", dict(precision="mixed", batched=False, graph=False)),
             ("mixed_graph_tuned", dict(precision="mixed", batched=True, graph=True, autotune=True)),
             ("fp64_graph_tuned", dict(precision="fp64", batched=True, graph=True, autotune=True))]
    if not args.no_mlp:
        specs.append(("mixed_graph_tuned_mlp", dict(precision="mixed", batched=True, graph=True, autotune=True,
                                                     mlp=True)))
    forms = measure_group(case, args, specs, traffic=traffic)
    if forms is None:
        return 3
    lusgs_out = None if args.no_lusgs else lusgs_form(case, args)
    log("[bench] mixed-precision error against the FP64 path")
    errors = mixed_errors(case, mlp=not args.no_mlp)
    c4_out = None
    if not args.no_mlp and not args.no_c4:
        log("[bench] generating the 384x256x256 surrogate workload")
        c4 = cases.c1(384, 256, 256)
        c4_forms = measure_group(c4, args, [
            ("fp64", dict(precision="fp64", batched=False, graph=False, roofline=False)),
            ("mixed_graph_tuned", dict(precision="mixed", batched=True, graph=True, autotune=True, roofline=False)),
            ("mixed_graph_tuned_mlp", dict(precision="mixed", batched=True, graph=True, autotune=True,
                                           roofline=False, mlp=True))])
        if c4_forms is None:
            return 3
        c4_out = {
            "workload": "3-D block 384x256x256, 9 species, mixed precision, graph-captured, autotuned; "
                        "Wilke vs surrogate transport",
            "cells": c4.n_cells,
            "forms": {k: {kk: v[kk] for kk in ("value", "ms_per_step", "kernel_ms_per_step") + tuple(
                x for x in ("rhs_tiles", "mlp", "create_ms") if x in v)} for k, v in c4_forms.items()},
            "surrogate_speedup_vs_wilke": round(c4_forms["mixed_graph_tuned_mlp"]["value"]
                                                / c4_forms["mixed_graph_tuned"]["value"], 3),
            "surrogate_props_ms_vs_wilke": [c4_forms["mixed_graph_tuned_mlp"]["kernel_ms_per_step"]["props"],
                                            c4_forms["mixed_graph_tuned"]["kernel_ms_per_step"]["props"]],
            "speedup_vs_fp64": round(c4_forms["mixed_graph_tuned_mlp"]["value"] / c4_forms["fp64"]["value"], 3),
        }
        del c4
    c3_out = None
    if not args.no_c3:
        log("[bench] generating the 64-block lattice workload")
        t0 = time.perf_counter()
        c3 = cases.c3()
        log(f"[bench] generated {c3.n_cells} points in {len(c3.blocks)} blocks in {time.perf_counter() - t0:.1f} s")
        c3_forms = measure_group(c3, args, [
            ("fp64_per_block", dict(precision="fp64", batched=False, graph=False, roofline=False)),
            ("fp64_batched_graph_tuned", dict(precision="fp64", batched=True, graph=True, autotune=True,
                                              roofline=False)),
            ("mixed_batched_graph_tuned", dict(precision="mixed", batched=True, graph=True, autotune=True,
                                               roofline=False))])
        if c3_forms is None:
            return 3
        sizes = sorted({(b.mx, b.my, b.mz) for b in c3.blocks})
        c3_out = {
            "workload": f"{len(c3.blocks)} blocks on a 4x4x4 face-matched lattice, extents drawn from [16,160] "
                        f"(smallest {sizes[0]}, largest {sizes[-1]}), {c3.n_cells} points, 9 species",
            "cells": c3.n_cells,
            "forms": {k: {kk: v[kk] for kk in ("value", "ms_per_step", "kernel_ms_per_step", "launches_per_step",
                                               "execution", "create_ms") + (("rhs_tiles",) if "rhs_tiles" in v else ())}
                      for k, v in c3_forms.items()},
            "batched_speedup_vs_per_block_fp64": round(c3_forms["fp64_batched_graph_tuned"]["value"]
                                                       / c3_forms["fp64_per_block"]["value"], 3),
            "mixed_batched_speedup_vs_per_block_fp64": round(c3_forms["mixed_batched_graph_tuned"]["value"]
                                                             / c3_forms["fp64_per_block"]["value"], 3),
        }
        del c3
    head = forms["fp64"]
    best = max((k for k in forms if k.startswith("mixed")), key=lambda k: forms[k]["value"])
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
        "dtype": "f64",
        "data": "synthetic (seeded smooth T/p/u fields, smoothed Dirichlet mass fractions, 9 synthetic species)",
        "config": {"workload": "plain FP64 explicit step, 3-D block 256x128x128, 9 species, Van Albada MUSCL, "
                               "AUSM+, viscous + Wilke, SSP-RK3 global CFL 0.5",
                   "cells": case.n_cells, "event_ms_per_step": head["event_ms_per_step"]},
        "roofline": head["roofline"],
        "roofline_valu": head.get("roofline_valu"),
        "kernel_ms_per_step": head["kernel_ms_per_step"],
        "create_ms": head.get("create_ms"),
        "variants": {
            **{k: {kk: v[kk] for kk in ("value", "ms_per_step", "kernel_ms_per_step") + tuple(
                x for x in ("rhs_tiles", "mlp", "roofline", "roofline_valu", "create_ms") if x in v)}
                for k, v in forms.items()
                if k != "fp64"},
            "changed_form_speedup_vs_fp64": {k: round(forms[k]["value"] / head["value"], 3)
                                             for k in forms if k != "fp64"},
            "best_changed_form": best,
            "mixed_error_vs_fp64": errors,
            "c3_batched": c3_out,
            "c4_surrogate": c4_out,
            "fp64_lusgs": lusgs_out,
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
