#!/usr/bin/env python3
"""Throughput of the MI355X LBM hot path (MLUPS and %HBM) on the 3.41M-cell case.

Workload (one step = one D3Q19 collide+stream pass over the whole lattice):
256x128x104 cells, FP32, TRT (magic 3/16) + Smagorinsky Cs 0.1, U0 0.08 along
+x, convective outlet, no body force, a synthetic ellipsoid STL (semi-axes
30/12/12 cells, 199,080 triangles, vertex jitter U(-0.05, 0.05) cells from a
seeded generator) voxelized by the BVH three-ray vote at the cell centres.
The lattice and its mask are resident in HBM before the timed region.

Prints one JSON line.  `value` is MLUPS = cells * steps / seconds / 1e6 over
the timed steps; `roofline` uses the step kernel's average duration from HIP
events on the solver's stream and the algorithmic 177 bytes per cell update
(19 FP32 loads + 19 stores + previous-step velocity read and current write for
the Smagorinsky stencil + 1 flag byte); `cpu_baseline` times the CPU
restatement on the host cores for a bounded number of steps of the same case.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "aeros-engine_amd"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {("fp32", True): 177, ("fp32", False): 153, ("fp16", True): 89, ("bf16", True): 89,
                  ("fp64", False): 305, ("fp64", True): 353}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_gpu(A, W, cfg, steps, warmup, label):
    t0 = time.time()
    L = A.Lattice(A.default_config(**cfg))
    log(f"[{label}] lattice {L.dims} ready in {time.time() - t0:.1f} s")
    L.step(warmup)
    L.kernel_timing(-1)
    L.kernel_timing(1)
    t1 = time.perf_counter()
    L.step(steps)
    dt = time.perf_counter() - t1
    kms, launches = L.kernel_timing(0)
    sc = L.scalars()
    cells = L.ncells
    res = dict(dims=L.dims, cells=cells, seconds=dt, mlups=cells * steps / dt / 1e6,
               kernel_ms=kms / max(1, launches), launches=launches, avg_rho=sc["avg_rho"],
               max_u=sc["max_u_lb"], step=sc["step"])
    L.close()
    return res


def cpu_baseline(W, cfg, budget_s=12.0, max_steps=40):
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    O.build()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    threads = O.threads(min(threads, 16))
    L = O.OracleLattice(O.default_config(**cfg))
    L.step(1)
    n = 0
    t0 = time.perf_counter()
    while n < max_steps and (n < 3 or time.perf_counter() - t0 < budget_s):
        L.step(1)
        n += 1
    dt = time.perf_counter() - t0
    cells = L.ncells
    L.close()
    return dict(value=round(cells * n / dt / 1e6, 3), unit="MLUPS", cores=threads, kind="port",
                sample=f"{n} steps of the same {L.dims[0]}x{L.dims[1]}x{L.dims[2]} case after 1 warm-up step, "
                       f"oracle (C++ restatement, OpenMP {threads} threads)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-secondary", action="store_true", help="skip the 33.5M-cell precision lines")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus}); "
              "see DESIGN.md for how the lattice would shard by x-slabs.", file=sys.stderr)
        sys.exit(2)

    import aeros_lbm as A
    import workloads as W

    info = A.probe()
    if not info["usable"]:
        raise SystemExit(f"no usable gfx950 device: {info['status']}")
    log(f"device: {info['status']}")

    t0 = time.time()
    c2 = W.c2_config(A.voxelize)
    log(f"C2 mask voxelized in {time.time() - t0:.1f} s ({int(c2['voxels'].sum())} solid cells)")
    r = run_gpu(A, W, c2, args.steps, args.warmup, "c2 fp32")
    bpc = BYTES_PER_CELL[("fp32", True)]
    achieved = bpc * r["cells"] / (r["kernel_ms"] * 1e-3) / 1e9
    out = {
        "metric": "MLUPS (D3Q19 TRT+Smagorinsky FP32, 3.41M cells) and %HBM peak",
        "value": round(r["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["seconds"] * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered ellipsoid STL, voxelized on the host)",
        "config": {"workload": "256x128x104 D3Q19 TRT(3/16)+Smagorinsky Cs0.1 FP32, ellipsoid 30/12/12 cells, "
                               "U0 0.08, tau 0.5192, convective outlet",
                   "cells": r["cells"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_step<float,float,float,LES>", "kernel_avg_ms": round(r["kernel_ms"], 5),
                     "bytes_per_cell": bpc},
        "pct_hbm_from_mlups": round(r["mlups"] * 1e6 * bpc / (HBM_PEAK_GBS * 1e9), 4),
    }
    if not args.no_secondary:
        sec = {}
        for prec in ("fp32", "fp16", "bf16"):
            c4 = W.c2_config(A.voxelize, scale=2.0, precision=prec)
            rr = run_gpu(A, W, c4, 200, 20, f"c4 {prec}")
            b = BYTES_PER_CELL[(prec, True)]
            sec[prec] = {"mlups": round(rr["mlups"], 1), "kernel_ms": round(rr["kernel_ms"], 4),
                         "pct_hbm": round(rr["mlups"] * 1e6 * b / (HBM_PEAK_GBS * 1e9), 4),
                         "achieved_gbs": round(b * rr["cells"] / (rr["kernel_ms"] * 1e-3) / 1e9, 1),
                         "bytes_per_cell": b}
        sec["fp16_over_fp32"] = round(sec["fp16"]["mlups"] / sec["fp32"]["mlups"], 3)
        sec["bf16_over_fp32"] = round(sec["bf16"]["mlups"] / sec["fp32"]["mlups"], 3)
        out["secondary"] = {"workload": "512x256x256 (33.5M cells), ellipsoid scaled x2, Re 1000", **sec}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(W, c2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
