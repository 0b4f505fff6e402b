#!/usr/bin/env python3
"""Throughput of the MI355X LBM hot path (MLUPS and %HBM) on the 33.5M-cell case.

Workload (one step = one D3Q19 collide+stream pass over the whole lattice):
512x256x256 cells, FP32, TRT (magic 3/16) + Smagorinsky Cs 0.1, U0 0.08 along
+x, convective outlet, no body force, tau 0.51152 (Re 1000 on the minor-axis
diameter, 48 cells), a synthetic ellipsoid STL (semi-axes 60/24/24 cells,
199,080 triangles, vertex jitter U(-0.05, 0.05) cells from a seeded generator)
voxelized by the BVH three-ray vote at the cell centres.  The lattice and its
mask are resident in HBM before the timed region.

Each timed run warms up, restarts the lattice from its initial state and times
steps 1..K of the case; throughput and kernel time count only the steps the call
advanced (`steps_done`: the reference's step loop breaks on a NaN storm or a
diverged mean density, and the launches after a break return at once).

Prints one JSON line.  `value` is MLUPS = cells * steps_done / seconds / 1e6
over the timed call; `roofline` uses the step kernel's average duration from one
HIP event pair on the solver's stream around the timed back-to-back launches
(an upper bound: it includes the dispatch gaps) and the algorithmic 177 bytes per cell update
(19 FP32 loads + 19 stores + previous-step velocity read and current write for
the Smagorinsky stencil + 1 flag byte).  `roofline.traffic` is the HBM bytes
per step-kernel launch from two rocprofv3 counter passes (FETCH_SIZE, then
WRITE_SIZE) over a short run of the same case, each scaled by what the same
counter reads for a plain copy kernel of known size in the same process.
`cpu_baseline` times the oracle's C++ restatement of the reference solver (the
reference itself refuses lattices above 20 Mi cells) on the host cores for a
bounded sample of steps of the same case, with the CPU model; the reference's own
solver, built from its sources, is timed on the 256x128x104 case beside it.
`secondary` holds the other precisions and cases (33.5M cells FP16/BF16, the
256x128x104 case, the FP64 sphere, the coupled thermal lattice), unless
--no-studies the host-driver leg and the short precision studies, and with
--studies the long ones (33.5M cells for 20k steps, the Nusselt number after
300k steps); `served`, `served_fp16_eq` and `served_fp16` the served-mode stream
in FP32, equilibrium-shifted FP16 and FP16 storage (saturation solves/s against
one-at-a-time solves, p50/p99 latency at 0.7 of the saturation rate; the FP16
line on the first 64 requests only).
"""
from __future__ import annotations

import argparse
import csv
import ctypes as C
import glob
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "aeros-engine_amd"))

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {("fp32", True): 177, ("fp32", False): 153, ("fp16", True): 89, ("bf16", True): 89,
                  ("fp64", False): 305, ("fp64", True): 353, ("thermal", False): 217}
COPY_BYTES = 256 << 20   # the counter-scaling copy


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_gpu(A, cfg, steps, warmup, label, T=None):
    """W warm-up steps, then the lattice restarts from its initial state and the K
    timed steps are steps 1..K of the case.  A step that breaks the reference's step
    loop (NaN storm, or mean rho outside 0.8..1.2) ends the call there and the later
    launches of the call return at once; throughput and the kernel average are taken
    over the steps the call really advanced."""
    t0 = time.time()
    L = A.Lattice(A.default_config(**cfg))
    if T is not None:
        L.thermal_upload(T)
    log(f"[{label}] lattice {L.dims} ready in {time.time() - t0:.1f} s")
    L.step(warmup)
    L.reset()
    if T is not None:
        L.thermal_upload(T)
    sc0 = L.scalars()
    # The K timed steps; the step kernel's duration from one HIP event pair on
    # the solver's stream around the K back-to-back launches (a pair around
    # every launch would itself hold back each next dispatch by several us)
    L.kernel_timing(-1)
    L.kernel_timing(2)
    t1 = time.perf_counter()
    L.step(steps)
    dt = time.perf_counter() - t1
    kms, launches = L.kernel_timing(0)
    sc = L.scalars()
    done = sc["step"] - sc0["step"]
    resets = sc["resets"] - sc0["resets"]
    cells = L.ncells
    res = dict(dims=L.dims, cells=cells, seconds=dt, mlups=cells * done / dt / 1e6,
               kernel_ms=kms / max(1, done), launches=launches, steps_done=done, resets=resets,
               avg_rho=sc["avg_rho"], max_u=sc["max_u_lb"], step=sc["step"], nusselt=sc.get("nusselt"))
    L.close()
    log(f"[{label}] {res['mlups']:.1f} MLUPS, step kernel {res['kernel_ms']:.5f} ms, {done} of {steps} steps "
        f"advanced, {resets} reset(s), max |u| {sc['max_u_lb']:.4f}")
    return res


def c1_results(A, W):
    """The C1 sphere's reported results: the FP64 rho sum after 120 steps and
    Cd after 2000 steps (control volume 0.45/0.40/0.40 of the domain around its
    centre, which holds the sphere; A = projected solid cells along x)."""
    L = A.Lattice(A.default_config(**W.c1_sphere_fp64()))
    L.step(120)
    rho_sum, _ = L.checksums()
    L.step(1880)
    fx, _, _ = L.cv_force(0.45, 0.40, 0.40)
    area = L.projected_area(0)
    cd = fx / (0.5 * 0.05 * 0.05 * area)
    L.close()
    log(f"[c1 fp64] rho sum after 120 steps {rho_sum:.12g}; Cd after 2000 steps {cd:.4f} (A {area:.0f} cells)")
    return {"rho_sum_120": float("%.12g" % rho_sum), "cd_2000": round(cd, 4), "area_cells": area}


def c3_nusselt(A, W, steps=300000, chunk=20000, precisions=("fp64", "fp32", "fp32_fp64c", "fp16")):
    """The C3 Rayleigh-Benard box run to steady state in each precision from the same
    seeded start: Nu at `steps` steps and its change over the last `chunk` steps.  The
    FP64 run stands for the FP64 oracle, which the FP64 kernel matches within 1e-12 and
    which cannot itself run 1.05M cells for 300k steps on the host."""
    out = {}
    for prec in precisions:
        cfg, T = W.c3_rayleigh_benard()
        cfg["precision"] = prec
        L = A.Lattice(A.default_config(**cfg))
        L.thermal_upload(T)
        t0 = time.perf_counter()
        nus = []
        for _ in range(steps // chunk):
            L.step(chunk)
            nus.append(float(L.scalars()["nusselt"]))
        sc = L.scalars()
        L.close()
        out[prec] = {"nu": round(nus[-1], 6), "dnu_last_chunk": float("%.3g" % (nus[-1] - nus[-2])),
                     "max_u": float("%.4g" % sc["max_u_lb"]), "steps": sc["step"],
                     "seconds": round(time.perf_counter() - t0, 1)}
        log(f"[c3 nu] {prec}: {out[prec]}")
    ref = out["fp64"]["nu"]
    for prec in precisions[1:]:
        out[prec]["rel_err_vs_fp64"] = float("%.4g" % ((out[prec]["nu"] - ref) / ref))
    out["note"] = (f"{steps} steps of 128x64x128, Ra 2500, Pr 0.71, tau 0.6, T linear + seeded U(-1e-3, 1e-3); "
                   "Nu sampled every %d steps" % chunk)
    return out


def precision_study(A, W, scale=2.0, re=None, steps=20000, window=10000, every=100,
                    precisions=("fp64", "fp32", "fp32_fp64c", "fp16", "bf16", "fp16_eq", "bf16_eq")):
    """One body, grid and Re in FP64, FP32, FP16 and BF16 storage from the same start:
    the drag coefficient (control volume 0.45/0.40/0.40 of the domain, A = projected
    solid cells along x) averaged over samples every `every` steps in the last `window`
    steps, and the velocity field's relative L2 over the fluid cells after `steps`, each
    against the FP64 run.  With the reference's open side faces the ellipsoid flows
    diverge after about 450 steps in every precision (the reference solver's own run of
    the 256x128x104 case does the same), so the study closes the y and z faces with
    free-slip walls.  scale 2, Re 1000 is the 33.5M-cell case; scale 1, Re 100 a steady
    companion whose drag settles, so its errors are those of the storage and not of an
    unsteady wake.  At scale 2 one flow-through is 512 / 0.08 = 6400 steps and the
    wake sheds with a period of about 3000 steps, so 20k steps with the drag averaged
    over the last 10k cover three flow-throughs and about three periods."""
    cfg0 = W.c2_config(A.voxelize_gpu, scale=scale, re=re)
    cfg0["free_slip"] = W.SIDE_WALLS
    fluid = (np.asarray(cfg0["voxels"]) == 0).reshape(-1)
    out, u64, cd64 = {}, None, None
    for prec in precisions:
        cfg = dict(cfg0, precision=prec)
        L = A.Lattice(A.default_config(**cfg))
        area = L.projected_area(0)
        t0 = time.perf_counter()
        L.step(steps - window)
        cds = []
        for _ in range(window // every):
            L.step(every)
            cds.append(L.cv_force(0.45, 0.40, 0.40)[0] / (0.5 * 0.08 * 0.08 * area))
        u = np.stack([L.field(k).reshape(-1)[fluid].astype(np.float64) for k in ("ux", "uy", "uz")])
        sc = L.scalars()
        dims = L.dims
        L.close()
        cd = float(np.mean(cds))
        rec = {"cd_mean": round(cd, 5), "cd_last": round(cds[-1], 5), "steps": sc["step"], "resets": sc["resets"],
               "seconds": round(time.perf_counter() - t0, 1)}
        if u64 is None:
            u64, cd64 = u, cd
        else:
            rec["cd_rel_err_vs_fp64"] = float("%.4g" % ((cd - cd64) / cd64))
            rec["u_rel_l2_vs_fp64"] = float("%.4g" % np.sqrt(((u - u64) ** 2).sum() / (u64 ** 2).sum()))
        out[prec] = rec
        log(f"[precision {dims[0]}x{dims[1]}x{dims[2]}] {prec}: {rec}")
    out["note"] = (f"{dims[0]}x{dims[1]}x{dims[2]}, ellipsoid {30 * scale:g}/{12 * scale:g}/{12 * scale:g}, "
                   f"tau {cfg0['tau']:.5f}, TRT+LES, convective outlet, free-slip y/z faces; Cd mean of "
                   f"{window // every} samples over steps {steps - window + every}..{steps}; L2 after {steps} steps")
    return out


def c1_precision(A, W, steps=4000, window=1000, every=50):
    """The C1 sphere (96x48x48, BGK, U0 0.05, free-slip side walls) in every storage:
    Cd averaged over the last `window` steps against FP64, and the velocity L2."""
    out, u64, cd64 = {}, None, None
    for prec in ("fp64", "fp32", "fp32_fp64c", "fp16", "bf16", "fp16_eq", "bf16_eq"):
        L = A.Lattice(A.default_config(**W.c1_sphere_fp64(precision=prec)))
        area = L.projected_area(0)
        L.step(steps - window)
        cds = []
        for _ in range(window // every):
            L.step(every)
            cds.append(L.cv_force(0.45, 0.40, 0.40)[0] / (0.5 * 0.05 * 0.05 * area))
        u = np.stack([L.field(k).reshape(-1).astype(np.float64) for k in ("ux", "uy", "uz")])
        L.close()
        cd = float(np.mean(cds))
        rec = {"cd_mean": round(cd, 5)}
        if u64 is None:
            u64, cd64 = u, cd
        else:
            rec["cd_rel_err_vs_fp64"] = float("%.4g" % ((cd - cd64) / cd64))
            rec["u_rel_l2_vs_fp64"] = float("%.4g" % np.sqrt(((u - u64) ** 2).sum() / (u64 ** 2).sum()))
        out[prec] = rec
    out["note"] = f"96x48x48 sphere D 12, Cd mean of {window // every} samples over the last {window} of {steps} steps"
    log(f"[precision c1] {out}")
    return out


def line(r, bpc):
    return {"mlups": round(r["mlups"], 1), "kernel_ms": round(r["kernel_ms"], 5),
            "steps_done": r["steps_done"], "resets": r["resets"],
            "pct_hbm": round(r["mlups"] * 1e6 * bpc / (HBM_PEAK_GBS * 1e9), 4),
            "achieved_gbs": round(bpc * r["cells"] / (r["kernel_ms"] * 1e-3) / 1e9, 1),
            "bytes_per_cell": bpc, "cells": r["cells"]}


def cpu_model():
    """The host CPU's model name (with its family / model / stepping numbers, which
    name the generation where a virtual machine shortens the name) and the threads
    this process may use."""
    model = "unknown"
    info = {}
    try:
        for ln in open("/proc/cpuinfo"):
            k, _, v = ln.partition(":")
            k = k.strip()
            if k in ("model name", "cpu family", "model", "stepping") and k not in info:
                info[k] = v.strip()
            if not ln.strip() and info:
                break
    except OSError:
        pass
    if info:
        model = (f"{info.get('model name', 'unknown')} (family {info.get('cpu family', '?')}, "
                 f"model {info.get('model', '?')}, stepping {info.get('stepping', '?')})")
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:") or ln.startswith("BIOS Model name:"):
                v = ln.split(":", 1)[1].strip()
                if v and v not in model:
                    model += f"; lscpu {ln.split(':', 1)[0]}: {v}"
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return model, avail


def cpu_baseline(cfg, budget_s=15.0, max_steps=60):
    """The oracle (C++ restatement of the reference solver, same algorithm and AoS
    layout, OpenMP) on the host cores: one warm-up step, then single steps until
    budget_s has passed (at least 3, at most max_steps)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    O.build()
    model, avail = cpu_model()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or avail
    threads = O.threads(min(threads, avail, 16))
    t0 = time.perf_counter()
    L = O.OracleLattice(O.default_config(**cfg))
    t_init = time.perf_counter() - t0
    L.step(1)
    n = 0
    t0 = time.perf_counter()
    while n < max_steps and (n < 3 or time.perf_counter() - t0 < budget_s):
        L.step(1)
        n += 1
    dt = time.perf_counter() - t0
    cells = L.ncells
    dims = L.dims
    L.close()
    return dict(value=round(cells * n / dt / 1e6, 3), unit="MLUPS", cores=threads, kind="port",
                cpu_model=model, cpus_available=avail, init_s=round(t_init, 1),
                sample=f"{n} single steps ({dt:.1f} s) of the same {dims[0]}x{dims[1]}x{dims[2]} case after 1 "
                       f"warm-up step: the oracle (C++ restatement of the reference solver, same algorithm and AoS "
                       f"layout, OpenMP {threads} threads); the reference itself refuses lattices above 20 Mi cells")


def cpu_reference(cfg, threads, steps=320):
    """The reference solver itself on the same case: its own lbm.cpp built from its
    sources with its bench's flags (-O3, OpenMP) by oracle/ref.mk into oracle/_ref/ref_box,
    fed the same mask and tau; one warm-up step, then one call of `steps` steps.  None
    when that binary was not built (the reference sources were absent at build time)."""
    exe = ROOT / "oracle" / "_ref" / "ref_box"
    if not exe.exists():
        return None
    vox = np.ascontiguousarray(cfg["voxels"], dtype=np.uint8)
    nz, ny, nx = vox.shape
    with tempfile.TemporaryDirectory() as d:
        vox.tofile(os.path.join(d, "mask.u8"))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        r = subprocess.run([str(exe), "box_time", d, str(nx), str(ny), str(nz), repr(cfg["tau"]), str(steps)],
                           env=env, capture_output=True, text=True, timeout=600)
    m = re.search(r"box_time (\d+)x(\d+)x(\d+) steps (\d+) seconds (\S+) threads (\d+)", r.stdout)
    if r.returncode != 0 or not m:
        log(f"[cpu] reference run failed ({r.returncode}): {r.stderr[-400:]}")
        return None
    cells = int(m.group(1)) * int(m.group(2)) * int(m.group(3))
    n, sec, used = int(m.group(4)), float(m.group(5)), int(m.group(6))
    return dict(value=round(cells * n / sec / 1e6, 3), unit="MLUPS", cores=used, kind="reference",
                sample=f"{n} steps (one call, after 1 warm-up step) of the same {nx}x{ny}x{nz} case by the "
                       f"reference's own CPU solver (stepLBMCPU, -O3, OpenMP {used} threads)")


def cpu_others(A, W, budget_s=6.0):
    """The restatement on the other configurations, a few seconds each: C1 in FP64,
    the 256x128x104 case and the coupled thermal C3 in FP32."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O
    out = {}
    c3, c3T = W.c3_rayleigh_benard()
    for name, cfg, T in (("c1_fp64", W.c1_sphere_fp64(), None), ("c2_fp32", W.c2_config(A.voxelize), None),
                         ("c3_thermal_fp32", c3, c3T)):
        L = O.OracleLattice(O.default_config(**cfg))
        if T is not None:
            L.thermal_upload(T)
        L.step(1)
        n, t0 = 0, time.perf_counter()
        while n < 40 and (n < 2 or time.perf_counter() - t0 < budget_s):
            L.step(1)
            n += 1
        out[name] = {"mlups": round(L.ncells * n / (time.perf_counter() - t0) / 1e6, 3), "steps": n}
        L.close()
        log(f"[cpu] {name}: {out[name]}")
    return out


# ------------------------------------------------------------------ served mode
def solve_alone(A, W, cfg, precision, round_steps, tol, max_steps):
    """One request on its own lattice with the plain solver: one call per round of
    round_steps, the drag after each, retired at |dCd| < tol over a round or at
    max_steps.  (The server steps the same launches in ticks; without a divergence
    reset the populations, fields and drag are the same bits.)"""
    L = A.Lattice(A.default_config(**cfg, precision=precision))
    area = L.projected_area(0)
    u0 = float(np.float32(np.sqrt(np.float32(cfg["U0"]) * np.float32(cfg["U0"]))))
    q = 0.5 * u0 * u0 * float(area)
    cd, prev, rounds = 0.0, None, 0
    while True:
        L.step(round_steps)
        steps = L.scalars()["step"]   # the step counter, which a divergence reset moves back
        rounds += 1
        prev, cd = cd, L.cv_force(W.C5_CV_HALF, W.C5_CV_HALF, W.C5_CV_HALF)[0] / q
        if (rounds >= 2 and abs(cd - prev) < tol) or steps >= max_steps:
            break
    L.close()
    return cd, steps


def host_driver_leg(A, cfg, frames=40, steps=10, fields=("rho", "ux", "uy", "uz")):
    """The host stepping driver on the headline case: frames of `steps` lattice steps
    whose fields reach host memory.  Synchronous: update() then a read of each field.
    Asynchronous: the frame worker stages the fields on the device and copies them to a
    pinned slot while the next frame steps; the caller takes the newest finished frame."""
    L = A.Lattice(A.default_config(**cfg))
    L.step(20)
    t0 = time.perf_counter()
    for _ in range(frames):
        L.update(0.016, steps, 3)
        for f in fields:
            L.field(f)
    t_sync = time.perf_counter() - t0
    L.frames_start(fields)
    t0 = time.perf_counter()
    for _ in range(frames):
        L.frame_submit(0.016, steps, 3)
        L.frame_latest(copy=False)
    L.frames_wait()
    t_async = time.perf_counter() - t0
    last = L.frame_latest(copy=False)["frame"]
    L.frames_stop()
    L.close()
    res = {"frames": frames, "steps_per_frame": steps, "fields": list(fields),
           "sync_frames_per_s": round(frames / t_sync, 1), "async_frames_per_s": round(frames / t_async, 1),
           "async_over_sync": round(t_sync / t_async, 2), "last_frame": last}
    log(f"[host driver] {res}")
    return res


def served_leg(A, W, precision="fp32", n=256, base_k=24, round_steps=200, tick=20, tol=1e-4, max_steps=20000,
               seed=5150, load=0.7, slots=256):
    """The C5 stream: 256 seeded requests (32^3 .. 160x80x80, spheres / ellipsoids of
    6-20 cells, Re drawn from [50, 500], 30 % thermal with Ra in [1800, 5000], free-slip
    side walls), each retired at |dCd| < tol over a 200-step round or at max_steps.
    One-at-a-time: the first base_k requests solved back to back on the plain solver.
    Saturation: all n submitted at once; solves/s = n / makespan.  Latency: the n
    requests arrive as a Poisson stream at `load` times the saturation rate, and
    latency = completion - arrival (time to converged Cd, queueing included)."""
    reqs = W.c5_requests(n, seed)
    cfgs = [A.default_config(**cfg, precision=precision) for cfg, _ in reqs]
    t0 = time.perf_counter()
    alone = [solve_alone(A, W, cfg, precision, round_steps, tol, max_steps) for cfg, _ in reqs[:base_k]]
    t_alone = time.perf_counter() - t0
    base_rate = base_k / t_alone
    log(f"[served {precision}] one-at-a-time: {base_k} requests in {t_alone:.2f} s = {base_rate:.3f} solves/s")

    def run(gaps):
        srv = A.Server(precision=precision, round_steps=round_steps, tick_steps=tick, cd_tolerance=tol,
                       max_steps=max_steps, cv_half=W.C5_CV_HALF, max_cases=slots)
        ids = []
        t_next = srv.clock()
        for k in range(n):
            t_next += gaps[k]
            while srv.clock() < t_next:
                time.sleep(min(0.002, max(0.0, t_next - srv.clock())))
            ids.append(srv.submit(cfgs[k]))
        if not srv.wait_all(900.0):
            raise RuntimeError("served requests did not finish")
        res = [srv.poll(i) for i in ids]
        st = srv.stats()
        srv.close()
        makespan = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
        lat = np.array([r["t_done"] - r["t_submit"] for r in res])
        return res, st, makespan, lat

    res_s, st_s, span_s, _ = run(np.zeros(n))
    sat_rate = n / span_s
    log(f"[served {precision}] saturation: {n} requests in {span_s:.2f} s = {sat_rate:.3f} solves/s, "
        f"peak resident {st_s['peak_resident']}")
    lam = load * sat_rate
    gaps = np.random.Generator(np.random.PCG64(seed + 1)).exponential(1.0 / lam, size=n)
    res, st, span, lat = run(gaps)
    wait = np.array([r["t_admit"] - r["t_submit"] for r in res])
    same = sum(1 for k in range(base_k) if np.float64(res[k]["cd"]).tobytes() == np.float64(alone[k][0]).tobytes()
               and res[k]["steps"] == alone[k][1])
    done = [r for r in res if r["state_name"] == "done"]
    steps_done = np.array([r["steps"] for r in done]) if done else np.zeros(1)
    re_req = np.array([m["re"] for _, m in reqs])
    re_eff = np.array([m["re_eff"] for _, m in reqs])
    out = {"precision": precision, "requests": n,
           "saturation_solves_per_s": round(sat_rate, 3), "saturation_makespan_s": round(span_s, 2),
           "one_at_a_time_solves_per_s": round(base_rate, 3), "one_at_a_time_sample": base_k,
           "speedup_vs_one_at_a_time": round(sat_rate / base_rate, 2),
           "poisson_lambda_per_s": round(lam, 3), "poisson_load": load,
           "p50_s": round(float(np.percentile(lat, 50)), 3), "p99_s": round(float(np.percentile(lat, 99)), 3),
           "p50_wait_for_admission_s": round(float(np.percentile(wait, 50)), 4),
           "converged": len(done), "stopped_at_cap": sum(r["state_name"] == "limit" for r in res),
           "failed": sum(r["state_name"] == "failed" for r in res),
           "steps_to_converge_p50": int(np.percentile(steps_done, 50)),
           "steps_to_converge_max": int(steps_done.max()),
           "converged_at_saturation": sum(r["state_name"] == "done" for r in res_s),
           "identical_to_alone": f"{same}/{base_k}",
           "served_mlups": round(st["cell_steps"] / max(st["busy_seconds"], 1e-9) / 1e6, 1),
           "peak_resident": st["peak_resident"], "peak_resident_at_saturation": st_s["peak_resident"],
           "peak_bytes": st["peak_bytes"], "slots": slots,
           "re_requested": [round(float(re_req.min()), 1), round(float(re_req.max()), 1)],
           "re_realised": [round(float(re_eff.min()), 1), round(float(re_eff.max()), 1)],
           "re_realised_equals_requested": int(np.sum(np.abs(re_eff - re_req) < 1e-6 * re_req)),
           "rule": f"retire at |dCd| < {tol} over a {round_steps}-step round or at {max_steps} steps; "
                   f"{tick}-step ticks (requests join between ticks); tau = 1/2 + 3 U0 D / Re, at least 0.51"}
    log(f"[served {precision}] {out}")
    return out


# ------------------------------------------------------------------ HBM traffic
def device_copy(nbytes, reps):
    """A plain float4 copy kernel over nbytes, reps times (aeros-engine_amd/csrc/copy_probe.hip)."""
    lib = C.CDLL(str(ROOT / "aeros-engine_amd" / "build" / "libaeros_copyprobe.so"))
    lib.aeros_probe_copy.restype = C.c_int
    lib.aeros_probe_copy.argtypes = [C.c_size_t, C.c_int]
    rc = lib.aeros_probe_copy(nbytes, reps)
    if rc:
        raise RuntimeError(f"copy probe failed ({rc})")


PMC_CASES = {"c2": dict(scale=1.0, precision="fp32"), "c4_fp32": dict(scale=2.0, precision="fp32"),
             "c4_fp16": dict(scale=2.0, precision="fp16"), "c4_bf16": dict(scale=2.0, precision="bf16"),
             "c4_fp16_eq": dict(scale=2.0, precision="fp16_eq")}
HEADLINE_CASE = "c4_fp32"


def pmc_child(steps, case=HEADLINE_CASE):
    """What each counter pass runs: one case (the headline one unless named) for a few
    steps, then the copy."""
    import aeros_lbm as A
    import workloads as W
    L = A.Lattice(A.default_config(**W.c2_config(A.voxelize_gpu, **PMC_CASES[case])))
    L.step(steps)
    L.close()
    device_copy(COPY_BYTES, 4)
    print("pmc child done", flush=True)


def read_counter(d, name):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    step, copy = [], []
    for f in files:
        for r in csv.DictReader(open(f)):
            if r["Counter_Name"] != name:
                continue
            v = float(r["Counter_Value"])
            if "k_step" in r["Kernel_Name"]:
                step.append((int(r["Dispatch_Id"]), v))
            elif "k_copy_probe" in r["Kernel_Name"]:
                copy.append(v)
    step.sort()
    step = [v for _, v in step[1:]]   # the first launch collides in place
    if not step or not copy:
        raise RuntimeError(f"{name}: no step or copy dispatches in {d}")
    return sum(step) / len(step), sum(copy) / len(copy)


def measure_traffic(steps=20, timeout=240, case=HEADLINE_CASE):
    """HBM bytes per step-kernel launch from FETCH_SIZE and WRITE_SIZE passes."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return None, "rocprofv3 not found"
    res = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="aeros_pmc_", dir="/tmp")
        cmd = [exe, "--pmc", ctr, "-d", d, "-o", "run", "-f", "csv", "--", sys.executable, str(ROOT / "bench.py"),
               "--pmc-child", str(steps), "--pmc-case", case]
        log(f"[traffic] counter pass {ctr}")
        try:
            r = subprocess.run(cmd, cwd="/tmp", env={**os.environ, "TMPDIR": "/tmp"}, timeout=timeout,
                               capture_output=True, text=True)
            if "pmc child done" not in r.stdout:
                return None, f"{ctr} pass failed (rc {r.returncode}): {r.stderr[-300:]}"
            kib, copy_kib = read_counter(d, ctr)
        except (subprocess.TimeoutExpired, RuntimeError, OSError) as e:
            return None, f"{ctr} pass: {e}"
        finally:
            shutil.rmtree(d, ignore_errors=True)
        scale = (COPY_BYTES / 1024.0) / copy_kib
        res[ctr] = dict(kib=kib, copy_kib=copy_kib, scale=scale, bytes=kib * 1024.0 * scale)
    return res, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-secondary", action="store_true", help="skip the other precisions and cases")
    ap.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 counter passes")
    ap.add_argument("--no-served", action="store_true", help="skip the served-mode leg")
    ap.add_argument("--no-studies", action="store_true",
                    help="skip the host-driver leg and the short precision studies of the secondary section")
    ap.add_argument("--studies", action="store_true",
                    help="add the long studies: 33.5M cells for 20k steps in five storages, and the Rayleigh-Benard "
                         "Nusselt number after 300k steps in four")
    ap.add_argument("--pmc-child", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--pmc-case", default=HEADLINE_CASE, choices=sorted(PMC_CASES), help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args.pmc_child, args.pmc_case)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus}); "
              "the lattice would shard by x-slabs with a five-population halo per face.", file=sys.stderr)
        sys.exit(2)

    import aeros_lbm as A
    import workloads as W

    info = A.probe()
    if not info["usable"]:
        raise SystemExit(f"no usable gfx950 device: {info['status']}")
    log(f"device: {info['status']}")

    # the 33.5M-cell case: the x2 ellipsoid's mask by the device vote (bit-identical to the
    # host vote and to the reference's own, by the voxelizer tests)
    t0 = time.perf_counter()
    c4 = W.c2_config(A.voxelize_gpu, scale=2.0)
    log(f"C4 mask voxelized on the device in {time.perf_counter() - t0:.2f} s (mesh generation included)")
    r = run_gpu(A, c4, args.steps, args.warmup, "c4 fp32")
    bpc = BYTES_PER_CELL[("fp32", True)]
    achieved = bpc * r["cells"] / (r["kernel_ms"] * 1e-3) / 1e9
    nx, ny, nz = r["dims"]
    out = {
        "metric": "MLUPS (D3Q19 TRT+Smagorinsky FP32, 33.5M cells) and %HBM peak",
        "value": round(r["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["seconds"] * 1e3 / max(1, r["steps_done"]), 4),
        "steps_done": r["steps_done"],
        "resets": r["resets"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered ellipsoid STL, voxelized by the BVH three-ray vote)",
        "config": {"workload": f"{nx}x{ny}x{nz} D3Q19 TRT(3/16)+Smagorinsky Cs0.1 FP32, ellipsoid 60/24/24 cells "
                               f"(199,080 triangles), U0 0.08, tau {c4['tau']:.5f} (Re 1000), convective outlet",
                   "cells": r["cells"], "parallelism": "single GPU",
                   "timed": "steps 1..K of the case: W warm-up steps, then a restart from the initial state"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_step<float,float,float,LES>", "kernel_avg_ms": round(r["kernel_ms"], 5),
                     "timing": "one HIP event pair around the timed launches on the solver stream",
                     "bytes_per_cell": bpc},
        "pct_hbm_from_mlups": round(r["mlups"] * 1e6 * bpc / (HBM_PEAK_GBS * 1e9), 4),
    }
    c2 = None
    if not args.no_secondary:
        sec = {}
        tris = W.ellipsoid_mesh(semi=(30.0, 12.0, 12.0), centre=(256 * 0.3, 64.0, 52.0), seed=12345, jitter=0.05)
        grid = ((256, 128, 104), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
        t0 = time.perf_counter()
        m_host = A.voxelize(tris, *grid)
        t_host = time.perf_counter() - t0
        A.voxelize_gpu(tris, *grid)   # first call: device set-up
        t0 = time.perf_counter()
        m_gpu = A.voxelize_gpu(tris, *grid)
        t_gpu = time.perf_counter() - t0
        if not np.array_equal(m_host, m_gpu):
            raise RuntimeError("device voxelizer disagrees with the host voxelizer")
        c2 = W.c2_config(lambda t, d, b, c: m_gpu)   # the same mesh and grid as above
        sec["voxelize"] = {"mesh": "199,080-triangle jittered ellipsoid on 256x128x104",
                           "host_s": round(t_host, 4), "host_threads": os.cpu_count(), "gpu_s": round(t_gpu, 4),
                           "note": "BVH build on the host included in both; masks identical"}
        sec["c2_fp32"] = line(run_gpu(A, c2, 1000, 50, "c2 fp32"), BYTES_PER_CELL[("fp32", True)])
        sec["c2_fp32"]["tau"] = round(c2["tau"], 5)
        sec["c4_fp32"] = line(r, bpc)
        for prec in ("fp16", "bf16", "fp16_eq", "bf16_eq", "fp32_fp64c"):
            c4p = dict(c4, precision=prec)
            bpc4 = BYTES_PER_CELL[("fp32", True)] if prec == "fp32_fp64c" else BYTES_PER_CELL[("fp16", True)]
            sec[f"c4_{prec}"] = line(run_gpu(A, c4p, 200, 20, f"c4 {prec}"), bpc4)
        for prec in ("fp16", "bf16", "fp16_eq", "bf16_eq", "fp32_fp64c"):
            sec[f"{prec}_over_fp32"] = round(sec[f"c4_{prec}"]["mlups"] / r["mlups"], 3)
        sec["c1_fp64"] = line(run_gpu(A, W.c1_sphere_fp64(), 1000, 50, "c1 fp64"), BYTES_PER_CELL[("fp64", False)])
        sec["c1_fp64"].update(c1_results(A, W))
        c3cfg, c3T = W.c3_rayleigh_benard()
        sec["c3_thermal_fp32"] = line(run_gpu(A, c3cfg, 500, 20, "c3 thermal", T=c3T), BYTES_PER_CELL[("thermal", False)])
        sec["workloads"] = {"c4": "512x256x256 (33.5M cells), ellipsoid scaled x2, Re 1000, TRT+LES",
                            "c2": f"256x128x104 (3.41M cells), ellipsoid 30/12/12, Re 100 (tau {c2['tau']:.5f}), "
                                  "TRT+LES, 1000 steps",
                            "c1": "96x48x48 FP64 BGK sphere D=12, U0 0.05, zero-gradient outflow, "
                                  "free-slip side walls",
                            "c3": "128x64x128 coupled D3Q19+D3Q7 FP32, Rayleigh-Benard Ra 2500 Pr 0.71"}
        if not args.no_studies:
            sec["host_driver"] = host_driver_leg(A, c2)
            sec["c2_re100_precision_study"] = precision_study(A, W, scale=1.0, re=100.0, steps=16000, window=4000,
                                                              every=250)
            sec["c1_precision_study"] = c1_precision(A, W)
        if args.studies:   # about four minutes more
            sec["c4_precision_study"] = precision_study(A, W)
            sec["c3_nusselt"] = c3_nusselt(A, W)
        out["secondary"] = sec
    if not args.no_served:
        out["served"] = served_leg(A, W, "fp32")
        out["served_fp16_eq"] = served_leg(A, W, "fp16_eq")
        out["served_fp16"] = served_leg(A, W, "fp16", n=64, max_steps=6000, base_k=8)
    if not args.no_traffic:
        tr, err = measure_traffic()
        if tr:
            total = tr["FETCH_SIZE"]["bytes"] + tr["WRITE_SIZE"]["bytes"]
            out["roofline"]["traffic"] = round(total)
            out["roofline"]["traffic_detail"] = {
                "case": HEADLINE_CASE,
                "read_bytes": round(tr["FETCH_SIZE"]["bytes"]), "write_bytes": round(tr["WRITE_SIZE"]["bytes"]),
                "bytes_per_cell": round(total / r["cells"], 1),
                "over_algorithmic": round(total / (bpc * r["cells"]), 3),
                "fetch_kib_per_launch": round(tr["FETCH_SIZE"]["kib"], 1),
                "write_kib_per_launch": round(tr["WRITE_SIZE"]["kib"], 1),
                "copy_fetch_kib": round(tr["FETCH_SIZE"]["copy_kib"], 1), "copy_write_kib": round(tr["WRITE_SIZE"]["copy_kib"], 1),
                "copy_bytes": COPY_BYTES,
                "fetch_scale": round(tr["FETCH_SIZE"]["scale"], 4), "write_scale": round(tr["WRITE_SIZE"]["scale"], 4)}
        else:
            log(f"[traffic] not measured: {err}")
            out["roofline"]["traffic_note"] = err
    if not args.no_cpu:
        port = cpu_baseline(c4)
        log(f"[cpu] restatement on the headline case: {port['value']} MLUPS on {port['cores']} threads "
            f"({port['cpu_model']})")
        out["cpu_baseline"] = port
        if c2 is None:
            c2 = W.c2_config(A.voxelize_gpu)
        ref = cpu_reference(c2, port["cores"])
        if ref:
            log(f"[cpu] reference solver on 256x128x104: {ref['value']} MLUPS on {ref['cores']} threads")
            out["cpu_baseline"]["reference_256x128x104"] = ref
        if not args.no_secondary:
            out["cpu_baseline"]["others"] = cpu_others(A, W)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
