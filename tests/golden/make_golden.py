#!/usr/bin/env python3
"""Regenerate the golden vectors under tests/golden/ from the reference's own binaries.

Runs the reference's bench and headless suites as built by ``oracle/ref.mk``
(into oracle/_ref/, from the sources under /root/reference) and records what
they print: the bench's sequential double checksums of the rho and u fields
for several (N, steps, thermal, LES) settings, and the numbers of the
Poiseuille / thermal / sphere-validation / solver suites.  Only numbers are
kept; the JSON files are the fixtures the CPU tests compare the oracle with.

    make -f oracle/ref.mk -j8 && python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
REF = ROOT / "oracle" / "_ref"
OUT = Path(__file__).resolve().parent

# (N, timed steps, thermal, LES): 30 warm-up steps are always added by the bench
BENCH_CASES = [
    (32, 120, 0, 1),
    (32, 120, 0, 0),
    (32, 120, 1, 1),
    (24, 60, 1, 0),
    (16, 200, 0, 1),
    (20, 170, 1, 1),
    (12, 400, 0, 1),
    (64, 100, 0, 1),
]

NUM = r"([-+]?\d+(?:\.\d+)?(?:[eE][-+]?\d+)?)"


def run(binary: str, *args: str, threads: int = 8) -> str:
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    os.makedirs("/tmp/tl", exist_ok=True)
    r = subprocess.run([str(REF / binary), *args], env=env, capture_output=True, text=True, timeout=1800)
    if r.returncode != 0:
        raise SystemExit(f"{binary} {args} failed ({r.returncode}):\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    return r.stdout


def grab(pattern: str, text: str, n: int = 1, cast=float):
    m = re.search(pattern, text, re.S)
    if not m:
        raise SystemExit(f"pattern not found: {pattern}")
    vals = [cast(m.group(i + 1)) for i in range(n)]
    return vals[0] if n == 1 else vals


def bench() -> dict:
    out = {"cases": []}
    for N, steps, thermal, les in BENCH_CASES:
        txt = run("bench", str(N), str(steps), str(thermal), str(les))
        grid = grab(r"(\d+)x(\d+)x(\d+) =", txt, 3, int)
        rho, u = grab(r"checksum: rho " + NUM + r"\s+u " + NUM, txt, 2, str)
        out["cases"].append({"N": N, "warmup": 30, "steps": steps, "thermal": thermal, "les": les,
                             "grid": grid, "rho_sum": rho, "u_sum": u})
        print(f"bench {N} {steps} {thermal} {les}: {grid} rho {rho} u {u}", flush=True)
    txt = run("bench_fma", "32", "120", "0", "1")
    out["published_fma_build"] = {
        "N": 32, "warmup": 30, "steps": 120, "thermal": 0, "les": 1,
        "rho_sum": grab(r"checksum: rho " + NUM, txt, 1, str),
        "u_sum": grab(r"checksum: rho " + NUM + r"\s+u " + NUM, txt, 2, str)[1],
        "note": "-O3 -march=native: GCC contracts a*b+c into FMA; not the uncontracted arithmetic",
    }
    return out


def physref() -> dict:
    t = run("physref")
    d = {}
    d["poiseuille_u_center"], d["poiseuille_u_exact"], d["poiseuille_l2_pct"], d["poiseuille_max_pct"] = grab(
        r"u_центр " + NUM + r" \(точно " + NUM + r"\), L2 " + NUM + r" %, макс " + NUM, t, 4)
    d["poiseuille_flux"], d["poiseuille_flux_exact"], d["poiseuille_mass_drift"] = grab(
        r"расход " + NUM + r" \(точно " + NUM + r"\), дрейф массы " + NUM, t, 3)
    d["decay_lambda"], d["decay_lambda_exact"] = grab(r"измерено " + NUM + r", аналитика nu\*pi\^2/H\^2 = " + NUM, t, 2)
    d["l2_h8_pct"] = grab(r"H=8:  L2 " + NUM, t)
    d["l2_h16_pct"] = grab(r"H=16: L2 " + NUM, t)
    d["l2_h32_pct"] = grab(r"H=32: L2 " + NUM, t)
    d["order_p"] = grab(r"p = " + NUM, t)
    d["trt_l2_pct"], d["trt_u_center"], d["trt_mass_drift"] = grab(
        r"TRT: L2 " + NUM + r" %, u_центр " + NUM + r", дрейф массы " + NUM, t, 3)
    d["checks_failed"] = grab(r"проверок (\d+), ошибок (\d+)", t, 2, int)[1]
    return d


def thermal() -> dict:
    t = run("thermal")
    d = {}
    d["diffusion_A0"], d["diffusion_A1"], d["diffusion_alpha"], d["diffusion_alpha_set"] = grab(
        r"A0 = " + NUM + r", A1 = " + NUM + r" за 1500 шагов; alpha измеренная " + NUM + r", заданная " + NUM, t, 4)
    d["advection_u_phase"], d["advection_u_field"], d["advection_phase_residual"], d["advection_A0"], d["advection_A1"] = grab(
        r"скорость по фазе " + NUM + r", средняя по полю " + NUM + r" \(задано [\d.]+\); остаток фазы " + NUM +
        r" рад; амплитуда " + NUM + r" -> " + NUM, t, 5)
    d["conduction_max_dev"], d["conduction_nu"] = grab(r"от линейного профиля " + NUM + r", Nu = " + NUM, t, 2)
    for ra in (1000, 1700, 2500):
        a, b, nu = grab(r"Ra = \s*" + str(ra) + r": амплитуда " + NUM + r" \(6000 шагов\) -> " + NUM +
                        r" \(24000\), Nu = " + NUM, t, 3)
        d[f"rb{ra}_amp_6000"], d[f"rb{ra}_amp_24000"], d[f"rb{ra}_nu"] = a, b, nu
    d["buoyancy_hot"], d["buoyancy_cold"] = grab(r"горячая ячейка " + NUM + r", холодная " + NUM, t, 2)
    d["checks_failed"] = grab(r"проверок (\d+), ошибок (\d+)", t, 2, int)[1]
    return d


def validate() -> dict:
    t = run("validate")
    d = {}
    d["sphere_grid"] = grab(r"сетка (\d+)x(\d+)x(\d+) = \d+ тыс\. ячеек, tau=0\.5480", t, 3, int)
    d["sphere_cd"] = grab(r"Cd измеренный " + NUM, t)
    d["sphere_area"], d["sphere_cd_amp"] = grab(r"A=" + NUM + r" ячеек, Re_lat=[\d.]+, шагов 550, [\d.]+ с, размах Cd " + NUM, t, 2)
    d["coarse_cd"] = grab(r"D=13: Cd=" + NUM, t)
    d["cylinder_cd"], d["cylinder_cl_amp"] = grab(r"Cd=" + NUM + r", размах Cl=" + NUM, t, 2)
    d["cylinder_divergence_reset"] = "Divergence detected" in t
    d["checks_failed"] = grab(r"проверок (\d+), ошибок (\d+)", t, 2, int)[1]
    return d


def lbm_suite() -> dict:
    t = run("lbm")
    d = {}
    (d["les_avg_rho"], d["les_tke"], d["les_vmax"], d["les_pmin"], d["les_pmax"], d["les_neg"], d["les_bad"],
     d["les_convergence"]) = grab(r"\[турбулентность\] rho=" + NUM + r"\s+TKE=" + NUM + r" м2/с2\s+vmax=" + NUM +
                                  r" м/с\s+p=\[" + NUM + r", " + NUM + r"\] Па\s+отриц\.популяций=" + NUM +
                                  r"\s+замен=" + NUM + r"\s+сходимость=" + NUM, t, 8)
    d["lam_avg_rho"], d["lam_vmax"] = grab(r"\[без турбулент\.\] rho=" + NUM + r"\s+vmax=" + NUM, t, 2)
    d["outlet_off_avg_rho"], d["outlet_on_avg_rho"], d["outlet_off_vmax"], d["outlet_on_vmax"] = grab(
        r"конвективный выход меняет поле: rho " + NUM + r" против " + NUM + r", vmax " + NUM + r" против " + NUM, t, 4)
    d["ground_off_vmax"], d["ground_on_vmax"] = grab(r"земля меняет решение: vmax " + NUM + r" против " + NUM, t, 2)
    d["thread_rho_sum"], d["thread_u_sum"] = grab(r"не зависит от числа потоков: rho " + NUM + r" против [\d.]+ \(отклонение [\w.+-]+\), u " + NUM, t, 2)
    d["padded_grid"] = grab(r"домен: размер (\d+)x(\d+)x(\d+)", t, 3, int)
    d["padded_solid_cells"] = grab(r"маска тела сохранена \((\d+) твёрдых", t, 1, int)
    d["mu_273"] = grab(r"mu\(273\.15 K\) = " + NUM, t)
    d["mu_288"] = grab(r"mu\(288\.15 K\) = " + NUM, t)
    d["re_phys"] = grab(r"initLBM: Re=" + NUM, t)
    d["checks_failed"] = grab(r"проверок: (\d+), ошибок: (\d+)", t, 2, int)[1]
    return d


# Per-cell field dumps of oracle/_ref/ref_fields (the reference solver run on
# small cases): recorded as SHA-256 digests of the raw float32 bytes, plus a
# few values for a readable failure.
FIELD_CASES = ["bench12_les", "bench12_les_thermal", "bench16", "rb_walls", "rb_storm", "bench12_auto",
               "bench12_auto_clamped"]


def fields() -> dict:
    import hashlib
    import numpy as np
    out = {"cases": {}}
    for case in FIELD_CASES:
        d = ROOT / "oracle" / "_ref" / "fields" / case
        d.mkdir(parents=True, exist_ok=True)
        text = run("ref_fields", case, str(d))
        m = re.search(r"(\d+)x(\d+)x(\d+) steps (\d+) step (\d+)", text)
        rec = {"grid": [int(g) for g in m.groups()[:3]], "steps": int(m.group(4)), "calls": "one per step",
               "final_step": int(m.group(5)), "digests": {}, "samples": {}}
        a = re.search(r"auto (.*)", text)
        if a:   # tau from Re: what initLBM derived, and the air state it used
            kv = a.group(1).split()
            rec["init"] = {kv[i]: float(kv[i + 1]) for i in range(0, len(kv), 2)}
        for f in sorted(d.glob("*.f32")):
            raw = f.read_bytes()
            v = np.frombuffer(raw, dtype="<f4")
            rec["digests"][f.stem] = hashlib.sha256(raw).hexdigest()
            idx = [0, v.size // 3, v.size // 2, v.size - 1]
            rec["samples"][f.stem] = {str(i): float(v[i]) for i in idx}
        out["cases"][case] = rec
    out["atmosphere"] = [{"altitude": float(a), "air_temperature": float(t), "air_density": float(r)}
                         for a, t, r in re.findall(r"altitude (\S+) air_temperature (\S+) air_density (\S+)",
                                                   run("ref_fields", "atmosphere", "-"))]
    return out


def voxels() -> dict:
    """The reference's own vote on the synthetic ellipsoid STLs of the measured cases (the
    199,080-triangle C2 body on 256x128x104 and its x2 scale on 512x256x256), as SHA-256
    digests of the uint8 masks."""
    import hashlib
    sys.path.insert(0, str(ROOT / "aeros-engine_amd"))
    import workloads as W
    out = {"cases": {}}
    d = ROOT / "oracle" / "_ref" / "voxels"
    d.mkdir(parents=True, exist_ok=True)
    for name, scale in (("c2", 1.0), ("c4", 2.0)):
        meshes = {}
        cfg = W.c2_config(lambda t, dims, mn, cs: meshes.setdefault("m", (t, dims, mn, cs)) and None,
                          scale=scale, with_mask=True)
        tris, dims, mn, cs = meshes["m"]
        (d / f"{name}.f32").write_bytes(tris.astype("<f4").tobytes())
        text = run("ref_voxels", str(d / f"{name}.f32"), *[str(x) for x in dims], *[repr(float(x)) for x in mn],
                   *[repr(float(x)) for x in cs], str(d / f"{name}.u8"))
        ntri, solid = (int(x) for x in re.search(r"triangles (\d+) solid (\d+)", text).groups())
        raw = (d / f"{name}.u8").read_bytes()
        out["cases"][name] = {"grid": list(dims), "min": list(mn), "cell": list(cs), "triangles": ntri,
                              "solid": solid, "sha256": hashlib.sha256(raw).hexdigest()}
        del cfg
    return out


def box_run() -> dict:
    """The reference solver on the measured ellipsoid case itself: 256x128x104, TRT +
    Smagorinsky, U0 0.08, convective outlet, the C2 body's mask (the host voxelizer's, whose
    digest equals the reference's own vote above).  One call per step for 1000 calls: the
    drag coefficient, the step counter and max |u| every 50 calls, and SHA-256 digests of
    the fields after 300 calls."""
    import hashlib
    import numpy as np
    sys.path.insert(0, str(ROOT / "aeros-engine_amd"))
    import aeros_lbm as A
    import workloads as W
    cfg = W.c2_config(A.voxelize, re=W.REF_DIVERGING_RE)
    nz, ny, nx = cfg["voxels"].shape
    d = ROOT / "oracle" / "_ref" / "fields" / "c2_box"
    d.mkdir(parents=True, exist_ok=True)
    (d / "mask.u8").write_bytes(np.ascontiguousarray(cfg["voxels"], dtype=np.uint8).tobytes())
    calls, dump_at = 1000, 300
    text = run("ref_fields", "box_les", str(d), str(nx), str(ny), str(nz), repr(cfg["tau"]), str(calls), str(dump_at))
    traj = [{"call": int(c), "step": int(s), "cd": float(cd), "area": float(a), "umax": float(u)}
            for c, s, cd, a, u in re.findall(r"call (\d+) step (\d+) Cd (\S+) A (\S+) max\|u\| (\S+)", text)]
    rec = {"grid": [nx, ny, nz], "tau": cfg["tau"], "calls": calls, "dump_after_calls": dump_at,
           "cv_box": [0.45, 0.40, 0.40], "trajectory": traj, "digests": {}, "samples": {}}
    for f in sorted(d.glob("*.f32")):
        raw = f.read_bytes()
        v = np.frombuffer(raw, dtype="<f4")
        rec["digests"][f.stem] = hashlib.sha256(raw).hexdigest()
        idx = [0, v.size // 3, v.size // 2, v.size - 1]
        rec["samples"][f.stem] = {str(i): float(v[i]) for i in idx}
    return rec


def main() -> None:
    if not (REF / "bench").exists():
        raise SystemExit("oracle/_ref is not built: make -f oracle/ref.mk -j8")
    which = sys.argv[1:] or ["bench", "physref", "thermal", "validate", "lbm", "fields", "voxels", "box_run"]
    funcs = {"bench": bench, "physref": physref, "thermal": thermal, "validate": validate, "lbm": lbm_suite,
             "fields": fields, "voxels": voxels, "box_run": box_run}
    for name in which:
        data = funcs[name]()
        data["_source"] = f"oracle/_ref/{ {'fields': 'ref_fields', 'voxels': 'ref_voxels', 'box_run': 'ref_fields'}.get(name, name)} built by oracle/ref.mk"
        (OUT / f"ref_{name}.json").write_text(json.dumps(data, indent=1, ensure_ascii=False) + "\n")
        print(f"wrote ref_{name}.json", flush=True)


if __name__ == "__main__":
    main()
