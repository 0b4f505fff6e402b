"""Measurements of the reference's headless suites, restated on field arrays.

Each function takes a lattice object (the HIP ``aeros_lbm.Lattice`` or the
CPU ``oracle.OracleLattice``; both expose step/field/scalars/cv_force) and
returns the number the reference's test prints, so the same measurement can
be compared against the golden vectors for either implementation.
"""
from __future__ import annotations

import math

import numpy as np


# ---------------------------------------------------------------- Poiseuille
def channel_profile(L, force, tau, ny):
    """runChannel's measurement (aeros/tests/physics_ref_headless.cpp:120-171)."""
    nx = 8
    nu = (tau - 0.5) / 3.0
    H = float(ny - 2)
    yc = 0.5 * (ny - 1)
    u_max_ref = force * H * H / (8.0 * nu)
    ux = L.field("ux").astype(np.float64)
    zc = L.dims[2] // 2
    err2 = ref2 = maxabs = flux = flux_ref = best = 0.0
    for y in range(1, ny - 1):
        u = float(np.sum(ux[zc, y, :], dtype=np.float64)) / nx
        s = y - yc
        u_ref = force / (2.0 * nu) * (0.25 * H * H - s * s)
        err2 += (u - u_ref) ** 2
        ref2 += u_ref ** 2
        maxabs = max(maxabs, abs(u - u_ref))
        flux += u
        flux_ref += u_ref
        if abs(s) <= 0.75:
            best = max(best, u)
    rho = L.field("rho").astype(np.float64)
    solid = L.field("solid") > 0.5
    ground = L.field("ground") > 0.5
    fl = ~(solid | ground)
    drift = abs(float(np.sum(rho[fl], dtype=np.float64)) / max(1, int(fl.sum())) - 1.0)
    return dict(l2=math.sqrt(err2 / max(1e-30, ref2)), max_pct=100.0 * maxabs / max(1e-30, u_max_ref),
                u_center=best, u_center_ref=u_max_ref, flux=flux, flux_ref=flux_ref, mass_drift=drift)


# ---------------------------------------------------------------- thermal
def harmonic_x(T, y, z):
    """First x harmonic amplitude and phase (aeros/tests/thermal_headless.cpp:101-128)."""
    nx = T.shape[2]
    k = 2.0 * math.pi / nx
    row = T[z, y, :].astype(np.float64)
    re = im = 0.0
    for x in range(nx):
        a = k * x
        re += row[x] * math.cos(a)
        im -= row[x] * math.sin(a)
    return 2.0 * math.sqrt(re * re + im * im) / nx, math.atan2(im, re)


def perturbation_rms(T, solid, hot=1.0, cold=0.0):
    """RMS deviation from the conduction profile (aeros/tests/thermal_headless.cpp:130-147)."""
    nz, ny, nx = T.shape
    span = float(ny - 2)
    acc = 0.0
    n = 0
    for y in range(1, ny - 1):
        s = (y - 0.5) / max(1.0, span)
        lin = hot + (cold - hot) * s
        m = ~solid[:, y, :]
        d = T[:, y, :].astype(np.float64)[m] - lin
        acc += float(np.sum(d * d))
        n += int(m.sum())
    return math.sqrt(acc / n) if n else 0.0


def cosine_T(shape):
    nz, ny, nx = shape
    k = 2.0 * math.pi / nx
    row = np.array([np.float32(0.5 + 0.5 * math.cos(k * x)) for x in range(nx)], dtype=np.float32)
    return np.broadcast_to(row[None, None, :], shape).copy()


# ---------------------------------------------------------------- drag
def cd_series(L, u0, warm, meas, box=0.30):
    """runCase's Cd / Cl over the measured steps (aeros/tests/validation_headless.cpp:137-160).

    One step per call, as runCase calls stepLBMCPU(1): a divergence reset ends
    the call it happens in, so the call pattern is part of the semantics."""
    for _ in range(warm):
        L.step(1)
    area = L.projected_area(0)
    q = np.float32(0.5) * np.float32(u0) * np.float32(u0) * np.float32(area)
    cds, cls = [], []
    for _ in range(meas):
        L.step(1)
        fx, fy, fz = L.cv_force(box, box, box)
        cds.append(float(np.float32(fx) / q))
        cls.append(float(np.float32(fy) / q))
    cd = np.array(cds, dtype=np.float32)
    cl = np.array(cls, dtype=np.float32)
    cd_mean = 0.0
    for v in cds:
        cd_mean += v
    return dict(area=area, cd=float(np.float32(cd_mean / len(cds))), cd_amp=float(cd.max() - cd.min()),
                cl_amp=float(cl.max() - cl.min()))


def host_checksums(L):
    """The bench's sequential double sums (aeros/tests/bench_lbm.cpp:103-107)."""
    rho = L.field("rho").ravel()
    ux, uy, uz = (L.field(k).ravel() for k in ("ux", "uy", "uz"))
    su3 = (ux + uy + uz).astype(np.float32)   # float sum per cell, as the reference does
    s_rho = 0.0
    s_u = 0.0
    # sequential double accumulation (math.fsum would not be the same operation)
    for a, b in zip(rho.tolist(), su3.tolist()):
        s_rho += a
        s_u += b
    return s_rho, s_u


# ------------------------------------------------- the measured ellipsoid case
def follow_box_run(L, ref, calls):
    """Steps L one call per step (the reference's frame loop) for `calls` calls and
    checks it against the reference solver's run of the same case (the box_run
    fixture, printed by oracle/ref_fields.cpp): every field after
    ref["dump_after_calls"] calls by SHA-256, and at each recorded call the step
    counter (calls broken by a reset do not advance it), the projected area, the
    drag coefficient Fx / (0.5 U0^2 A) and max |u| over the lattice, each computed
    in float as the reference's harness prints them (6 and 5 decimals)."""
    import hashlib
    f32 = np.float32
    area = L.projected_area(0)
    traj = {t["call"]: t for t in ref["trajectory"]}
    for call in range(1, calls + 1):
        L.step(1)
        if call == ref["dump_after_calls"]:
            for name, digest in ref["digests"].items():
                v = np.ascontiguousarray(L.field(name), dtype="<f4").ravel()
                for i, want in ref["samples"][name].items():
                    assert float(v[int(i)]) == want, (call, name, i, float(v[int(i)]), want)
                assert hashlib.sha256(v.tobytes()).hexdigest() == digest, (call, name)
        t = traj.get(call)
        if t is None:
            continue
        fx = L.cv_force(*ref["cv_box"])[0]
        cd = float(f32(fx) / (f32(0.5) * f32(0.08) * f32(0.08) * f32(area)))
        ux, uy, uz = (L.field(k) for k in ("ux", "uy", "uz"))
        umax = float(np.sqrt(ux * ux + uy * uy + uz * uz).max())
        step = int(L.scalars()["step"])
        print(f"call {call} step {step} Cd {cd:.6f} (reference {t['cd']:.6f}) max|u| {umax:.5f}", flush=True)
        assert step == t["step"], (call, step, t["step"])
        assert area == t["area"]
        assert abs(cd - t["cd"]) <= 1e-6 + 1e-6 * abs(t["cd"]), (call, cd, t["cd"])
        assert abs(umax - t["umax"]) <= 6e-6, (call, umax, t["umax"])
