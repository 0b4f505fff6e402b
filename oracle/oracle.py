"""ctypes binding of the CPU oracle (oracle/build/liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker.  The product path never
loads it.  The oracle takes the same ``aeros_lbm_config`` struct as the HIP
library (the struct is the ABI, declared once in aeros-engine_amd/aeros_lbm.py)
and exposes the same operations, so a parity test can run both on identical
inputs.
"""
from __future__ import annotations

import ctypes as C
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent / "aeros-engine_amd"))
import aeros_lbm as A  # noqa: E402  (ABI structs only)

LIB_PATH = HERE / "build" / "liblbm_oracle.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "ora_config_defaults": (None, [P(A.Config)]),
        "ora_create": (vp, [P(A.Config), P(C.c_int)]),
        "ora_destroy": (None, [vp]),
        "ora_dims": (C.c_int, [vp, P(C.c_int32), P(C.c_float), P(C.c_float)]),
        "ora_step": (C.c_int, [vp, C.c_int]),
        "ora_set_mask": (C.c_int, [vp, P(C.c_uint8)]),
        "ora_read_field": (C.c_int, [vp, C.c_int, P(C.c_double)]),
        "ora_read_populations": (C.c_int, [vp, P(C.c_double)]),
        "ora_scalars": (C.c_int, [vp, P(A.Scalars)]),
        "ora_cv_force": (C.c_int, [vp, C.c_float, C.c_float, C.c_float, P(C.c_double)]),
        "ora_projected_area": (C.c_float, [vp, C.c_int]),
        "ora_checksums": (C.c_int, [vp, P(C.c_double), P(C.c_double)]),
        "ora_thermal_upload": (C.c_int, [vp, P(C.c_float)]),
        "ora_thermal_reset": (C.c_int, [vp]),
        "ora_equilibrium_moments": (C.c_int, [C.c_float] * 4 + [P(C.c_float)] * 4),
        "ora_q_half": (C.c_float, [C.c_float]),
        "ora_storage_shift": (C.c_int, [vp, P(C.c_double)]),
        "ora_sutherland": (C.c_float, [C.c_float]),
        "ora_q_bf16": (C.c_float, [C.c_float]),
        "ora_voxelize": (C.c_int, [P(C.c_float), C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                   P(C.c_float), P(C.c_float), P(C.c_int32)]),
        "ora_threads": (C.c_int, [C.c_int]),
        "ora_last_divergence": (C.c_float, [vp, P(C.c_int)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def threads(n: int = 0) -> int:
    return load().ora_threads(n)


def default_config(**overrides) -> A.Config:
    c = A.Config()
    load().ora_config_defaults(C.byref(c))
    A.apply_overrides(c, overrides)
    return c


class OracleLattice:
    """The oracle with the same surface as aeros_lbm.Lattice."""

    def __init__(self, config: A.Config | None = None, **overrides):
        self.lib = load()
        cfg = config if config is not None else default_config()
        if overrides:
            A.apply_overrides(cfg, overrides)
        self.config = cfg
        err = C.c_int(0)
        h = self.lib.ora_create(C.byref(cfg), C.byref(err))
        if not h:
            raise RuntimeError(f"oracle create failed: {A.ERRORS.get(err.value, err.value)}")
        self.h = C.c_void_p(h)
        d = (C.c_int32 * 3)()
        mn = (C.c_float * 3)()
        cs = (C.c_float * 3)()
        self.lib.ora_dims(self.h, d, mn, cs)
        self.dims = (d[0], d[1], d[2])
        self.box_min = tuple(mn)
        self.cell_size = tuple(cs)

    @property
    def shape(self):
        nx, ny, nz = self.dims
        return (nz, ny, nx)

    @property
    def ncells(self):
        nx, ny, nz = self.dims
        return nx * ny * nz

    def close(self):
        if getattr(self, "h", None):
            self.lib.ora_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        self.lib.ora_step(self.h, int(n))

    def set_mask(self, solid):
        solid = np.ascontiguousarray(solid, dtype=np.uint8)
        assert solid.shape == self.shape
        rc = self.lib.ora_set_mask(self.h, solid.ctypes.data_as(C.POINTER(C.c_uint8)))
        if rc:
            raise RuntimeError(f"oracle set_mask failed {rc}")

    def field(self, name, dtype=np.float32):
        fid = A.FIELD_NAMES[name] if isinstance(name, str) else int(name)
        out = np.empty(self.ncells, dtype=np.float64)
        rc = self.lib.ora_read_field(self.h, fid, out.ctypes.data_as(C.POINTER(C.c_double)))
        if rc:
            raise RuntimeError(f"oracle field {name} unavailable")
        return out.astype(dtype).reshape(self.shape)

    def populations(self):
        out = np.empty(self.ncells * 19, dtype=np.float64)
        self.lib.ora_read_populations(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out.reshape(self.shape + (19,))

    def scalars(self):
        s = A.Scalars()
        self.lib.ora_scalars(self.h, C.byref(s))
        return s.as_dict()

    def storage_shift(self):
        """The 19 values stored populations are taken relative to (w_q, or the inflow
        equilibrium for the equilibrium-shifted 16-bit storages)."""
        out = np.empty(19, dtype=np.float64)
        self.lib.ora_storage_shift(self.h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def cv_force(self, hx=0.30, hy=0.30, hz=0.30):
        out = (C.c_double * 3)()
        self.lib.ora_cv_force(self.h, hx, hy, hz, out)
        return tuple(out)

    def projected_area(self, axis=0):
        return self.lib.ora_projected_area(self.h, axis)

    def checksums(self):
        a, b = C.c_double(), C.c_double()
        self.lib.ora_checksums(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def thermal_upload(self, T):
        T = np.ascontiguousarray(T, dtype=np.float32).ravel()
        rc = self.lib.ora_thermal_upload(self.h, T.ctypes.data_as(C.POINTER(C.c_float)))
        if rc:
            raise RuntimeError("oracle thermal upload failed")

    def thermal_reset(self):
        self.lib.ora_thermal_reset(self.h)

    def last_divergence(self):
        """(step, avg rho) of the last divergence reset, step -1 if none."""
        st = C.c_int(-1)
        r = self.lib.ora_last_divergence(self.h, C.byref(st))
        return st.value, r


def equilibrium_moments(rho, ux, uy, uz):
    lib = load()
    o = [C.c_float() for _ in range(4)]
    ok = lib.ora_equilibrium_moments(rho, ux, uy, uz, *[C.byref(v) for v in o])
    return bool(ok), tuple(v.value for v in o)


def sutherland(t_kelvin: float) -> float:
    """Dynamic viscosity of air (Pa s) at t_kelvin, airSutherlandViscosity."""
    return load().ora_sutherland(t_kelvin)


def q_half(x: float) -> float:
    return load().ora_q_half(x)


def q_bf16(x: float) -> float:
    return load().ora_q_bf16(x)


def voxelize(tris, dims, box_min, cell):
    lib = load()
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    nx, ny, nz = dims
    out = np.empty(nx * ny * nz, dtype=np.int32)
    mn = (C.c_float * 3)(*box_min)
    cs = (C.c_float * 3)(*cell)
    lib.ora_voxelize(tris.ctypes.data_as(C.POINTER(C.c_float)), tris.shape[0], nx, ny, nz, mn, cs,
                     out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out.reshape(nz, ny, nx)


# ---------------------------------------------------------------- post-processing
# Restatements of the reference's field post-processing, in float32 with the
# reference's operation order (numpy rounds every float32 operation), for the
# GPU post-processing tests.

f32 = np.float32


def vorticity_q(uxw, uyw, uzw, solid, cell):
    """computeLBMVorticityAndQ (aeros/src/lbm.cpp:2005-2093) on (nz, ny, nx) float32
    world velocities; solid = solid | ground.  Returns (vort, q, strain, interior)
    where vort/q are 0 on solid interior cells and strain is defined on the fluid
    interior cells whose x-derivative of u_x is finite (the reference leaves the
    other cells of these arrays untouched)."""
    cs = [f32(c) if c >= 1e-6 else f32(0.1) for c in cell]
    ic = [f32(1) / c for c in cs]
    u = [np.asarray(a, dtype=f32) for a in (uxw, uyw, uzw)]
    I = (slice(1, -1),) * 3

    def d(a, axis, icv):   # (a[+] - a[-]) * 0.5 * ic over the interior, axis in (z, y, x) order
        hi = [slice(1, -1)] * 3
        lo = [slice(1, -1)] * 3
        hi[axis] = slice(2, None)
        lo[axis] = slice(0, -2)
        return (a[tuple(hi)] - a[tuple(lo)]) * f32(0.5) * icv

    dux_dx, dux_dy, dux_dz = d(u[0], 2, ic[0]), d(u[0], 1, ic[1]), d(u[0], 0, ic[2])
    duy_dx, duy_dy, duy_dz = d(u[1], 2, ic[0]), d(u[1], 1, ic[1]), d(u[1], 0, ic[2])
    duz_dx, duz_dy, duz_dz = d(u[2], 2, ic[0]), d(u[2], 1, ic[1]), d(u[2], 0, ic[2])
    with np.errstate(all="ignore"):
        wx, wy, wz = duz_dy - duy_dz, dux_dz - duz_dx, duy_dx - dux_dy
        vm = np.sqrt(wx * wx + wy * wy + wz * wz)
        vm = np.where(np.isfinite(vm), vm, f32(0))
        vm = np.minimum(vm, f32(1000))
        sxy, sxz, syz = f32(0.5) * (dux_dy + duy_dx), f32(0.5) * (dux_dz + duz_dx), f32(0.5) * (duy_dz + duz_dy)
        oxy, oxz, oyz = f32(0.5) * (dux_dy - duy_dx), f32(0.5) * (dux_dz - duz_dx), f32(0.5) * (duy_dz - duz_dy)
        s2 = dux_dx * dux_dx + duy_dy * duy_dy + duz_dz * duz_dz + f32(2) * (sxy * sxy + sxz * sxz + syz * syz)
        o2 = f32(2) * (oxy * oxy + oxz * oxz + oyz * oyz)
        q = f32(0.5) * (o2 - s2)
        q = np.where(np.isfinite(q), q, f32(0))
        q = np.clip(q, f32(-1e6), f32(1e6))
        st = np.sqrt(f32(2) * s2)
        st = np.where(np.isfinite(st), st, f32(0))
    sol = np.asarray(solid, dtype=bool)[I]
    good = ~sol & np.isfinite(dux_dx)
    vort = np.zeros(u[0].shape, f32)
    qc = np.zeros(u[0].shape, f32)
    strain = np.full(u[0].shape, np.nan, f32)
    vort[I] = np.where(good, vm, f32(0))
    qc[I] = np.where(good, q, f32(0))
    strain[I] = np.where(good, st, np.nan)
    return vort, qc, strain


def sample_field(field, solid, box_min, cell, p, clamp_density=False):
    """sampleFieldTrilinear with solid cells skipped (aeros/src/lbm.cpp:1871-1925) at
    one world point; returns (value as the named sampler returns it, valid)."""
    nz, ny, nx = field.shape
    x, y, z = (f32(v) for v in p)
    default = f32(1) if clamp_density else f32(0)
    if not (np.isfinite(x) and np.isfinite(y) and np.isfinite(z)):
        return default, False
    cs = [f32(c) for c in cell]
    if min(cs) < f32(1e-9) or nx < 2 or ny < 2 or nz < 2:
        return default, False
    fx, fy, fz = (x - f32(box_min[0])) / cs[0], (y - f32(box_min[1])) / cs[1], (z - f32(box_min[2])) / cs[2]
    ix, iy, iz = int(np.floor(fx)), int(np.floor(fy)), int(np.floor(fz))
    if ix < 0 or iy < 0 or iz < 0 or ix > nx - 2 or iy > ny - 2 or iz > nz - 2:
        return default, False
    cl = lambda t: f32(min(max(t, f32(0)), f32(1)))  # noqa: E731
    tx, ty, tz = cl(fx - f32(ix)), cl(fy - f32(iy)), cl(fz - f32(iz))
    wx, wy, wz = (f32(1) - tx, tx), (f32(1) - ty, ty), (f32(1) - tz, tz)
    acc, wsum = f32(0), f32(0)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                if solid[iz + dz, iy + dy, ix + dx]:
                    continue
                v = f32(field[iz + dz, iy + dy, ix + dx])
                if not np.isfinite(v):
                    continue
                w = wx[dx] * wy[dy] * wz[dz]
                acc = f32(acc + v * w)
                wsum = f32(wsum + w)
    if wsum < f32(1e-6):
        return default, False
    acc = f32(acc / wsum)
    if not np.isfinite(acc):
        return default, False
    if clamp_density:
        acc = f32(min(max(acc, f32(0.8)), f32(1.2)))
    return acc, True


def velocity_lb(ux, uy, uz, solid, box_min, cell, p):
    """getLBMVelocityLB (aeros/src/lbm.cpp:1781-1843) at one world point."""
    nz, ny, nx = ux.shape
    x, y, z = (f32(v) for v in p)
    zero = np.zeros(3, f32)
    if not (np.isfinite(x) and np.isfinite(y) and np.isfinite(z)):
        return zero
    cs = [f32(c) for c in cell]
    fx, fy, fz = (x - f32(box_min[0])) / cs[0], (y - f32(box_min[1])) / cs[1], (z - f32(box_min[2])) / cs[2]
    ix, iy, iz = int(np.floor(fx)), int(np.floor(fy)), int(np.floor(fz))
    cl = lambda t: f32(min(max(t, f32(0)), f32(1)))  # noqa: E731
    tx, ty, tz = cl(fx - f32(ix)), cl(fy - f32(iy)), cl(fz - f32(iz))
    if ix < 0 or ix >= nx - 1 or iy < 0 or iy >= ny - 1 or iz < 0 or iz >= nz - 1:
        return zero
    one = f32(1)
    w = [(one - tx) * (one - ty) * (one - tz), tx * (one - ty) * (one - tz), (one - tx) * ty * (one - tz),
         tx * ty * (one - tz), (one - tx) * (one - ty) * tz, tx * (one - ty) * tz, (one - tx) * ty * tz, tx * ty * tz]
    samples, tot = [], f32(0)
    for k in range(8):
        cx, cy, cz = ix + (k & 1), iy + ((k >> 1) & 1), iz + ((k >> 2) & 1)
        if solid[cz, cy, cx]:
            continue
        v = np.array([ux[cz, cy, cx], uy[cz, cy, cx], uz[cz, cy, cx]], dtype=f32)
        if not np.isfinite(v[0]):
            continue
        samples.append((v, w[k]))
        tot = f32(tot + w[k])
    if not samples:
        return zero
    if tot < f32(1e-6):
        return samples[0][0]
    r = np.zeros(3, f32)
    for v, wk in samples:
        r = (r + v * f32(wk / tot)).astype(f32)
    return r if np.isfinite(r[0]) else zero


def velocity_world(ux, uy, uz, solid, box_min, cell, p, inlet, flow_speed, free_stream):
    """getLBMVelocityWorld (aeros/src/lbm.cpp:1845-1869) at one world point; inlet is
    the lattice inflow velocity, free_stream flowParams.v."""
    nz, ny, nx = ux.shape
    x, y, z = (f32(v) for v in p)
    mn = [f32(b) for b in box_min]
    cs = [f32(c) for c in cell]
    mx = [mn[0] + f32(nx) * cs[0], mn[1] + f32(ny) * cs[1], mn[2] + f32(nz) * cs[2]]
    fs = np.array(free_stream, dtype=f32)
    if x < mn[0] or x > mx[0] or y < mn[1] or y > mx[1] or z < mn[2] or z > mx[2]:
        return fs
    if np.isfinite(x) and np.isfinite(y) and np.isfinite(z) and min(cs) >= f32(1e-9):
        ix = int(np.floor((x - mn[0]) / cs[0]))
        iy = int(np.floor((y - mn[1]) / cs[1]))
        iz = int(np.floor((z - mn[2]) / cs[2]))
        if 0 <= ix < nx and 0 <= iy < ny and 0 <= iz < nz and solid[iz, iy, ix]:
            return np.zeros(3, f32)
    v = velocity_lb(ux, uy, uz, solid, box_min, cell, p)
    if v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < f32(1e-12):
        return np.zeros(3, f32)
    a, b, c = (f32(t) for t in inlet)
    u0 = np.sqrt(a * a + b * b + c * c)
    if not np.isfinite(u0) or u0 < f32(1e-6):
        u0 = f32(0.08)
    scale = f32(flow_speed) / u0
    if not np.isfinite(scale) or scale > f32(1000):
        scale = f32(20)
    w = (v * scale).astype(f32)
    if not np.isfinite(w[0]):
        return fs
    mag = np.sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2])
    vmax = f32(flow_speed) * f32(2)
    if mag > vmax:
        w = (w * (vmax / mag)).astype(f32)
    return w
