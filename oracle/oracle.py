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
