"""Python binding of the MI355X LBM solver's C ABI (include/aeros_lbm.h).

The solver itself is C++ host code plus hand-written HIP kernels built into
``aeros-engine_amd/build/libaeros_lbm.so``; this module only declares the ABI
structs and calls the library through ctypes.  It mirrors the reference's
operator surface: ``initLBM`` -> :class:`Lattice` construction,
``stepLBMCPU`` -> :meth:`Lattice.step`, the ``lbmRho/lbmUx/...`` globals ->
:meth:`Lattice.field`, ``lbmControlVolumeForce`` -> :meth:`Lattice.cv_force`
(aeros/src/lbm.h:156-202 of the reference).

Importing this module needs no GPU; creating a :class:`Lattice` does, and
fails loudly if the HIP library is missing or no device is usable.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "libaeros_lbm.so"

FP32, FP64, FP16, BF16, FP32_FP64C, FP16_EQ, BF16_EQ = 0, 1, 2, 3, 4, 5, 6
# Domain faces (Config.free_slip): x = 0, x = Nx-1, y = 0, y = Ny-1, z = 0, z = Nz-1.
FACE_X0, FACE_X1, FACE_Y0, FACE_Y1, FACE_Z0, FACE_Z1 = 1, 2, 4, 8, 16, 32
PRECISIONS = {"fp32": FP32, "fp64": FP64, "fp16": FP16, "bf16": BF16, "fp32_fp64c": FP32_FP64C,
              "fp16_eq": FP16_EQ, "bf16_eq": BF16_EQ}

(FIELD_RHO, FIELD_UX, FIELD_UY, FIELD_UZ, FIELD_UX_WORLD, FIELD_UY_WORLD, FIELD_UZ_WORLD,
 FIELD_PRESSURE, FIELD_TKE, FIELD_STRAIN, FIELD_SOLID, FIELD_GROUND, FIELD_WALLDIST,
 FIELD_T, FIELD_BUOYANCY, FIELD_VORTICITY, FIELD_QCRIT) = range(17)
FIELD_NAMES = {
    "rho": FIELD_RHO, "ux": FIELD_UX, "uy": FIELD_UY, "uz": FIELD_UZ,
    "ux_world": FIELD_UX_WORLD, "uy_world": FIELD_UY_WORLD, "uz_world": FIELD_UZ_WORLD,
    "pressure": FIELD_PRESSURE, "tke": FIELD_TKE, "strain": FIELD_STRAIN, "solid": FIELD_SOLID,
    "ground": FIELD_GROUND, "walldist": FIELD_WALLDIST, "T": FIELD_T, "buoyancy": FIELD_BUOYANCY,
    "vorticity": FIELD_VORTICITY, "qcrit": FIELD_QCRIT,
}

ERRORS = {0: "ok", -1: "bad argument", -2: "out of memory", -3: "HIP error", -4: "bad state",
          -5: "domain too large", -6: "unsupported", -7: "no device"}


class Config(C.Structure):
    """aeros_lbm_config: everything initLBM reads (aeros/src/lbm.cpp:405-886)."""
    _fields_ = [
        ("struct_size", C.c_int32),
        ("voxels", C.POINTER(C.c_int32)),
        ("vox_n", C.c_int32 * 3),
        ("vox_min", C.c_float * 3),
        ("vox_max", C.c_float * 3),
        ("voxel_resolution", C.c_int32),
        ("flow_speed", C.c_float),
        ("flow_azimuth_deg", C.c_float),
        ("flow_elevation_deg", C.c_float),
        ("air_density", C.c_float),
        ("air_temperature", C.c_float),
        ("max_dim", C.c_float),
        ("frontal_area", C.c_float),
        ("ground_effect", C.c_int32),
        ("ground_effect_height", C.c_float),
        ("tau", C.c_float),
        ("U0", C.c_float),
        ("smagorinsky_c", C.c_float),
        ("convergence_threshold", C.c_float),
        ("domain_padding", C.c_float),
        ("inlet_turbulence", C.c_float),
        ("ground_height", C.c_float),
        ("body_force", C.c_float * 3),
        ("use_les", C.c_int32),
        ("use_trt", C.c_int32),
        ("convective_outlet", C.c_int32),
        ("periodic_x", C.c_int32),
        ("periodic_z", C.c_int32),
        ("auto_viscosity", C.c_int32),
        ("use_ground", C.c_int32),
        ("thermal", C.c_int32),
        ("rayleigh", C.c_float),
        ("prandtl", C.c_float),
        ("fixed_wall_temps", C.c_int32),
        ("hot_temp", C.c_float),
        ("cold_temp", C.c_float),
        ("perturbation", C.c_float),
        ("precision", C.c_int32),
        ("cell_cap", C.c_int64),
        ("device", C.c_int32),
        ("free_slip", C.c_int32),
    ]


class Scalars(C.Structure):
    """aeros_lbm_scalars: the reference's scalar globals (aeros/src/lbm.h:77-125)."""
    _fields_ = [
        ("step", C.c_int32), ("converged", C.c_int32), ("resets", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("avg_rho", C.c_float), ("avg_ke", C.c_float), ("max_u_lb", C.c_float),
        ("max_u_world", C.c_float), ("tke", C.c_float), ("convergence", C.c_float),
        ("neg_pops", C.c_int64), ("bad_pops", C.c_int64), ("rho_out_of_range", C.c_int64),
        ("rho_non_finite", C.c_int64), ("u_limited", C.c_int64), ("pressure_non_finite", C.c_int64),
        ("tau", C.c_float), ("viscosity", C.c_float), ("U0", C.c_float),
        ("reynolds_phys", C.c_float), ("reynolds_lb", C.c_float), ("blockage", C.c_float),
        ("pressure_min", C.c_float), ("pressure_max", C.c_float),
        ("residual_rho", C.c_float), ("residual_u", C.c_float), ("time_ms", C.c_float),
        ("fluid_cells", C.c_int64),
        ("nusselt", C.c_float), ("thermal_mean_t", C.c_float), ("thermal_alpha", C.c_float),
        ("thermal_gravity", C.c_float), ("thermal_tau", C.c_float), ("thermal_rayleigh", C.c_float),
        ("thermal_prandtl", C.c_float), ("thermal_ready", C.c_int32),
        ("thermal_ylo", C.c_int32), ("thermal_yhi", C.c_int32),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PhysClockState(C.Structure):
    """aeros_phys_clock: the fixed-tick physics clock (aeros/src/phys_clock.h:48-250)."""
    _fields_ = [
        ("fixed", C.c_int32), ("hz", C.c_float), ("max_substeps", C.c_int32),
        ("particle_substeps", C.c_int32), ("max_frame_dt", C.c_float),
        ("accumulator", C.c_float), ("lattice_accum", C.c_float),
        ("sim_time", C.c_double), ("real_time", C.c_double), ("dropped_time", C.c_double),
        ("total_ticks", C.c_int64), ("total_lattice_steps", C.c_int64),
        ("ticks", C.c_int32), ("lattice_steps", C.c_int32), ("part_steps", C.c_int32),
        ("tick_dt", C.c_float), ("achieved_hz", C.c_float), ("time_scale", C.c_float),
        ("wall_scale", C.c_float), ("display_ratio", C.c_float), ("saturated", C.c_int32),
    ]


class Frame(C.Structure):
    """aeros_frame: one finished asynchronous frame."""
    _fields_ = [("frame", C.c_int64), ("steps", C.c_int32), ("scalars", Scalars),
                ("field", C.POINTER(C.c_float) * 17)]


class BackendInfo(C.Structure):
    _fields_ = [
        ("compiled_with_hip", C.c_int32), ("device_count", C.c_int32), ("usable", C.c_int32),
        ("gfx_major", C.c_int32), ("gfx_minor", C.c_int32), ("gfx_stepping", C.c_int32),
        ("runtime_version", C.c_int32), ("driver_version", C.c_int32),
        ("total_memory_mb", C.c_int64), ("compute_units", C.c_int32), ("last_error", C.c_int32),
        ("device_name", C.c_char * 128), ("arch_name", C.c_char * 64), ("status", C.c_char * 512),
    ]


def default_config(**overrides) -> Config:
    """Reference defaults (LBMParams, ThermalParams, globals), then overrides.

    Mirrors aeros_lbm_config_defaults without needing the library loaded."""
    c = Config()
    c.struct_size = C.sizeof(Config)
    c.voxel_resolution = 48
    for i in range(3):
        c.vox_min[i] = -1.0
        c.vox_max[i] = 1.0
    c.flow_speed = 2.0
    c.air_density = 1.225
    c.air_temperature = 288.15
    c.max_dim = 1.0
    c.tau = 0.6
    c.U0 = 0.1
    c.smagorinsky_c = 0.12
    c.convergence_threshold = 1e-5
    c.domain_padding = 0.25
    c.inlet_turbulence = 0.02
    c.use_les = 1
    c.convective_outlet = 1
    c.auto_viscosity = 1
    c.rayleigh = 1.0e5
    c.prandtl = 0.71
    c.fixed_wall_temps = 1
    c.hot_temp = 1.0
    c.cold_temp = 0.0
    c.perturbation = 0.02
    c.precision = FP32
    apply_overrides(c, overrides)
    return c


def apply_overrides(c: Config, overrides: dict) -> Config:
    for k, v in overrides.items():
        if k == "voxels":
            set_voxels(c, v)
        elif k in ("vox_n", "vox_min", "vox_max", "body_force"):
            arr = getattr(c, k)
            for i in range(3):
                arr[i] = v[i]
        elif k == "precision" and isinstance(v, str):
            c.precision = PRECISIONS[v]
        else:
            if not hasattr(c, k):
                raise AttributeError(f"unknown config field {k}")
            setattr(c, k, v)
    return c


def set_voxels(c: Config, vox) -> None:
    """Attach a voxel grid (z, y, x) int32 array; the array is kept alive on the struct."""
    if vox is None:
        c.voxels = None
        c._vox_keep = None
        return
    vox = np.ascontiguousarray(vox, dtype=np.int32)
    nz, ny, nx = vox.shape
    c.vox_n[0], c.vox_n[1], c.vox_n[2] = nx, ny, nz
    c.voxels = vox.ctypes.data_as(C.POINTER(C.c_int32))
    c._vox_keep = vox


_lib = None


class ServeConfig(C.Structure):
    """aeros_serve_config (include/aeros_serve.h)."""
    _fields_ = [("struct_size", C.c_int32), ("precision", C.c_int32), ("memory_budget", C.c_int64),
                ("max_cases", C.c_int32), ("round_steps", C.c_int32), ("cd_tolerance", C.c_double),
                ("max_steps", C.c_int32), ("cv_half", C.c_float), ("device", C.c_int32),
                ("tick_steps", C.c_int32), ("cd_average", C.c_int32)]


class ServeResult(C.Structure):
    _fields_ = [("state", C.c_int32), ("steps", C.c_int32), ("rounds", C.c_int32), ("resets", C.c_int32),
                ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("error", C.c_int32),
                ("cd", C.c_double), ("cd_prev", C.c_double), ("area", C.c_float), ("nusselt", C.c_float),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double), ("bytes", C.c_int64),
                ("tick_admit", C.c_int64), ("tick_done", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64), ("rounds", C.c_int64),
                ("launches", C.c_int64), ("ticks", C.c_int64), ("resident", C.c_int32), ("queued", C.c_int32),
                ("peak_resident", C.c_int32), ("bytes_in_use", C.c_int64), ("peak_bytes", C.c_int64),
                ("busy_seconds", C.c_double), ("cell_steps", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


REQ_STATES = {0: "queued", 1: "running", 2: "done", 3: "limit", 4: "failed"}


class StlInfo(C.Structure):
    """aeros_stl_info: stlformat::Detection (aeros/src/stl_format.h:36-42)."""
    _fields_ = [("format", C.c_int32), ("reason", C.c_int32), ("declared_triangles", C.c_uint32),
                ("pad", C.c_int32), ("payload_bytes", C.c_int64), ("trailing_bytes", C.c_int64)]


STL_FORMATS = {0: "unknown", 1: "binary", 2: "ascii"}
STL_REASONS = {1: "empty", 2: "solid+keywords", 3: "binary exact", 4: "binary with tail", 5: "text without solid",
               6: "text without keywords", 7: "truncated", 8: "unrecognised"}


def load_library(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load libaeros_lbm.so and declare its prototypes.  Raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"HIP library {p} is not built; run __graft_entry__.build()")
    lib = C.CDLL(str(p))
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "aeros_lbm_config_defaults": (None, [P(Config)]),
        "aeros_lbm_create": (vp, [P(Config), P(C.c_int)]),
        "aeros_lbm_destroy": (C.c_int, [vp]),
        "aeros_lbm_reset": (C.c_int, [vp]),
        "aeros_lbm_dims": (C.c_int, [vp, P(C.c_int32), P(C.c_float), P(C.c_float)]),
        "aeros_lbm_set_mask": (C.c_int, [vp, P(C.c_uint8), C.c_int32, C.c_int32, C.c_int32]),
        "aeros_lbm_step": (C.c_int, [vp, C.c_int32]),
        "aeros_lbm_update": (C.c_int, [vp, C.c_float, C.c_int32, C.c_int32]),
        "aeros_lbm_update_steps": (C.c_int32, [C.c_float, C.c_int32, C.c_int32]),
        "aeros_lbm_lattice_dt": (C.c_int, [vp, P(C.c_float)]),
        "aeros_clock_defaults": (None, [P(PhysClockState)]),
        "aeros_clock_reset": (None, [P(PhysClockState)]),
        "aeros_clock_sanitize": (None, [P(PhysClockState)]),
        "aeros_clock_advance": (C.c_int32, [P(PhysClockState), C.c_float, C.c_float, C.c_float, P(C.c_int32)]),
        "aeros_clock_particle_dt": (C.c_float, [P(PhysClockState)]),
        "aeros_clock_frame_sim_dt": (C.c_float, [P(PhysClockState)]),
        "aeros_lbm_frames_start": (C.c_int, [vp, C.c_uint32]),
        "aeros_lbm_frame_submit": (C.c_int, [vp, C.c_float, C.c_int32, C.c_int32]),
        "aeros_lbm_frame_latest": (C.c_int, [vp, P(Frame)]),
        "aeros_lbm_frames_wait": (C.c_int, [vp]),
        "aeros_lbm_frames_stop": (C.c_int, [vp]),
        "aeros_lbm_read_field": (C.c_int, [vp, C.c_int32, P(C.c_float)]),
        "aeros_lbm_read_field_f64": (C.c_int, [vp, C.c_int32, P(C.c_double)]),
        "aeros_lbm_read": (C.c_int, [vp, C.c_uint32, P(C.c_float), P(C.c_float), P(C.c_float),
                                     P(C.c_float), P(C.c_float)]),
        "aeros_lbm_get_scalars": (C.c_int, [vp, P(Scalars)]),
        "aeros_lbm_read_populations": (C.c_int, [vp, P(C.c_double)]),
        "aeros_lbm_cv_force": (C.c_int, [vp, C.c_float, C.c_float, C.c_float, P(C.c_double)]),
        "aeros_lbm_projected_area": (C.c_int, [vp, C.c_int32, P(C.c_float)]),
        "aeros_lbm_rho_sum": (C.c_int, [vp, P(C.c_double)]),
        "aeros_lbm_checksums": (C.c_int, [vp, P(C.c_double), P(C.c_double)]),
        "aeros_lbm_thermal_upload": (C.c_int, [vp, P(C.c_float)]),
        "aeros_lbm_thermal_reset": (C.c_int, [vp]),
        "aeros_lbm_equilibrium_moments": (C.c_int, [C.c_float] * 4 + [P(C.c_float)] * 4),
        "aeros_voxelize": (C.c_int, [P(C.c_float), C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                     P(C.c_float), P(C.c_float), P(C.c_int32)]),
        "aeros_voxelize_gpu": (C.c_int, [P(C.c_float), C.c_int64, C.c_int32, C.c_int32, C.c_int32,
                                         P(C.c_float), P(C.c_float), C.c_int32, P(C.c_int32)]),
        "aeros_stl_detect": (C.c_int, [C.c_char_p, C.c_int64, P(StlInfo)]),
        "aeros_lbm_vorticity_q": (C.c_int, [vp]),
        "aeros_lbm_sample": (C.c_int, [vp, C.c_int32, P(C.c_float), C.c_int64, P(C.c_float), P(C.c_uint8)]),
        "aeros_lbm_sample_noalloc": (C.c_int, [vp, C.c_int32, P(C.c_float), C.c_int64, P(C.c_float), P(C.c_uint8)]),
        "aeros_lbm_sample_velocity": (C.c_int, [vp, C.c_int32, P(C.c_float), C.c_int64, P(C.c_float)]),
        "aeros_stl_parse": (C.c_int, [C.c_char_p, C.c_int64, P(P(C.c_float)), P(P(C.c_float)), P(C.c_int64),
                                      P(C.c_int64)]),
        "aeros_stl_load": (C.c_int, [C.c_char_p, P(P(C.c_float)), P(P(C.c_float)), P(C.c_int64), P(C.c_int64)]),
        "aeros_stl_free": (None, [P(C.c_float)]),
        "aeros_stl_write_binary": (C.c_int, [C.c_char_p, P(C.c_float), C.c_int64, C.c_char_p]),
        "aeros_voxel_plan": (C.c_int, [P(C.c_float), P(C.c_float), C.c_float, C.c_int32,
                                       P(C.c_int32), P(C.c_float), P(C.c_float)]),
        "aeros_lbm_last_error": (C.c_int, [C.c_char_p, C.c_int32]),
        "aeros_probe": (C.c_int, [P(BackendInfo)]),
        "aeros_release_all": (None, []),
        "aeros_lbm_kernel_timing": (C.c_int, [vp, C.c_int32, P(C.c_double), P(C.c_int64)]),
        "setVoxelData": (None, [P(C.c_int), P(C.c_float), C.c_int, C.c_int, C.c_int] + [C.c_float] * 6),
        "aeros_serve_config_defaults": (None, [P(ServeConfig)]),
        "aeros_serve_create": (vp, [P(ServeConfig), P(C.c_int)]),
        "aeros_serve_submit": (C.c_int, [vp, P(Config), P(C.c_int64)]),
        "aeros_serve_poll": (C.c_int, [vp, C.c_int64, P(ServeResult)]),
        "aeros_serve_wait_all": (C.c_int, [vp, C.c_double]),
        "aeros_serve_clock": (C.c_double, [vp]),
        "aeros_serve_stats_get": (C.c_int, [vp, P(ServeStats)]),
        "aeros_serve_destroy": (C.c_int, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib if path is None else _lib
    return lib


def exported_symbols() -> list[str]:
    """Names of every entry point include/aeros_lbm.h declares."""
    hdr = (HERE.parent / "include" / "aeros_lbm.h").read_text()
    import re
    names = re.findall(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b(aeros_\w+|setVoxelData)\s*\(", hdr, re.M)
    return sorted(set(names))


def last_error() -> str:
    lib = load_library()
    buf = C.create_string_buffer(1024)
    code = lib.aeros_lbm_last_error(buf, 1024)
    return f"[{code}] {buf.value.decode(errors='replace')}" if code else ""


class AerosError(RuntimeError):
    pass


class PhysClock:
    """The fixed-tick physics clock of the host stepping driver (host code)."""

    def __init__(self, **settings):
        self.lib = load_library()
        self.c = PhysClockState()
        self.lib.aeros_clock_defaults(C.byref(self.c))
        for k, v in settings.items():
            setattr(self.c, k, v)

    def __getattr__(self, k):
        if k in ("lib", "c"):
            raise AttributeError(k)
        return getattr(self.c, k)

    def set(self, **settings):
        for k, v in settings.items():
            setattr(self.c, k, v)

    def advance(self, frame_dt: float, lattice_dt: float = 0.0, fps: float = 60.0):
        """Returns (ticks, lattice steps) for one rendered frame."""
        lbm = C.c_int32(0)
        t = self.lib.aeros_clock_advance(C.byref(self.c), float(frame_dt), float(lattice_dt), float(fps), C.byref(lbm))
        return t, lbm.value

    def sanitize(self):
        self.lib.aeros_clock_sanitize(C.byref(self.c))

    def reset(self):
        self.lib.aeros_clock_reset(C.byref(self.c))

    def particle_dt(self) -> float:
        return self.lib.aeros_clock_particle_dt(C.byref(self.c))

    def frame_sim_dt(self) -> float:
        return self.lib.aeros_clock_frame_sim_dt(C.byref(self.c))


def update_steps(dt: float, force_steps: int = -1, steps_per_frame: int = 3) -> int:
    """updateLBM's step count for these arguments (host rule)."""
    return load_library().aeros_lbm_update_steps(float(dt), int(force_steps), int(steps_per_frame))


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise AerosError(f"{what} failed: {ERRORS.get(rc, rc)} {last_error()}")


def probe() -> dict:
    lib = load_library()
    info = BackendInfo()
    lib.aeros_probe(C.byref(info))
    d = {k: getattr(info, k) for k, _ in info._fields_}
    for k in ("device_name", "arch_name", "status"):
        d[k] = d[k].decode(errors="replace")
    return d


class Lattice:
    """One solver instance on the GPU (the reference's initLBM .. shutdownLBM)."""

    def __init__(self, config: Config | None = None, **overrides):
        self.lib = load_library()
        cfg = config if config is not None else default_config()
        if overrides:
            apply_overrides(cfg, overrides)
        self.config = cfg
        err = C.c_int(0)
        h = self.lib.aeros_lbm_create(C.byref(cfg), C.byref(err))
        if not h:
            raise AerosError(f"aeros_lbm_create failed: {ERRORS.get(err.value, err.value)} {last_error()}")
        self.h = C.c_void_p(h)
        d = (C.c_int32 * 3)()
        mn = (C.c_float * 3)()
        cs = (C.c_float * 3)()
        _check(self.lib.aeros_lbm_dims(self.h, d, mn, cs), "aeros_lbm_dims")
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
            self.lib.aeros_lbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def step(self, n: int = 1):
        _check(self.lib.aeros_lbm_step(self.h, int(n)), "aeros_lbm_step")

    def update(self, dt: float, force_steps: int = -1, steps_per_frame: int = 3):
        _check(self.lib.aeros_lbm_update(self.h, float(dt), int(force_steps), int(steps_per_frame)),
               "aeros_lbm_update")

    def lattice_dt(self) -> float:
        v = C.c_float()
        _check(self.lib.aeros_lbm_lattice_dt(self.h, C.byref(v)), "aeros_lbm_lattice_dt")
        return v.value

    # asynchronous frames (host stepping driver)
    def frames_start(self, fields=("rho", "ux", "uy", "uz")):
        mask = 0
        for f in fields:
            mask |= 1 << (FIELD_NAMES[f] if isinstance(f, str) else int(f))
        self._frame_fields = [FIELD_NAMES[f] if isinstance(f, str) else int(f) for f in fields]
        _check(self.lib.aeros_lbm_frames_start(self.h, mask), "aeros_lbm_frames_start")

    def frame_submit(self, dt: float, force_steps: int = -1, steps_per_frame: int = 3):
        _check(self.lib.aeros_lbm_frame_submit(self.h, float(dt), int(force_steps), int(steps_per_frame)),
               "aeros_lbm_frame_submit")

    def frame_latest(self, copy: bool = True) -> dict:
        """Newest finished frame: {'frame', 'steps', 'scalars', 'fields': {name: array}}.
        copy=False returns views of the pinned slot (valid until the next call)."""
        fr = Frame()
        _check(self.lib.aeros_lbm_frame_latest(self.h, C.byref(fr)), "aeros_lbm_frame_latest")
        names = {v: k for k, v in FIELD_NAMES.items()}
        out = {"frame": fr.frame, "steps": fr.steps, "scalars": fr.scalars.as_dict(), "fields": {}}
        if fr.frame:
            for fid in self._frame_fields:
                a = np.ctypeslib.as_array(fr.field[fid], shape=(self.ncells,)).reshape(self.shape)
                out["fields"][names[fid]] = a.copy() if copy else a
        return out

    def frames_wait(self):
        _check(self.lib.aeros_lbm_frames_wait(self.h), "aeros_lbm_frames_wait")

    def frames_stop(self):
        _check(self.lib.aeros_lbm_frames_stop(self.h), "aeros_lbm_frames_stop")

    def reset(self):
        _check(self.lib.aeros_lbm_reset(self.h), "aeros_lbm_reset")

    def set_mask(self, solid):
        solid = np.ascontiguousarray(solid, dtype=np.uint8)
        nz, ny, nx = solid.shape
        _check(self.lib.aeros_lbm_set_mask(self.h, solid.ctypes.data_as(C.POINTER(C.c_uint8)), nx, ny, nz),
               "aeros_lbm_set_mask")

    def field(self, name, dtype=np.float32):
        fid = FIELD_NAMES[name] if isinstance(name, str) else int(name)
        if dtype == np.float64:
            out = np.empty(self.ncells, dtype=np.float64)
            _check(self.lib.aeros_lbm_read_field_f64(self.h, fid, out.ctypes.data_as(C.POINTER(C.c_double))),
                   "aeros_lbm_read_field_f64")
        else:
            out = np.empty(self.ncells, dtype=np.float32)
            _check(self.lib.aeros_lbm_read_field(self.h, fid, out.ctypes.data_as(C.POINTER(C.c_float))),
                   "aeros_lbm_read_field")
        return out.reshape(self.shape)

    def populations(self):
        out = np.empty(self.ncells * 19, dtype=np.float64)
        _check(self.lib.aeros_lbm_read_populations(self.h, out.ctypes.data_as(C.POINTER(C.c_double))),
               "aeros_lbm_read_populations")
        return out.reshape(self.shape + (19,))

    def scalars(self) -> dict:
        s = Scalars()
        _check(self.lib.aeros_lbm_get_scalars(self.h, C.byref(s)), "aeros_lbm_get_scalars")
        return s.as_dict()

    def cv_force(self, hx=0.30, hy=0.30, hz=0.30):
        out = (C.c_double * 3)()
        _check(self.lib.aeros_lbm_cv_force(self.h, hx, hy, hz, out), "aeros_lbm_cv_force")
        return tuple(out)

    def projected_area(self, axis=0):
        v = C.c_float()
        _check(self.lib.aeros_lbm_projected_area(self.h, axis, C.byref(v)), "aeros_lbm_projected_area")
        return v.value

    def checksums(self):
        a, b = C.c_double(), C.c_double()
        _check(self.lib.aeros_lbm_checksums(self.h, C.byref(a), C.byref(b)), "aeros_lbm_checksums")
        return a.value, b.value

    def rho_sum(self) -> float:
        a = C.c_double()
        _check(self.lib.aeros_lbm_rho_sum(self.h, C.byref(a)), "aeros_lbm_rho_sum")
        return a.value

    def thermal_upload(self, T):
        T = np.ascontiguousarray(T, dtype=np.float32).ravel()
        if T.size != self.ncells:
            raise ValueError("temperature field size mismatch")
        _check(self.lib.aeros_lbm_thermal_upload(self.h, T.ctypes.data_as(C.POINTER(C.c_float))),
               "aeros_lbm_thermal_upload")

    def thermal_reset(self):
        _check(self.lib.aeros_lbm_thermal_reset(self.h), "aeros_lbm_thermal_reset")

    def vorticity_q(self):
        """computeLBMVorticityAndQ on the device (fields "vorticity", "qcrit", "strain")."""
        _check(self.lib.aeros_lbm_vorticity_q(self.h), "aeros_lbm_vorticity_q")

    def sample(self, name: str, pos, noalloc: bool = False):
        """World-space trilinear sampler of a scalar field at (N, 3) points: (values, valid).
        noalloc: through aeros_lbm_sample_noalloc (fixed device staging, at most 65536 points)."""
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        out = np.empty(pos.shape[0], dtype=np.float32)
        valid = np.empty(pos.shape[0], dtype=np.uint8)
        F = C.POINTER(C.c_float)
        fn = "aeros_lbm_sample_noalloc" if noalloc else "aeros_lbm_sample"
        _check(getattr(self.lib, fn)(self.h, FIELD_NAMES[name], pos.ctypes.data_as(F), pos.shape[0],
                                     out.ctypes.data_as(F), valid.ctypes.data_as(C.POINTER(C.c_uint8))), fn)
        return out, valid.astype(bool)

    def sample_velocity(self, pos, world: bool = True):
        """getLBMVelocityWorld (world) or getLBMVelocityLB at (N, 3) points: (N, 3)."""
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        out = np.empty((pos.shape[0], 3), dtype=np.float32)
        F = C.POINTER(C.c_float)
        _check(self.lib.aeros_lbm_sample_velocity(self.h, 1 if world else 0, pos.ctypes.data_as(F), pos.shape[0],
                                                  out.ctypes.data_as(F)), "aeros_lbm_sample_velocity")
        return out

    def kernel_timing(self, enable: int):
        """enable 1/0/-1 (start/stop/reset); returns (total ms, launches) of the step kernel."""
        ms, n = C.c_double(), C.c_int64()
        _check(self.lib.aeros_lbm_kernel_timing(self.h, int(enable), C.byref(ms), C.byref(n)),
               "aeros_lbm_kernel_timing")
        return ms.value, n.value


def equilibrium_moments(rho, ux, uy, uz):
    lib = load_library()
    o = [C.c_float() for _ in range(4)]
    ok = lib.aeros_lbm_equilibrium_moments(rho, ux, uy, uz, *[C.byref(v) for v in o])
    return bool(ok), tuple(v.value for v in o)


def voxelize(tris, dims, box_min, cell):
    """BVH three-ray vote at cell centres; tris is (T, 9) float32. Returns (nz, ny, nx) int32."""
    lib = load_library()
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    nx, ny, nz = dims
    out = np.empty(nx * ny * nz, dtype=np.int32)
    mn = (C.c_float * 3)(*box_min)
    cs = (C.c_float * 3)(*cell)
    _check(lib.aeros_voxelize(tris.ctypes.data_as(C.POINTER(C.c_float)), tris.shape[0], nx, ny, nz, mn, cs,
                              out.ctypes.data_as(C.POINTER(C.c_int32))), "aeros_voxelize")
    return out.reshape(nz, ny, nx)


def voxelize_gpu(tris, dims, box_min, cell, device: int = 0):
    """The same vote on the device (bit-identical to voxelize); needs a gfx950 GPU."""
    lib = load_library()
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    nx, ny, nz = dims
    out = np.empty(nx * ny * nz, dtype=np.int32)
    mn = (C.c_float * 3)(*box_min)
    cs = (C.c_float * 3)(*cell)
    _check(lib.aeros_voxelize_gpu(tris.ctypes.data_as(C.POINTER(C.c_float)), tris.shape[0], nx, ny, nz, mn, cs,
                                  int(device), out.ctypes.data_as(C.POINTER(C.c_int32))), "aeros_voxelize_gpu")
    return out.reshape(nz, ny, nx)


def stl_detect(data: bytes) -> dict:
    """Format of an STL buffer (stlformat::detect)."""
    lib = load_library()
    info = StlInfo()
    _check(lib.aeros_stl_detect(data, len(data), C.byref(info)), "aeros_stl_detect")
    return dict(format=STL_FORMATS[info.format], reason=STL_REASONS.get(info.reason, str(info.reason)),
                declared_triangles=info.declared_triangles, payload_bytes=info.payload_bytes,
                trailing_bytes=info.trailing_bytes)


def _stl_take(lib, rc, what, pt, pn, nt, dropped):
    _check(rc, what)
    try:
        n = nt.value
        tris = np.ctypeslib.as_array(pt, shape=(n * 9,)).reshape(n, 9).copy()
        norms = np.ctypeslib.as_array(pn, shape=(n * 9,)).reshape(n, 9).copy()
    finally:
        lib.aeros_stl_free(pt)
        lib.aeros_stl_free(pn)
    return tris, norms, dropped.value


def stl_parse(data: bytes):
    """loadSTL on a buffer: ((T, 9) vertices, (T, 9) per-vertex normals, dropped count)."""
    lib = load_library()
    pt, pn = C.POINTER(C.c_float)(), C.POINTER(C.c_float)()
    nt, dropped = C.c_int64(), C.c_int64()
    rc = lib.aeros_stl_parse(data, len(data), C.byref(pt), C.byref(pn), C.byref(nt), C.byref(dropped))
    return _stl_take(lib, rc, "aeros_stl_parse", pt, pn, nt, dropped)


def stl_load(path):
    """loadSTL from a file."""
    lib = load_library()
    pt, pn = C.POINTER(C.c_float)(), C.POINTER(C.c_float)()
    nt, dropped = C.c_int64(), C.c_int64()
    rc = lib.aeros_stl_load(os.fsencode(path), C.byref(pt), C.byref(pn), C.byref(nt), C.byref(dropped))
    return _stl_take(lib, rc, "aeros_stl_load", pt, pn, nt, dropped)


def stl_write_binary(path, tris, header: str = ""):
    lib = load_library()
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    _check(lib.aeros_stl_write_binary(os.fsencode(path), tris.ctypes.data_as(C.POINTER(C.c_float)), tris.shape[0],
                                      header.encode()), "aeros_stl_write_binary")


class Server:
    """Served mode (include/aeros_serve.h): a server thread advances many
    resident cases in batched launches; submit() queues a case, poll() and
    wait_all() report results."""

    def __init__(self, **overrides):
        self.lib = load_library()
        cfg = ServeConfig()
        self.lib.aeros_serve_config_defaults(C.byref(cfg))
        for k, v in overrides.items():
            if k == "precision" and isinstance(v, str):
                v = PRECISIONS[v]
            setattr(cfg, k, v)
        self.config = cfg
        err = C.c_int(0)
        h = self.lib.aeros_serve_create(C.byref(cfg), C.byref(err))
        if not h:
            raise AerosError(f"aeros_serve_create failed: {ERRORS.get(err.value, err.value)} {last_error()}")
        self.h = C.c_void_p(h)
        self._keep = {}

    def submit(self, config: Config) -> int:
        rid = C.c_int64(0)
        _check(self.lib.aeros_serve_submit(self.h, C.byref(config), C.byref(rid)), "aeros_serve_submit")
        return rid.value

    def poll(self, rid: int) -> dict:
        r = ServeResult()
        _check(self.lib.aeros_serve_poll(self.h, int(rid), C.byref(r)), "aeros_serve_poll")
        d = r.as_dict()
        d["state_name"] = REQ_STATES.get(d["state"], str(d["state"]))
        return d

    def wait_all(self, timeout_s: float = 0.0) -> bool:
        return self.lib.aeros_serve_wait_all(self.h, float(timeout_s)) == 0

    def clock(self) -> float:
        return self.lib.aeros_serve_clock(self.h)

    def stats(self) -> dict:
        st = ServeStats()
        _check(self.lib.aeros_serve_stats_get(self.h, C.byref(st)), "aeros_serve_stats_get")
        return st.as_dict()

    def close(self):
        if getattr(self, "h", None):
            self.lib.aeros_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
