"""Synthetic cases: the reference's own test geometries and the measured workloads.

Each builder returns a dict of ``aeros_lbm_config`` overrides (voxel grid,
box, flow and solver parameters) that both the HIP solver and the CPU oracle
accept.  Geometry is evaluated in float32 exactly as the reference's C++ test
generators do, so the voxel masks are bit-identical to theirs.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def _centres(n: int, lo: float, cs: np.float32) -> np.ndarray:
    # lo + (i + 0.5f) * cs, all float32
    return (f32(lo) + (np.arange(n, dtype=f32) + f32(0.5)) * cs).astype(f32)


def sphere_voxels(n: int, lo: float, hi: float, radius: float) -> np.ndarray:
    """(n,n,n) voxel grid of a sphere of radius ``radius`` centred at 0 in the cube [lo,hi]^3.

    Same rule as aeros/tests/bench_lbm.cpp:45-60 and
    aeros/tests/validation_headless.cpp:57-81: solid iff |p| - R <= 0 at the
    voxel centre, evaluated in float."""
    cs = f32(f32(hi - lo) / f32(n))
    c = _centres(n, lo, cs)
    pz, py, px = np.meshgrid(c, c, c, indexing="ij")
    d = np.sqrt(px * px + py * py + pz * pz).astype(f32) - f32(radius)
    return (d <= f32(0)).astype(np.int32)


def cylinder_voxels(n: int, lo: float, hi: float, radius: float) -> np.ndarray:
    """Cylinder along x (aeros/tests/validation_headless.cpp:68-72)."""
    cs = f32(f32(hi - lo) / f32(n))
    c = _centres(n, lo, cs)
    pz, py, _px = np.meshgrid(c, c, c, indexing="ij")
    d = np.sqrt(py * py + pz * pz).astype(f32) - f32(radius)
    return (d <= f32(0)).astype(np.int32)


def bench_sphere(n: int = 32, les: bool = True, thermal: bool = False) -> dict:
    """The reference bench (aeros/tests/bench_lbm.cpp:35-79): sphere R=0.35 in [-1,1]^3 at n^3,
    padding 0.25, tau 0.6, U0 0.08, LES on, convective outlet, flow 20 m/s.
    n=32 gives the 80x72x72 domain of the published checksum."""
    vox = sphere_voxels(n, -1.0, 1.0, 0.35)
    return dict(
        voxels=vox, vox_min=(-1.0, -1.0, -1.0), vox_max=(1.0, 1.0, 1.0), voxel_resolution=n,
        max_dim=0.7, flow_speed=20.0, flow_azimuth_deg=0.0, flow_elevation_deg=0.0,
        U0=0.08, tau=0.6, use_les=int(les), use_ground=0, inlet_turbulence=0.0,
        domain_padding=0.25, auto_viscosity=0, convective_outlet=1, periodic_x=0,
        body_force=(0.0, 0.0, 0.0), thermal=int(thermal), rayleigh=2500.0, prandtl=1.0,
    )


def lbm_suite_sphere(les: bool = True) -> dict:
    """makeSphereCase(0.7) of aeros/tests/lbm_physics_headless.cpp:48-63 with the
    suite's parameters (tau 0.55, U0 0.08, Cs 0.14, padding 0.25)."""
    vox = sphere_voxels(24, -3.0, 3.0, 0.7)
    return dict(
        voxels=vox, vox_min=(-3.0, -3.0, -3.0), vox_max=(3.0, 3.0, 3.0), voxel_resolution=24,
        max_dim=1.4, flow_speed=20.0, tau=0.55, auto_viscosity=0, U0=0.08, smagorinsky_c=0.14,
        use_ground=0, domain_padding=0.25, use_les=int(les),
    )


def channel(ny: int = 34, tau: float = 0.65, force: float = 7.8e-6, trt: bool = False) -> dict:
    """Poiseuille channel 8 x ny x 8 (aeros/tests/physics_ref_headless.cpp:65-118):
    solid rows y=0 and y=ny-1, periodic x and z, Guo body force along x."""
    nx = nz = 8
    vox = np.zeros((nz, ny, nx), dtype=np.int32)
    vox[:, 0, :] = 1
    vox[:, ny - 1, :] = 1
    return dict(
        voxels=vox, vox_min=(-1.0, -1.0, -1.0), vox_max=(1.0, 1.0, 1.0), voxel_resolution=nx,
        max_dim=2.0, flow_speed=1.0, U0=0.0, tau=float(tau), use_les=0, use_ground=0,
        inlet_turbulence=0.0, domain_padding=0.0, auto_viscosity=0, convective_outlet=0,
        use_trt=int(trt), periodic_x=1, periodic_z=1, body_force=(float(force), 0.0, 0.0),
    )


def thermal_box(nx=32, ny=16, nz=8, walls_y=False, u0=0.0, tau=0.65, periodic_z=True,
                rayleigh=0.0, prandtl=1.0, fixed_wall_temps=False, perturbation=0.0,
                hot=1.0, cold=0.0) -> dict:
    """setupBox of aeros/tests/thermal_headless.cpp:54-99 plus the thermal parameters."""
    vox = np.zeros((nz, ny, nx), dtype=np.int32)
    if walls_y:
        vox[:, 0, :] = 1
        vox[:, ny - 1, :] = 1
    return dict(
        voxels=vox, vox_min=(-1.0, -1.0, -1.0), vox_max=(1.0, 1.0, 1.0), voxel_resolution=nx,
        max_dim=2.0, flow_speed=1.0, U0=float(u0), tau=float(tau), use_les=0, use_ground=0,
        inlet_turbulence=0.0, domain_padding=0.0, auto_viscosity=0, convective_outlet=0,
        periodic_x=1, periodic_z=int(periodic_z), body_force=(0.0, 0.0, 0.0),
        thermal=1, rayleigh=float(rayleigh), prandtl=float(prandtl),
        fixed_wall_temps=int(fixed_wall_temps), hot_temp=float(hot), cold_temp=float(cold),
        perturbation=float(perturbation),
    )


def validate_sphere(res=48, diameter_cells=20, re_lat=100.0, pad=0.15, u0=0.08) -> dict:
    """Sphere of aeros/tests/validation_headless.cpp:99-135 (BGK, no LES, convective outlet):
    res 48, D 20 gives the 92x84x84 domain of the published Cd 1.128."""
    radius = f32(3.0) * f32(diameter_cells) / f32(res)
    vox = sphere_voxels(res, -3.0, 3.0, float(radius))
    D = f32(diameter_cells)
    tau = f32(0.5) + f32(3.0) * f32(u0) * D / f32(re_lat)
    tau = max(f32(0.505), min(tau, f32(1.4)))
    return dict(
        voxels=vox, vox_min=(-3.0, -3.0, -3.0), vox_max=(3.0, 3.0, 3.0), voxel_resolution=res,
        max_dim=float(f32(2.0) * radius), flow_speed=20.0, U0=float(u0), tau=float(tau), use_les=0,
        use_ground=0, inlet_turbulence=0.0, domain_padding=float(pad), auto_viscosity=0,
        convective_outlet=1, periodic_x=0,
    )


# --------------------------------------------------------------- measured workloads

def sphere_in_box(nx, ny, nz, centre, diameter) -> np.ndarray:
    """Sphere by the cell-centre test |x - c| <= D/2 in lattice units (the bench rule)."""
    z, y, x = np.meshgrid(np.arange(nz, dtype=f32) + f32(0.5), np.arange(ny, dtype=f32) + f32(0.5),
                          np.arange(nx, dtype=f32) + f32(0.5), indexing="ij")
    cx, cy, cz = (f32(v) for v in centre)
    dx, dy, dz = x - cx, y - cy, z - cz
    d = np.sqrt(dx * dx + dy * dy + dz * dz).astype(f32)
    return (d <= f32(diameter) * f32(0.5)).astype(np.int32)


def lattice_box(nx, ny, nz, cell=1.0):
    """Voxel box whose cells coincide with the lattice (padding 0): world = lattice units."""
    return dict(vox_min=(0.0, 0.0, 0.0), vox_max=(nx * cell, ny * cell, nz * cell), domain_padding=0.0)


SIDE_WALLS = 4 | 8 | 16 | 32   # free-slip on y = 0, y = Ny-1, z = 0, z = Nz-1


def c1_sphere_fp64(precision="fp64") -> dict:
    """96x48x48, sphere D=12 at (24,24,24), U0 0.05 along +x, BGK, tau = 0.5 + 3 U0 D / Re (Re 100),
    uniform inflow, zero-gradient outflow (copy in place), free-slip side walls (y and z faces)."""
    nx, ny, nz = 96, 48, 48
    vox = sphere_in_box(nx, ny, nz, (24, 24, 24), 12)
    tau = float(f32(0.5) + f32(3.0) * f32(0.05) * f32(12.0) / f32(100.0))
    return dict(voxels=vox, **lattice_box(nx, ny, nz), voxel_resolution=nx, max_dim=12.0,
                flow_speed=20.0, U0=0.05, tau=tau, use_les=0, use_trt=0, auto_viscosity=0,
                inlet_turbulence=0.0, convective_outlet=0, free_slip=SIDE_WALLS, precision=precision)


def slip_box(nx, ny, nz, centre, diameter, free_slip, U0=0.06, tau=0.56, **over) -> dict:
    """Small sphere-in-a-box case for the free-slip walls: lattice-aligned box, BGK, no LES,
    uniform inflow along +x, zero-gradient outflow."""
    vox = sphere_in_box(nx, ny, nz, centre, diameter)
    cfg = dict(voxels=vox, **lattice_box(nx, ny, nz), voxel_resolution=nx, max_dim=float(diameter),
               flow_speed=10.0, U0=U0, tau=tau, use_les=0, use_trt=0, auto_viscosity=0,
               inlet_turbulence=0.0, convective_outlet=0, free_slip=free_slip)
    cfg.update(over)
    return cfg


def ellipsoid_mesh(semi=(30.0, 12.0, 12.0), centre=(0.0, 0.0, 0.0), n_u=316, n_v=316,
                   jitter=0.05, seed=12345) -> np.ndarray:
    """Closed UV-parametrised ellipsoid, (T, 9) float32 triangles, T = 2 n_u (n_v - 1).

    n_u = n_v = 316 gives 199,080 triangles.  Every vertex (poles included) gets
    a jitter drawn from U(-jitter, jitter) per axis by a seeded PCG64 generator;
    shared vertices move together so the surface stays closed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a, b, c = semi
    th = np.linspace(0.0, np.pi, n_v + 1)[1:-1]          # interior rings
    ph = np.linspace(0.0, 2.0 * np.pi, n_u, endpoint=False)
    T, P = np.meshgrid(th, ph, indexing="ij")
    ring = np.stack([a * np.cos(T), b * np.sin(T) * np.cos(P), c * np.sin(T) * np.sin(P)], axis=-1)
    north = np.array([a, 0.0, 0.0])
    south = np.array([-a, 0.0, 0.0])
    verts = np.concatenate([north[None], ring.reshape(-1, 3), south[None]], axis=0)
    verts = verts + rng.uniform(-jitter, jitter, size=verts.shape)
    verts = verts + np.asarray(centre)[None, :]
    nr = n_v - 1
    idx = lambda i, j: 1 + i * n_u + (j % n_u)  # noqa: E731
    tris = []
    for j in range(n_u):
        tris.append((0, idx(0, j), idx(0, j + 1)))
    for i in range(nr - 1):
        for j in range(n_u):
            a0, a1, b0, b1 = idx(i, j), idx(i, j + 1), idx(i + 1, j), idx(i + 1, j + 1)
            tris.append((a0, b0, b1))
            tris.append((a0, b1, a1))
    last = len(verts) - 1
    for j in range(n_u):
        tris.append((last, idx(nr - 1, j + 1), idx(nr - 1, j)))
    tri = np.asarray(tris, dtype=np.int64)
    return verts[tri].reshape(-1, 9).astype(f32)


# Re of the 256x128x104 case whose divergence and reset the reference solver's own
# run records (tests/golden/ref_box_run.json): tau 0.5192
REF_DIVERGING_RE = 300.0


def c2_config(voxelize_fn, nx=256, ny=128, nz=104, scale=1.0, seed=12345, re=None,
              precision="fp32", with_mask=True) -> dict:
    """256x128x104, TRT (magic 3/16) + Smagorinsky Cs 0.1, U0 0.08, no body force, an ellipsoid
    STL (semi-axes 30/12/12 cells, ~200k triangles, vertex jitter 0.05 cells, seeded) voxelized
    by the BVH three-ray vote at cell centres.  tau is stated: 0.5 + 3 U0 D / Re with D = 24 (the
    minor-axis diameter) and Re = 100 -> 0.5576, at which the case runs its 1000 steps with the
    reference's open side faces (at Re 300, tau 0.5192, it diverges after about 450 steps, as the
    reference solver does on it: REF_DIVERGING_RE).  scale=2 scales the body and the x/y extents and
    gives the 512x256x256 (33.5M-cell) case with Re 1000 -> tau 0.51152 (nz is 256 there); other
    scales scale all three extents (0.25: 64x32x26)."""
    if scale == 2.0 and nz == 104:
        nx, ny, nz = int(nx * scale), int(ny * scale), 256
    else:
        nx, ny, nz = int(nx * scale), int(ny * scale), int(nz * scale)
    semi = (30.0 * scale, 12.0 * scale, 12.0 * scale)
    centre = (nx * 0.3, ny * 0.5, nz * 0.5)
    cfg = dict(**lattice_box(nx, ny, nz), voxel_resolution=nx, max_dim=float(2 * semi[0]),
               flow_speed=20.0, U0=0.08, use_les=1, smagorinsky_c=0.1, use_trt=1,
               auto_viscosity=0, inlet_turbulence=0.0, convective_outlet=1, body_force=(0.0, 0.0, 0.0),
               precision=precision, cell_cap=1 << 31)
    D = 24.0 * scale
    Re = re if re is not None else (100.0 if scale == 1.0 else 1000.0)
    cfg["tau"] = float(f32(0.5) + f32(3.0) * f32(0.08) * f32(D) / f32(Re))
    if with_mask:
        tris = ellipsoid_mesh(semi=semi, centre=centre, seed=seed, jitter=0.05)
        cfg["voxels"] = voxelize_fn(tris, (nx, ny, nz), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    else:
        cfg["voxels"] = np.zeros((nz, ny, nx), dtype=np.int32)
    return cfg


def c3_rayleigh_benard(nx=128, ny=64, nz=128, seed=2024, rayleigh=2500.0, prandtl=0.71, tau=0.6):
    """Coupled D3Q19 + D3Q7 Rayleigh-Benard.  The reference's gravity and walls are along y, so
    the 128x128x64 case runs as x=128, y=64 (vertical, solid rows y=0 and y=ny-1 hot/cold),
    z=128, periodic in x and z.  Returns (config overrides, initial T field): T linear between
    the walls plus a seeded U(-1e-3, 1e-3) perturbation."""
    cfg = thermal_box(nx=nx, ny=ny, nz=nz, walls_y=True, u0=0.0, tau=tau, periodic_z=True,
                      rayleigh=rayleigh, prandtl=prandtl, fixed_wall_temps=True, perturbation=0.0)
    cfg["cell_cap"] = 1 << 31
    rng = np.random.Generator(np.random.PCG64(seed))
    y = np.arange(ny, dtype=np.float64)
    s = np.clip((y - 1.0) / max(1.0, ny - 3.0), 0.0, 1.0)
    lin = 1.0 - s
    T = np.broadcast_to(lin[None, :, None], (nz, ny, nx)).copy()
    T += rng.uniform(-1e-3, 1e-3, size=T.shape)
    T[:, 0, :] = 1.0
    T[:, ny - 1, :] = 0.0
    return cfg, T.astype(f32)


def ellipsoid_in_box(nx, ny, nz, centre, semi) -> np.ndarray:
    """Ellipsoid by the cell-centre test sum((x - c)^2 / a^2) <= 1 in lattice units."""
    z, y, x = np.meshgrid(np.arange(nz, dtype=f32) + f32(0.5), np.arange(ny, dtype=f32) + f32(0.5),
                          np.arange(nx, dtype=f32) + f32(0.5), indexing="ij")
    q = sum(((v - f32(c)) / f32(a)) ** 2 for v, c, a in zip((x, y, z), centre, semi))
    return (q <= f32(1.0)).astype(np.int32)


def c5_request(rng, u0=0.05, thermal_frac=0.3, size=None, tau_min=0.51, dcap=0.3, xc=1.0 / 3.0,
               side_walls=True) -> tuple[dict, dict]:
    """One served-mode request: grid on the line 32^3 .. 160x80x80, a sphere or an
    ellipsoid (3:2:2) of diameter U[6, 20] cells (at most dcap of the cross-section) at a
    third of the length, Re U[50, 500], and with probability thermal_frac a thermal
    lattice with Ra U[1800, 5000], Pr 0.71, hot bottom / cold top walls.  TRT +
    Smagorinsky, zero-gradient outflow, U0 = 0.05, free-slip side walls (y and z faces).

    tau = 1/2 + 3 U0 D / Re, raised to tau_min where it falls below (the set-up's own
    floor is 0.51, aeros/src/lbm.cpp:741-815), so the realised Reynolds number is
    3 U0 D / (tau - 1/2): at most 15 D at U0 = 0.05.  With the reference's open side
    faces (a source outside the domain keeps the cell's own population) these small
    lattices diverge and reset every few hundred steps at tau near 0.51 (measured);
    free-slip side walls keep them stable (every non-thermal request of the seeded
    stream converges in FP32, measured).  The drag is measured on a control volume of
    0.40 of the domain around its centre, which encloses the body at every size."""
    t = rng.uniform(0.0, 1.0) if size is None else size
    nx = int(round(32 + 128 * t))
    ny = nz = int(round(32 + 48 * t))
    d = min(float(rng.uniform(6.0, 20.0)), dcap * ny)
    ellip = bool(rng.uniform() < 0.5)
    re = float(rng.uniform(50.0, 500.0))
    thermal = bool(rng.uniform() < thermal_frac)
    ra = float(rng.uniform(1800.0, 5000.0)) if thermal else 0.0
    centre = (nx * xc, ny / 2.0, nz / 2.0)
    if ellip:
        vox = ellipsoid_in_box(nx, ny, nz, centre, (0.75 * d, 0.5 * d, 0.5 * d))
    else:
        vox = sphere_in_box(nx, ny, nz, centre, d)
    tau = max(tau_min, 0.5 + 3.0 * u0 * d / re)
    cfg = dict(voxels=vox, **lattice_box(nx, ny, nz), voxel_resolution=nx, max_dim=d, flow_speed=20.0, U0=u0,
               tau=float(tau), use_les=1, use_trt=1, smagorinsky_c=0.1, auto_viscosity=0, inlet_turbulence=0.0,
               convective_outlet=0, body_force=(0.0, 0.0, 0.0), free_slip=SIDE_WALLS if side_walls else 0)
    if thermal:
        cfg.update(thermal=1, rayleigh=ra, prandtl=0.71, fixed_wall_temps=1, hot_temp=1.0, cold_temp=0.0,
                   perturbation=0.02)
    meta = dict(nx=nx, ny=ny, nz=nz, diameter=d, ellipsoid=ellip, re=re, re_eff=3.0 * u0 * d / (max(tau, 0.51) - 0.5),
                tau=tau, thermal=thermal, rayleigh=ra, cells=nx * ny * nz)
    return cfg, meta


C5_CV_HALF = 0.40


def c5_requests(n=256, seed=5150, **kw) -> list[tuple[dict, dict]]:
    """The served-mode stream: n seeded requests (PCG64(seed))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return [c5_request(rng, **kw) for _ in range(n)]


def reference_field_cases() -> dict:
    """The small cases whose per-cell fields the reference solver dumped
    (tests/golden/ref_fields.json): name -> (config overrides, steps)."""
    return {
        "bench12_les": (bench_sphere(12, les=True), 90),
        "bench12_les_thermal": (bench_sphere(12, les=True, thermal=True), 90),
        "bench16": (bench_sphere(16, les=False), 150),
        "rb_walls": (thermal_box(32, 18, 8, walls_y=True, rayleigh=2500.0, fixed_wall_temps=True,
                                 perturbation=0.02), 200),
        # buoyancy large enough to overflow the populations: the NaN-storm reset
        "rb_storm": (thermal_box(32, 18, 8, walls_y=True, rayleigh=1.0e30, fixed_wall_temps=True,
                                 perturbation=0.02), 12),
        # tau from the Reynolds number (autoViscosity, aeros/src/lbm.cpp:741-815) in the
        # standard atmosphere at 1500 m: at 0.0004 m/s tau lands inside [0.51, 1.5], at
        # the bench's 20 m/s it clamps to 0.51
        "bench12_auto": (bench_sphere_auto(0.0004), 60),
        "bench12_auto_clamped": (bench_sphere_auto(20.0), 60),
    }


# airTemperature and airDensity of the reference's standard atmosphere
# (updateAtmosphereParams, aeros/src/atmosphere.cpp:176-191) at sea level and
# at 1500 m, as float32
AIR_SEA_LEVEL = (288.149994, 1.2250123)
AIR_1500M = (278.399994, 1.05807602)


def bench_sphere_auto(flow_speed: float) -> dict:
    """The bench sphere at n = 12 with LES and tau from the Reynolds number
    (autoViscosity): Sutherland viscosity at the air temperature, Re = rho V L / mu
    with L = max_dim, then tau = 1/2 + 3 U0 L_LB / Re clamped to [0.51, 1.5]."""
    cfg = bench_sphere(12, les=True)
    cfg.update(auto_viscosity=1, flow_speed=float(np.float32(flow_speed)), air_temperature=AIR_1500M[0],
               air_density=AIR_1500M[1])
    return cfg
