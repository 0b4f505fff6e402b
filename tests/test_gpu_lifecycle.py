"""GPU: re-initialisation entry points through the C ABI, against a fresh
lattice and the oracle.

reset (resetLBM, aeros/src/lbm.h:157), set-mask (initLBM after a new mask,
aeros/src/lbm.cpp:596-815), thermalUpload and thermalResetField
(aeros/src/thermal_lbm.cpp:154-197): each must leave the lattice exactly as
a fresh create (or the oracle given the same calls) would, bit for bit in
FP32.  Also tau from the Reynolds number and the FP64 channel's closed form.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import aeros_lbm as A
import workloads as W
from test_set_mask import ground_case

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
FIELDS = ("rho", "ux", "uy", "uz", "tke", "strain")


def assert_same(a, b, names=FIELDS):
    for name in names:
        x, y = a.field(name), b.field(name)
        bad = np.argwhere(x != y)
        assert bad.size == 0, f"{name}: {len(bad)} cells differ, first {bad[:3].tolist()}"


@pytest.mark.parametrize("precision", ["fp32", "fp16", "fp64"])
def test_reset_then_step_equals_a_fresh_lattice(precision):
    cfg = dict(W.bench_sphere(12, les=True, thermal=precision != "fp64"), precision=precision)
    print(f"reset {precision}", flush=True)
    L = A.Lattice(A.default_config(**cfg))
    L.step(37)
    L.reset()
    assert L.scalars()["step"] == 0
    fresh = A.Lattice(A.default_config(**cfg))
    for X in (L, fresh):
        X.step(1)
        X.step(22)
    names = FIELDS + (("T", "buoyancy") if precision != "fp64" else ())
    assert_same(L, fresh, names)
    np.testing.assert_array_equal(L.populations(), fresh.populations())
    sl, sf = L.scalars(), fresh.scalars()
    for k in ("step", "avg_rho", "avg_ke", "max_u_lb", "tke", "neg_pops", "nusselt"):
        assert sl[k] == sf[k], (k, sl[k], sf[k])


def test_set_mask_keeps_the_ground_slab(oracle):
    """A use_ground case: set-mask with the lattice's own solid mask rebuilds the
    ground slab, so the lattice equals a fresh one and the oracle after the same calls."""
    print("set_mask ground", flush=True)
    fresh = A.Lattice(A.default_config(**ground_case()))
    L = A.Lattice(A.default_config(**ground_case()))
    O = oracle.OracleLattice(oracle.default_config(**ground_case()))
    assert int(fresh.field("ground").sum()) > 0
    L.step(9)
    O.step(9)
    solid = L.field("solid").astype(np.uint8)
    L.set_mask(solid)
    O.set_mask(solid)
    for X in (L, fresh, O):
        X.step(30)
    assert_same(L, fresh, FIELDS + ("ground", "walldist"))
    assert_same(L, O, FIELDS + ("ground", "walldist"))
    assert L.scalars()["step"] == fresh.scalars()["step"] == 30


@pytest.mark.parametrize("thermal", [False, True])
def test_set_mask_equals_a_fresh_lattice_with_that_mask(oracle, thermal):
    if thermal:
        cfg = W.thermal_box(32, 18, 8, walls_y=True, rayleigh=2500.0, fixed_wall_temps=True, perturbation=0.02)
        b = np.array(cfg["voxels"])
        b[:, 6:9, 10:14] = 1
    else:
        cfg = W.slip_box(40, 20, 18, (10, 10, 9), 6, 0, use_les=1, auto_viscosity=1, air_temperature=W.AIR_1500M[0],
                         air_density=W.AIR_1500M[1], flow_speed=0.0004)
        b = W.sphere_in_box(40, 20, 18, (16, 9, 9), 8)
    print(f"set_mask new mask thermal={thermal}", flush=True)
    L = A.Lattice(A.default_config(**cfg))
    O = oracle.OracleLattice(oracle.default_config(**cfg))
    L.step(5)
    O.step(5)
    L.set_mask(b.astype(np.uint8))
    O.set_mask(b.astype(np.uint8))
    fresh = A.Lattice(A.default_config(**dict(cfg, voxels=b)))
    names = FIELDS + ("solid", "walldist") + (("T", "buoyancy") if thermal else ())
    for X in (L, fresh, O):
        X.step(30)
    assert_same(L, fresh, names)
    assert_same(L, O, names)
    for k in ("tau", "reynolds_phys", "reynolds_lb"):
        assert L.scalars()[k] == fresh.scalars()[k], k


def test_set_mask_rejects_a_wrong_shape_and_keeps_the_lattice():
    cfg = W.bench_sphere(12)
    L = A.Lattice(A.default_config(**cfg))
    ref = A.Lattice(A.default_config(**cfg))
    L.step(10)
    ref.step(10)
    nz, ny, nx = L.shape
    with pytest.raises(A.AerosError):
        L.set_mask(np.zeros((nz, ny, nx + 1), dtype=np.uint8))
    for X in (L, ref):
        X.step(10)
    assert_same(L, ref)


def seeded_T(shape, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return (0.5 + rng.uniform(-0.4, 0.4, size=shape)).astype(np.float32)


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_thermal_upload_against_the_oracle(oracle, precision):
    """thermalUpload of a seeded field (solid rows included) on the Rayleigh-Benard
    box, then steps: FP32 bit for bit with the oracle, the temperature of solid
    cells kept as uploaded; FP16 storage (FP32 thermal lattice) to the error of the
    quantised oracle against FP64."""
    cfg = dict(W.thermal_box(32, 18, 8, walls_y=True, rayleigh=2500.0, fixed_wall_temps=False, perturbation=0.0),
               precision=precision)
    print(f"thermal upload {precision}", flush=True)
    L = A.Lattice(A.default_config(**cfg))
    O = oracle.OracleLattice(oracle.default_config(**cfg))
    T = seeded_T(L.shape, 99)
    L.step(3)
    O.step(3)
    L.thermal_upload(T)
    O.thermal_upload(T)
    assert L.scalars()["thermal_mean_t"] == O.scalars()["thermal_mean_t"]
    solid = L.field("solid") > 0.5
    for n in (1, 40):
        L.step(n)
        O.step(n)
        np.testing.assert_array_equal(L.field("T")[solid], T[solid])
        if precision == "fp32":
            assert_same(L, O, FIELDS + ("T", "buoyancy"))
            assert L.scalars()["nusselt"] == O.scalars()["nusselt"]
    if precision != "fp32":
        O64 = oracle.OracleLattice(oracle.default_config(**dict(cfg, precision="fp64")))
        O64.step(3)
        O64.thermal_upload(T)
        O64.step(41)
        for name in ("T", "uy"):
            e = O64.field(name, dtype=np.float64)
            g = np.abs(L.field(name, dtype=np.float64) - e)
            q = np.abs(O.field(name, dtype=np.float64) - e)
            assert np.sqrt((g ** 2).mean()) <= 1.25 * np.sqrt((q ** 2).mean()) + 1e-7, name


def test_thermal_reset_against_the_oracle(oracle):
    cfg = W.thermal_box(32, 18, 8, walls_y=True, rayleigh=2500.0, fixed_wall_temps=True, perturbation=0.02)
    print("thermal reset", flush=True)
    L = A.Lattice(A.default_config(**cfg))
    O = oracle.OracleLattice(oracle.default_config(**cfg))
    for X in (L, O):
        X.step(50)
        X.thermal_reset()
    assert L.scalars()["thermal_mean_t"] == O.scalars()["thermal_mean_t"]
    for X in (L, O):
        X.step(30)
    assert_same(L, O, FIELDS + ("T", "buoyancy"))
    assert L.scalars()["nusselt"] == O.scalars()["nusselt"]


def test_reynolds_number_and_tau_from_re():
    """initLBM's Re for the lbm suite's sphere at sea level (%.0f as the suite prints
    it), and tau from Re in the 1500 m atmosphere, inside and at the clamp, as the
    reference's initLBM derived them."""
    lbm = json.loads((GOLD / "ref_lbm.json").read_text())
    t, rho = W.AIR_SEA_LEVEL
    L = A.Lattice(A.default_config(**W.lbm_suite_sphere(), air_temperature=t, air_density=rho))
    assert "%.0f" % L.scalars()["reynolds_phys"] == "%.0f" % lbm["re_phys"]
    L.close()
    fields = json.loads((GOLD / "ref_fields.json").read_text())["cases"]
    for case, speed in (("bench12_auto", 0.0004), ("bench12_auto_clamped", 20.0)):
        sc = A.Lattice(A.default_config(**W.bench_sphere_auto(speed))).scalars()
        for key, name in (("tau", "tau"), ("viscosity", "viscosity"), ("re_phys", "reynolds_phys"),
                          ("re_lb", "reynolds_lb")):
            assert np.float32(sc[name]) == np.float32(fields[case]["init"][key]), (case, key)


@pytest.mark.parametrize("tau,trt,steps", [(0.65, False, 40000), (0.65, True, 40000)])
def test_poiseuille_fp64_closed_form_gpu(tau, trt, steps):
    """The FP64 HIP channel against u(s) = F/(2 nu) (H^2/4 - s^2 + (16 L - 3)/12),
    L = (tau - 1/2)(1/s- - 1/2) with the reference's s- rule, within 1e-6 of u_max."""
    F, ny = 7.8e-6, 34
    L = A.Lattice(A.default_config(**W.channel(ny, tau, F, trt=trt), precision="fp64"))
    L.step(steps)
    ux = L.field("ux", dtype=np.float64)
    nu = (tau - 0.5) / 3.0
    s_minus = 1.0 / max(0.05, 0.1875 / (1.0 / tau - 0.5) + 0.5) if trt else 1.0 / tau
    lam = (tau - 0.5) * (1.0 / s_minus - 0.5)
    s = np.arange(1, ny - 1) - 0.5 * (ny - 1)
    exact = F / (2.0 * nu) * ((ny - 2) ** 2 / 4.0 - s * s + (16.0 * lam - 3.0) / 12.0)
    dev = np.abs(ux[:, 1:ny - 1, :] - exact[None, :, None]).max() / exact.max()
    assert dev < 1e-6, dev
