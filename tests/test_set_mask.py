"""CPU: the oracle's set-mask path (initLBM after a new mask, aeros/src/lbm.cpp:596-815).

A new solid mask on a live lattice must give exactly the lattice a fresh
create with that mask gives: the ground slab goes back on top of the mask
(aeros/src/lbm.cpp:624-643), then wall distance, fluid count, tau and the
thermal module are rebuilt from the caller's parameters.
"""
import numpy as np
import pytest

import workloads as W

FIELDS = ("rho", "ux", "uy", "uz", "tke", "strain", "solid", "ground", "walldist")


def ground_case():
    """The lbm suite's sphere with its ground slab on (the suite's ground_on variant)."""
    return dict(W.lbm_suite_sphere(), use_ground=1, ground_height=0.0)


def assert_same(a, b, names=FIELDS):
    for name in names:
        np.testing.assert_array_equal(a.field(name), b.field(name), err_msg=name)


def test_set_mask_keeps_the_ground_slab(oracle):
    fresh = oracle.OracleLattice(oracle.default_config(**ground_case()))
    L = oracle.OracleLattice(oracle.default_config(**ground_case()))
    assert int(fresh.field("ground").sum()) > 0
    L.step(7)
    L.set_mask(L.field("solid").astype(np.uint8))
    assert_same(L, fresh)
    assert L.scalars()["step"] == 0
    for X in (L, fresh):
        X.step(25)
    assert_same(L, fresh)
    np.testing.assert_array_equal(L.populations(), fresh.populations())


@pytest.mark.parametrize("thermal", [False, True])
def test_set_mask_equals_a_fresh_lattice_with_that_mask(oracle, thermal):
    """Lattice-aligned boxes: set_mask(B) on a lattice made with mask A equals a
    lattice made with B (tau from Re on, so tau is derived again from the same
    parameters and not carried over from A's init)."""
    if thermal:
        cfg = W.thermal_box(32, 18, 8, walls_y=True, rayleigh=2500.0, fixed_wall_temps=True, perturbation=0.02)
        b = np.array(cfg["voxels"])
        b[:, 6:9, 10:14] = 1
    else:
        cfg = W.slip_box(40, 20, 18, (10, 10, 9), 6, 0, use_les=1, auto_viscosity=1, air_temperature=W.AIR_1500M[0],
                         air_density=W.AIR_1500M[1], flow_speed=0.0004)
        b = W.sphere_in_box(40, 20, 18, (16, 9, 9), 8)
    L = oracle.OracleLattice(oracle.default_config(**cfg))
    L.step(5)
    L.set_mask(b.astype(np.uint8))
    fresh = oracle.OracleLattice(oracle.default_config(**dict(cfg, voxels=b)))
    names = FIELDS + (("T", "buoyancy") if thermal else ())
    assert_same(L, fresh, names)
    sl, sf = L.scalars(), fresh.scalars()
    for k in ("tau", "viscosity", "reynolds_phys", "reynolds_lb"):
        assert sl[k] == sf[k], k
    for X in (L, fresh):
        X.step(30)
    assert_same(L, fresh, names)
