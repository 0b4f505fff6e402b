"""GPU: device post-processing against its restatement, bit for bit.

aeros_lbm_vorticity_q against oracle.vorticity_q on the same world velocity
(aeros/src/lbm.cpp:2005-2093); aeros_lbm_sample / aeros_lbm_sample_velocity
against the restated samplers (aeros/src/lbm.cpp:1781-1999) on the fields the
GPU holds, at random points in and around the lattice, in solid cells and on
lattice nodes."""
import numpy as np
import pytest

import aeros_lbm as A
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lattice():
    case = W.bench_sphere(12)
    print("case bench_sphere(12), 30 steps", flush=True)
    L = A.Lattice(A.default_config(**case))
    L.step(30)
    yield L
    L.close()


def solid_of(L):
    return (L.field("solid") != 0) | (L.field("ground") != 0)


def test_vorticity_q_matches_restatement(lattice):
    L = lattice
    strain_before = L.field("strain")
    L.vorticity_q()
    vg, qg, sg = L.field("vorticity"), L.field("qcrit"), L.field("strain")
    vo, qo, so = O.vorticity_q(L.field("ux_world"), L.field("uy_world"), L.field("uz_world"), solid_of(L),
                               L.cell_size)
    np.testing.assert_array_equal(vg, vo)
    np.testing.assert_array_equal(qg, qo)
    written = ~np.isnan(so)
    np.testing.assert_array_equal(sg[written], so[written])
    np.testing.assert_array_equal(sg[~written], strain_before[~written])   # the rest keeps the solver's |S|
    assert (vg > 0).any() and (qg != 0).any()


def points(L, n=400, seed=9):
    rng = np.random.Generator(np.random.PCG64(seed))
    mn, cs, dims = np.array(L.box_min), np.array(L.cell_size), np.array(L.dims)
    mx = mn + dims * cs
    p = rng.uniform(mn - 2 * cs, mx + 2 * cs, size=(n, 3))
    nodes = mn + rng.integers(0, dims, size=(64, 3)) * cs            # exactly on lattice nodes
    solid = np.argwhere(solid_of(L))[:32][:, ::-1]                     # (x, y, z) of solid cells
    inside = mn + (solid + 0.5) * cs
    return np.concatenate([p, nodes, inside]).astype(np.float32)


@pytest.mark.parametrize("name", ["pressure", "rho", "tke", "vorticity", "qcrit"])
def test_scalar_samplers(lattice, name):
    L = lattice
    L.vorticity_q()
    pos = points(L)
    vals, valid = L.sample(name, pos)
    field, solid = L.field(name), solid_of(L)
    ref = [O.sample_field(field, solid, L.box_min, L.cell_size, p, clamp_density=name == "rho") for p in pos]
    np.testing.assert_array_equal(vals, np.array([r[0] for r in ref], np.float32))
    np.testing.assert_array_equal(valid, np.array([r[1] for r in ref]))
    assert valid.any() and not valid.all()


def test_velocity_samplers(lattice):
    L = lattice
    pos = points(L)
    ux, uy, uz, solid = L.field("ux"), L.field("uy"), L.field("uz"), solid_of(L)
    lb = L.sample_velocity(pos, world=False)
    ref = np.array([O.velocity_lb(ux, uy, uz, solid, L.box_min, L.cell_size, p) for p in pos])
    np.testing.assert_array_equal(lb, ref)
    cfg = A.default_config(**W.bench_sphere(12))
    inlet = (cfg.U0, 0.0, 0.0)   # flow along +x at azimuth/elevation 0
    wv = L.sample_velocity(pos, world=True)
    ref = np.array([O.velocity_world(ux, uy, uz, solid, L.box_min, L.cell_size, p, inlet, cfg.flow_speed,
                                     (cfg.flow_speed, 0.0, 0.0)) for p in pos])
    np.testing.assert_array_equal(wv, ref)
