"""Post-processing restatements (host, no GPU), pinned by closed forms: the
vorticity / Q / strain pass (aeros/src/lbm.cpp:2005-2093) and the world-space
samplers (aeros/src/lbm.cpp:1781-1999) that the device kernels are tested
against bit for bit."""
import numpy as np

import oracle as O

f32 = np.float32


def grid(nz=6, ny=7, nx=8):
    return np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")


def test_solid_body_rotation():
    """u = omega x r about z: |vorticity| = 2 w / ... in 1/s with the cell size, Q = w^2, no strain."""
    z, y, x = grid()
    w, cs = 0.3, 0.5
    ux = (-w * (y - 3) * cs).astype(f32)
    uy = (w * (x - 4) * cs).astype(f32)
    uz = np.zeros_like(ux)
    v, q, s = O.vorticity_q(ux, uy, uz, np.zeros(ux.shape, bool), (cs, cs, cs))
    I = (slice(1, -1),) * 3
    np.testing.assert_allclose(v[I], 2 * w, rtol=1e-6)
    np.testing.assert_allclose(q[I], w * w, rtol=1e-5)
    np.testing.assert_allclose(s[I], 0.0, atol=1e-6)
    assert (v[0] == 0).all() and (q[:, :, -1] == 0).all()   # shell untouched (zero)


def test_simple_shear_and_solid_cells():
    """u = (a y, 0, 0): |vorticity| = a, Q = 0, strain sqrt(2 S:S) = a; solid cells 0."""
    z, y, x = grid()
    a = 0.7
    ux = (a * y).astype(f32)
    zero = np.zeros_like(ux)
    solid = np.zeros(ux.shape, bool)
    solid[3, 3, 3] = True
    v, q, s = O.vorticity_q(ux, zero, zero, solid, (1.0, 1.0, 1.0))
    I = (slice(1, -1),) * 3
    fluid = ~solid[I]
    np.testing.assert_allclose(v[I][fluid], a, rtol=1e-6)
    np.testing.assert_allclose(q[I][fluid], 0.0, atol=1e-6)
    np.testing.assert_allclose(s[I][fluid], a, rtol=1e-6)
    assert v[3, 3, 3] == 0 and q[3, 3, 3] == 0


def test_trilinear_sampler_reproduces_linear_fields():
    z, y, x = grid(5, 6, 7)
    mn, cs = (-1.0, 2.0, 0.5), (0.5, 0.25, 1.0)
    wx, wy, wz = mn[0] + x * cs[0], mn[1] + y * cs[1], mn[2] + z * cs[2]
    field = (1.0 + 2.0 * wx + 3.0 * wy - wz).astype(f32)
    solid = np.zeros(field.shape, bool)
    rng = np.random.Generator(np.random.PCG64(3))
    for _ in range(50):
        p = [mn[k] + rng.uniform(0.0, (n - 1.0001) * cs[k]) for k, n in zip(range(3), (7, 6, 5))]
        v, ok = O.sample_field(field, solid, mn, cs, p)
        assert ok and abs(v - (1 + 2 * p[0] + 3 * p[1] - p[2])) < 2e-5
    # exact at lattice nodes
    v, ok = O.sample_field(field, solid, mn, cs, (mn[0] + 2 * cs[0], mn[1] + 3 * cs[1], mn[2] + 1 * cs[2]))
    assert ok and v == field[1, 3, 2]
    # outside the lattice: no data, the named defaults
    assert O.sample_field(field, solid, mn, cs, (mn[0] - 0.01, 2.5, 1.0)) == (0.0, False)
    assert O.sample_field(field, solid, mn, cs, (mn[0] - 0.01, 2.5, 1.0), clamp_density=True) == (1.0, False)


def test_sampler_skips_solid_corners():
    field = np.zeros((2, 2, 2), f32)
    field[:, :, 1] = 4.0
    solid = np.zeros(field.shape, bool)
    solid[:, :, 0] = True          # only the x = 1 corners are fluid
    v, ok = O.sample_field(field, solid, (0, 0, 0), (1, 1, 1), (0.25, 0.5, 0.5))
    assert ok and v == 4.0         # renormalised over the fluid corners
    solid[:] = True
    assert O.sample_field(field, solid, (0, 0, 0), (1, 1, 1), (0.25, 0.5, 0.5)) == (0.0, False)


def test_velocity_samplers():
    shape = (4, 5, 6)
    ux = np.full(shape, 0.05, f32)
    uy = np.zeros(shape, f32)
    uz = np.zeros(shape, f32)
    solid = np.zeros(shape, bool)
    solid[1, 2, 3] = True
    mn, cs = (0.0, 0.0, 0.0), (0.1, 0.1, 0.1)
    v = O.velocity_lb(ux, uy, uz, solid, mn, cs, (0.21, 0.13, 0.17))
    np.testing.assert_allclose(v, [0.05, 0, 0], rtol=1e-6)
    inlet, speed, fs = (0.05, 0.0, 0.0), 20.0, (20.0, 0.0, 0.0)
    w = O.velocity_world(ux, uy, uz, solid, mn, cs, (0.21, 0.13, 0.17), inlet, speed, fs)
    np.testing.assert_allclose(w, [20.0, 0, 0], rtol=1e-6)                      # scaled by flowSpeed / U0
    assert (O.velocity_world(ux, uy, uz, solid, mn, cs, (0.35, 0.25, 0.15), inlet, speed, fs) == 0).all()  # solid cell
    np.testing.assert_array_equal(O.velocity_world(ux, uy, uz, solid, mn, cs, (-1, 0.2, 0.2), inlet, speed, fs), fs)
    fast = np.full(shape, 0.2, f32)   # 80 m/s > 2 flowSpeed: capped at 40
    w = O.velocity_world(fast, uy, uz, solid, mn, cs, (0.21, 0.13, 0.17), inlet, speed, fs)
    np.testing.assert_allclose(np.linalg.norm(w), 40.0, rtol=1e-6)
