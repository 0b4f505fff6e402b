"""GPU: the allocation-free sampler (aeros_lbm_sample_noalloc) against
aeros_lbm_sample, through the C ABI.

Both run the same kernel; the allocation-free one stages positions, values and
flags in the library's fixed device array instead of per-call buffers.  Results
must be the same bits, on two calls in a row with the second smaller than the
first (the array keeps the first call's contents), and on a second lattice of
another size afterwards.
"""
import numpy as np
import pytest

import aeros_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("pressure", "rho", "tke", "strain", "vorticity", "qcrit")
MAX_POINTS = 65536


def points(L, n, seed):
    """n points in and up to two cells around the lattice box, lattice nodes among them."""
    rng = np.random.Generator(np.random.PCG64(seed))
    mn, cs, dims = np.array(L.box_min), np.array(L.cell_size), np.array(L.dims)
    p = rng.uniform(mn - 2 * cs, mn + dims * cs + 2 * cs, size=(n, 3))
    p[: n // 8] = mn + rng.integers(0, dims, size=(n // 8, 3)) * cs
    return p.astype(np.float32)


def same(L, name, pos):
    a, va = L.sample(name, pos)
    b, vb = L.sample(name, pos, noalloc=True)
    assert a.tobytes() == b.tobytes(), (name, int(np.count_nonzero(a != b)))
    np.testing.assert_array_equal(va, vb)
    return va


def test_noalloc_sampler_matches_the_allocating_one():
    print("noalloc sampler: bench_sphere(16), 40 steps", flush=True)
    L = A.Lattice(A.default_config(**W.bench_sphere(16)))
    L.step(40)
    with pytest.raises(A.AerosError):   # vorticity/Q not on the device yet, and nothing may be allocated
        L.sample("vorticity", points(L, 16, 1), noalloc=True)
    L.vorticity_q()
    for name in FIELDS:
        valid = same(L, name, points(L, MAX_POINTS, 2))      # the largest call the staging holds
        assert valid.any() and not valid.all()
        same(L, name, points(L, 777, 3))                      # then a smaller one
    with pytest.raises(A.AerosError):
        L.sample("rho", points(L, MAX_POINTS + 1, 4), noalloc=True)
    same(L, "rho", np.zeros((0, 3), np.float32))
    L.close()
    print("noalloc sampler: a second lattice (bench_sphere(12), thermal)", flush=True)
    L2 = A.Lattice(A.default_config(**W.bench_sphere(12, thermal=True)))
    L2.step(25)
    L2.vorticity_q()
    for name in FIELDS:
        same(L2, name, points(L2, 5000, 5))
    L2.close()
