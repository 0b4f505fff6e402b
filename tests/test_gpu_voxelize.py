"""GPU: the device voxelizer (aeros_voxelize_gpu) against the host one.

Both cast the three jittered axis rays of the reference's vote
(aeros/src/voxel_grid.cpp:337-351) through the same BVH with the same float
slab and double-precision triangle tests, so the masks must be identical bit
for bit; the host voxelizer itself is pinned against brute force and the
reference's solid counts by the CPU tests.
"""
import numpy as np
import pytest

import aeros_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu


def cube_mesh(lo, hi):
    """Closed axis-aligned box, 12 triangles, (12, 9) float32."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], dtype=np.float32)
    f = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4),
         (2, 3, 7), (2, 7, 6), (1, 2, 6), (1, 6, 5), (0, 4, 7), (0, 7, 3)]
    return v[np.array(f)].reshape(-1, 9)


def check(tris, dims, box_min, cell):
    print(f"voxelize {len(tris)} triangles on {dims}", flush=True)
    host = A.voxelize(tris, dims, box_min, cell)
    gpu = A.voxelize_gpu(tris, dims, box_min, cell)
    bad = np.argwhere(host != gpu)
    assert bad.size == 0, f"{len(bad)} voxels differ, first {bad[:3].tolist()}"
    return host


def test_c2_ellipsoid_stl_mask():
    """The 199,080-triangle jittered ellipsoid of the 3.41M-cell case."""
    tris = W.ellipsoid_mesh(semi=(30.0, 12.0, 12.0), centre=(256 * 0.3, 64.0, 52.0), seed=12345, jitter=0.05)
    m = check(tris, (256, 128, 104), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    # volume of the voxelized body against the analytic ellipsoid volume
    assert abs(int(m.sum()) / (4.0 / 3.0 * np.pi * 30 * 12 * 12) - 1.0) < 0.02


def test_offset_grid_and_coarse_mesh():
    tris = W.ellipsoid_mesh(semi=(9.0, 5.0, 4.0), centre=(1.0, -2.0, 0.5), n_u=40, n_v=24, seed=7, jitter=0.2)
    m = check(tris, (48, 32, 28), (-10.3, -9.1, -6.7), (0.45, 0.5, 0.55))
    assert 0 < int(m.sum()) < m.size


def test_box_with_faces_on_cell_centres():
    tris = cube_mesh((3.5, 2.5, 4.5), (11.5, 9.5, 8.5))
    m = check(tris, (16, 12, 14), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    assert int(m.sum()) > 0


def test_empty_degenerate_and_non_finite_meshes():
    dims = (8, 8, 8)
    assert int(check(np.zeros((0, 9), np.float32), dims, (0, 0, 0), (1, 1, 1)).sum()) == 0
    degenerate = np.array([[1, 1, 1, 2, 2, 2, 3, 3, 3]], dtype=np.float32)   # collinear
    assert int(check(degenerate, dims, (0, 0, 0), (1, 1, 1)).sum()) == 0
    box = cube_mesh((1.2, 1.3, 1.4), (6.6, 6.7, 6.8))
    nan_tri = np.full((1, 9), np.nan, dtype=np.float32)
    a = check(np.concatenate([box, nan_tri]), dims, (0, 0, 0), (1, 1, 1))
    b = A.voxelize(box, dims, (0, 0, 0), (1, 1, 1))
    np.testing.assert_array_equal(a, b)


def test_device_vote_matches_the_reference_on_the_ellipsoid_stls():
    """The device voxelizer against the reference's own vote (SHA-256 of the
    uint8 masks of the C2 body and its x2 scale)."""
    import hashlib
    import json
    from pathlib import Path
    ref = json.loads((Path(__file__).resolve().parent / "golden" / "ref_voxels.json").read_text())["cases"]
    for case, scale in (("c2", 1.0), ("c4", 2.0)):
        print(f"reference vote {case}", flush=True)
        m = np.ascontiguousarray(W.c2_config(A.voxelize_gpu, scale=scale)["voxels"]).astype(np.uint8)
        assert int(m.sum()) == ref[case]["solid"]
        assert hashlib.sha256(m.tobytes()).hexdigest() == ref[case]["sha256"], case
