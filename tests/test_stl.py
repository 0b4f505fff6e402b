"""STL input/output (host, no GPU): the cases of the reference's STL suite
(aeros/tests/stl_headless.cpp:103-233) on the same synthetic buffers, and a
binary round trip of the synthetic ellipsoid STL through the voxelizer."""
import math
import struct

import numpy as np
import pytest

import aeros_lbm as A
import workloads as W


def make_binary(triangles, header="binary aeros test", trailing=0, nan_first=False):
    """Header (80 bytes), count, 50 bytes per triangle: normal (0,0,1), vertices
    (2i,0,0), (2i+1,0,0), (2i,1,0), attribute 0 (aeros/tests/stl_headless.cpp:48-73)."""
    d = bytearray(80)
    h = header.encode()[:80]
    d[:len(h)] = h
    d += struct.pack("<I", triangles)
    for i in range(triangles):
        b = 2.0 * i
        d += struct.pack("<12f", 0, 0, 1, b, 0, 0, b + 1, 0, 0, b, 1, 0) + b"\0\0"
    if nan_first and len(d) >= 84 + 24:
        d[96:100] = struct.pack("<f", math.nan)
    d += bytes([0xA5]) * trailing
    return bytes(d)


def make_ascii(triangles, solid=True):
    s = "solid test\n" if solid else ""
    for _ in range(triangles):
        s += "  facet normal 0 0 1\n    outer loop\n"
        s += "      vertex 0 0 0\n      vertex 1 0 0\n      vertex 0 1 0\n"
        s += "    endloop\n  endfacet\n"
    if solid:
        s += "endsolid test\n"
    return s.encode()


def test_binary_exact_size():
    d = make_binary(2)
    info = A.stl_detect(d)
    assert info["format"] == "binary" and info["declared_triangles"] == 2 and info["trailing_bytes"] == 0
    v, n, dropped = A.stl_parse(d)
    assert v.shape == (2, 9) and n.shape == (2, 9) and dropped == 0
    np.testing.assert_array_equal(v[1], [2, 0, 0, 3, 0, 0, 2, 1, 0])


def test_binary_with_tail():
    d = make_binary(3, trailing=16)
    info = A.stl_detect(d)
    assert info["format"] == "binary" and info["trailing_bytes"] == 16
    assert A.stl_parse(d)[0].shape == (3, 9)


def test_binary_with_zero_triangles_is_rejected():
    d = make_binary(0)
    assert A.stl_detect(d)["format"] != "binary"
    with pytest.raises(A.AerosError):
        A.stl_parse(d)


def test_ascii():
    d = make_ascii(2)
    assert A.stl_detect(d)["format"] == "ascii"
    v, n, _ = A.stl_parse(d)
    assert v.shape == (2, 9)
    assert abs(n[0, 2] - 1.0) < 1e-5   # normal normalised to (0, 0, 1)


def test_binary_header_starting_with_solid():
    d = make_binary(2, header="solid exported by CAD")
    assert A.stl_detect(d)["format"] == "binary"
    assert A.stl_parse(d)[0].shape == (2, 9)


def test_truncated_binary_is_rejected():
    d = make_binary(100)[:84 + 10 * 50 + 7]
    info = A.stl_detect(d)
    assert info["format"] != "binary" and info["reason"] == "truncated"
    with pytest.raises(A.AerosError):
        A.stl_parse(d)


def test_nan_triangle_dropped():
    v, _, dropped = A.stl_parse(make_binary(2, header="binary", nan_first=True))
    assert v.shape == (1, 9) and dropped == 1 and np.isfinite(v).all()


def test_millimetre_coordinates_kept():
    d = bytearray(80)
    d[:15] = b"binary mm model"
    d += struct.pack("<I", 1) + struct.pack("<12f", 0, 0, 1, 1250, 0, 0, 1251, 0, 0, 1250, 1, 0) + b"\0\0"
    v, _, _ = A.stl_parse(bytes(d))
    assert v.shape == (1, 9) and abs(v[0, 0] - 1250.0) < 1e-3 and abs(v[0, 3] - 1251.0) < 1e-3


def test_junk_and_empty():
    junk = b"hello, this is not an stl file at all\n"
    assert A.stl_detect(junk)["format"] == "unknown"
    with pytest.raises(A.AerosError):
        A.stl_parse(junk)
    assert A.stl_detect(b"")["format"] == "unknown"


def test_ellipsoid_binary_round_trip_and_mask(tmp_path):
    """The synthetic jittered ellipsoid written as binary STL and read back gives
    the same vertices, unit normals and the same voxel mask."""
    tris = W.ellipsoid_mesh(semi=(9.0, 5.0, 4.0), centre=(16.0, 8.0, 8.0), n_u=60, n_v=30, seed=3, jitter=0.05)
    path = tmp_path / "ellipsoid.stl"
    A.stl_write_binary(path, tris, header="aeros synthetic ellipsoid")
    assert path.stat().st_size == 84 + 50 * len(tris)
    v, n, dropped = A.stl_load(path)
    assert dropped == 0
    np.testing.assert_array_equal(v, tris)
    np.testing.assert_allclose(np.linalg.norm(n[:, :3], axis=1), 1.0, rtol=1e-5)
    grid = ((32, 16, 16), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    np.testing.assert_array_equal(A.voxelize(v, *grid), A.voxelize(tris, *grid))
