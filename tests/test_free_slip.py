"""Free-slip (specular) walls: no counterpart in the reference, so the oracle's
rule is pinned by symmetry against the reference's own periodic wrap.

A free-slip wall half a cell outside a face is a mirror: the flow in a box
with free-slip walls on z = 0 and z = Nz-1 is exactly the flow in a box of
twice the depth, periodic in z, holding the geometry and its mirror image.
The y walls follow from the z walls by exchanging the y and z axes, and the
edges where a y wall meets a z wall from the same mirror construction with
y walls on both sides.  All comparisons are FP64; they differ only in the
order in which the moments add populations.
"""
import numpy as np
import pytest

import aeros_lbm as A
import workloads as W

STEPS = 80
RTOL = 1e-12


def fields(L, names=("rho", "ux", "uy", "uz")):
    return {k: L.field(k, dtype=np.float64) for k in names}


def mirrored_z(cfg):
    """The same case on 2 Nz cells, periodic in z, with the voxels mirrored below."""
    v = cfg["voxels"]
    nz, ny, nx = v.shape
    big = np.concatenate([v[::-1], v], axis=0)
    out = dict(cfg, voxels=big, **W.lattice_box(nx, ny, 2 * nz), periodic_z=1,
               free_slip=cfg["free_slip"] & ~(A.FACE_Z0 | A.FACE_Z1))
    return out


def run(oracle, cfg, steps=STEPS):
    L = oracle.OracleLattice(A.default_config(**cfg, precision="fp64"))
    L.step(steps)
    return L


def close(a, b, what):
    scale = max(1.0, float(np.abs(b).max()))
    err = float(np.abs(a - b).max())
    assert err <= RTOL * scale, (what, err)
    return err


@pytest.mark.parametrize("slip_y", [False, True])
def test_z_walls_are_the_mirror_of_a_periodic_box(oracle, slip_y):
    walls = (A.FACE_Z0 | A.FACE_Z1) | ((A.FACE_Y0 | A.FACE_Y1) if slip_y else 0)
    cfg = W.slip_box(28, 12, 10, (9.0, 5.0, 2.8), 4.0, walls)   # sphere a cell off the z = 0 wall
    a = fields(run(oracle, cfg))
    b = fields(run(oracle, mirrored_z(cfg)))
    nz = 10
    for k in ("rho", "ux", "uy"):
        close(a[k], b[k][nz:], k)
        close(a[k], b[k][nz - 1::-1], k + " (mirror half)")
    close(a["uz"], b["uz"][nz:], "uz")
    close(a["uz"], -b["uz"][nz - 1::-1], "uz (mirror half)")
    assert float(np.abs(a["uz"]).max()) > 1e-4      # the walls are felt
    # the walls differ from the reference's copy-in-place faces
    c = fields(run(oracle, dict(cfg, free_slip=0)))
    assert float(np.abs(c["ux"] - a["ux"]).max()) > 1e-5


def test_y_walls_are_the_z_walls_transposed(oracle):
    cz = W.slip_box(28, 12, 10, (9.0, 5.0, 2.8), 4.0, A.FACE_Z0 | A.FACE_Z1)
    vz = cz["voxels"]                                  # (z, y, x)
    cy = dict(cz, voxels=np.ascontiguousarray(vz.transpose(1, 0, 2)), **W.lattice_box(28, 10, 12),
              free_slip=A.FACE_Y0 | A.FACE_Y1)
    fz = fields(run(oracle, cz))
    fy = fields(run(oracle, cy))
    for k in ("rho", "ux"):
        close(fy[k], fz[k].transpose(1, 0, 2), k)
    close(fy["uy"], fz["uz"].transpose(1, 0, 2), "uy")
    close(fy["uz"], fz["uy"].transpose(1, 0, 2), "uz")


def test_uniform_stream_passes_free_slip_walls_unchanged(oracle):
    """No obstacle: the free-slip box keeps the inflow's uniform state (to rounding)."""
    cfg = W.slip_box(16, 8, 8, (-50.0, -50.0, -50.0), 1.0, W.SIDE_WALLS)
    L = run(oracle, cfg, 40)
    f = fields(L)
    u0 = float(np.float32(0.06))
    assert float(np.abs(f["ux"] - u0).max()) < 1e-12
    assert float(np.abs(f["uy"]).max()) < 1e-14 and float(np.abs(f["uz"]).max()) < 1e-14
    assert float(np.abs(f["rho"] - 1.0).max()) < 1e-12


@pytest.mark.parametrize("over,why", [
    (dict(free_slip=64), "unknown face"),
    (dict(free_slip=A.FACE_X0, periodic_x=1), "face of a periodic axis"),
    (dict(free_slip=A.FACE_Z1, periodic_z=1), "face of a periodic axis"),
])
def test_invalid_free_slip_is_rejected_before_the_device(over, why):
    with pytest.raises(A.AerosError) as e:
        A.Lattice(A.default_config(**over))
    msg = str(e.value)
    assert "bad argument" in msg and why in msg, msg
