"""GPU: free-slip walls against the oracle (whose rule the CPU suite pins by
the mirror-image and axis-exchange symmetries).  FP32 bit for bit, FP64 to
1e-12 relative, through the C ABI."""
import numpy as np
import pytest

import aeros_lbm as A
import workloads as W
from test_gpu_parity import assert_same, make_pair

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "ux", "uy", "uz", "tke", "strain", "ux_world", "uy_world", "uz_world", "pressure")


@pytest.mark.parametrize("walls,extra", [
    (W.SIDE_WALLS, {}),
    (A.FACE_Y0 | A.FACE_Z1, dict(use_les=1, use_trt=1)),
    (A.FACE_Z0 | A.FACE_Z1, dict(convective_outlet=1)),
])
def test_free_slip_fp32_bitwise(oracle, walls, extra):
    G, O = make_pair(oracle, **W.slip_box(40, 20, 18, (12.0, 9.0, 3.5), 6.0, walls, **extra))
    for n in (1, 47):
        G.step(n)
        O.step(n)
        assert_same(G, O, FIELDS)
    np.testing.assert_array_equal(G.populations(), O.populations())
    sg, so = G.scalars(), O.scalars()
    for k in ("avg_rho", "avg_ke", "max_u_lb", "tke", "neg_pops", "u_limited"):
        assert sg[k] == so[k], (k, sg[k], so[k])


def test_free_slip_periodic_channel_with_force_bitwise(oracle):
    """Guo forcing along x, periodic x and z, free-slip y walls: a frictionless
    channel accelerates uniformly."""
    cfg = W.channel(18, 0.7, 2e-6)
    cfg.update(voxels=np.zeros_like(cfg["voxels"]), free_slip=A.FACE_Y0 | A.FACE_Y1)
    G, O = make_pair(oracle, **cfg)
    G.step(60)
    O.step(60)
    u60 = float(G.field("ux", dtype=np.float64).mean())
    G.step(60)
    O.step(60)
    assert_same(G, O, ("rho", "ux", "uy", "uz"))
    np.testing.assert_array_equal(G.populations(), O.populations())
    ux = G.field("ux", dtype=np.float64)
    assert float(ux.max() - ux.min()) <= 1e-6 * float(ux.max())      # uniform: the walls exert no drag
    np.testing.assert_allclose(float(ux.mean()) - u60, 60 * 2e-6, rtol=1e-3)   # momentum grows by F a step


def test_c1_fp64_against_fp64_oracle(oracle):
    """The C1 case (96x48x48 FP64 BGK sphere, free-slip side walls): the
    FP64 rho sum over 120 steps and the fields within 1e-12."""
    cfg = W.c1_sphere_fp64()
    G = A.Lattice(A.default_config(**cfg))
    O = oracle.OracleLattice(A.default_config(**cfg))
    G.step(120)
    O.step(120)
    rg = G.rho_sum()
    ro, _ = O.checksums()
    assert abs(rg - ro) <= 1e-12 * abs(ro), (rg, ro)
    assert_same(G, O, names=("rho", "ux", "uy", "uz"), rtol=1e-12)


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_free_slip_16bit_runs_the_slip_kernel(oracle, prec):
    case = W.slip_box(40, 20, 18, (12.0, 9.0, 3.5), 6.0, W.SIDE_WALLS)
    G = A.Lattice(A.default_config(**case, precision=prec))
    O64 = oracle.OracleLattice(A.default_config(**case, precision="fp64"))
    G.step(60)
    O64.step(60)
    for name in ("rho", "ux", "uy", "uz"):
        err = float(np.abs(G.field(name, dtype=np.float64) - O64.field(name, dtype=np.float64)).max())
        assert err < (2e-3 if prec == "fp16" else 8e-3), (name, err)


@pytest.mark.parametrize("les", [0, 1])
def test_free_slip_with_the_thermal_lattice_fp32_bitwise(oracle, les):
    """Free-slip side walls on the D3Q19 lattice with the D3Q7 lattice beside it
    (hot and cold y walls, x and z wrapping for the temperature), as the served
    thermal requests run: fields, temperature and populations bit for bit."""
    rng = np.random.Generator(np.random.PCG64(21))
    cfg, _ = W.c5_request(rng, thermal_frac=1.0, size=0.0)
    cfg["use_les"] = les
    G, O = make_pair(oracle, **cfg)
    for n in (1, 39):
        G.step(n)
        O.step(n)
        assert_same(G, O, FIELDS + ("T", "buoyancy"))
    np.testing.assert_array_equal(G.populations(), O.populations())
    assert G.scalars()["nusselt"] == O.scalars()["nusselt"]
