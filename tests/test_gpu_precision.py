"""GPU: the stated error of each storage precision against the FP64 path (which
equals the FP64 oracle within 1e-12) on the 96x48x48 sphere at Re 100 (BGK,
U0 0.05, free-slip side walls) after 2000 steps: drag coefficient (control
volume 0.45/0.40/0.40 of the domain) and the velocity field's relative L2.

Measured on MI355X: FP32 Cd -0.56 %, L2 2.4e-4; FP16 deviations Cd -12.6 %,
L2 4.6e-3; BF16 deviations Cd +9.7 %, L2 1.8e-2.  The tolerances below state
those errors with a margin; the 0.5 % drag target is met by no precision
below FP64 on this case.
"""
import numpy as np
import pytest

import aeros_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu

# precision: (|Cd error| bound, velocity L2 bound), both relative to FP64
TOL = {"fp32": (0.01, 5e-4), "fp16": (0.15, 1e-2), "bf16": (0.12, 3e-2), "fp32_fp64c": (1e-4, 1e-6),
       "fp16_eq": (0.03, 1.5e-3), "bf16_eq": (0.06, 6e-3)}


def _run(prec):
    L = A.Lattice(A.default_config(**W.c1_sphere_fp64(precision=prec)))
    area = L.projected_area(0)
    L.step(2000)
    cd = L.cv_force(0.45, 0.40, 0.40)[0] / (0.5 * 0.05 * 0.05 * area)
    u = np.stack([L.field(k).astype(np.float64) for k in ("ux", "uy", "uz")])
    L.close()
    return cd, u


def test_storage_precision_errors_against_fp64():
    print("precision errors: 96x48x48 sphere, 2000 steps", flush=True)
    cd64, u64 = _run("fp64")
    assert np.isfinite(cd64) and cd64 > 0
    for prec, (tol_cd, tol_l2) in TOL.items():
        cd, u = _run(prec)
        err = (cd - cd64) / cd64
        l2 = float(np.sqrt(((u - u64) ** 2).sum() / (u64 ** 2).sum()))
        print(f"{prec}: Cd {cd:.6f} against {cd64:.6f} ({err:+.3e}), velocity L2 {l2:.3e}", flush=True)
        assert abs(err) <= tol_cd, (prec, err)
        assert l2 <= tol_l2, (prec, l2)
