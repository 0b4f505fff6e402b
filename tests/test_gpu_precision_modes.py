"""GPU: the slow decay of the reference's 0.08 start flow between walls 30 rows
apart (see test_precision_modes.py).  The FP32 kernel, bit-identical to the
reference's arithmetic, freezes it near 2e-5; the FP32-storage mode with FP64
collision decays at the closed-form rate nu pi^2 / H^2, as the FP64 kernel does."""
import math

import numpy as np
import pytest

import aeros_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu

NY, TAU, STEPS = 32, 0.6, 40000


def decay(precision):
    L = A.Lattice(A.default_config(**W.channel(NY, TAU, 0.0), precision=precision))
    out = []
    for _ in range(8):
        L.step(STEPS // 8)
        out.append(float(np.abs(L.field("ux", dtype=np.float64)).max()))
    L.close()
    return out


def test_slow_decay_by_precision(oracle):
    print("channel decay 8x32x8, 40k steps per precision", flush=True)
    runs = {p: decay(p) for p in ("fp32", "fp64", "fp32_fp64c")}
    lam = (TAU - 0.5) / 3.0 * math.pi ** 2 / (NY - 2) ** 2
    for p in ("fp64", "fp32_fp64c"):
        u = runs[p]
        rate = math.log(u[3] / u[7]) / (4 * STEPS / 8)
        assert abs(rate - lam) <= 0.01 * lam, (p, rate, lam, u)
    u = runs["fp32"]
    assert u[-1] / u[-3] > 0.9 and u[-1] > 1e-5, u
    # the FP32 kernel is the oracle's FP32 arithmetic bit for bit
    O = oracle.OracleLattice(oracle.default_config(**W.channel(NY, TAU, 0.0), precision="fp32"))
    O.step(STEPS)
    assert float(np.abs(O.field("ux", dtype=np.float64)).max()) == u[-1]
