"""GPU: served mode (include/aeros_serve.h) against the single-lattice solver.

A served case advances in batched launches beside other cases; its block
count depends on its own size only, so it must compute exactly what it
computes alone, and its drag must equal the single-lattice solver's drag on
the same fields.
"""
import numpy as np
import pytest

import aeros_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu


def small_cases(seed=11):
    """Five small requests of different sizes; the third and fifth are thermal."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for k, size in enumerate((0.05, 0.12, 0.08, 0.2, 0.1)):
        cfg, meta = W.c5_request(rng, thermal_frac=1.0 if k in (2, 4) else 0.0, size=size)
        out.append((cfg, meta))
    return out


def served(cases, **over):
    kw = dict(precision="fp16", round_steps=200, cd_tolerance=1e30, max_steps=400, cv_half=W.C5_CV_HALF)
    kw.update(over)
    srv = A.Server(**kw)
    ids = [srv.submit(A.default_config(**cfg, precision="fp16")) for cfg, _ in cases]
    assert srv.wait_all(120.0)
    res = [srv.poll(i) for i in ids]
    st = srv.stats()
    srv.close()
    return res, st


def same(a, b):
    for k in ("state", "steps", "rounds", "resets", "nx", "ny", "nz"):
        assert a[k] == b[k], (k, a[k], b[k])
    for k in ("cd", "cd_prev", "area", "nusselt"):
        assert np.float64(a[k]).tobytes() == np.float64(b[k]).tobytes(), (k, a[k], b[k])


def test_served_alone_equals_mix():
    cases = small_cases()
    mix, st = served(cases)
    assert st["peak_resident"] >= 2
    for k in (0, 2):   # a plain and a thermal case, each alone
        alone, _ = served([cases[k]])
        print(f"case {k}: alone cd {alone[0]['cd']:.9f}, in the mix {mix[k]['cd']:.9f}", flush=True)
        same(alone[0], mix[k])
    for r in mix:
        assert r["state_name"] == "done" and r["steps"] == 400 and r["rounds"] == 2
        assert np.isfinite(r["cd"])   # transient after 400 steps: the sign is not settled


def test_served_drag_equals_single_lattice():
    cfg, meta = small_cases()[1]
    (res,), _ = served([(cfg, meta)])
    L = A.Lattice(A.default_config(**cfg, precision="fp16"))
    area = L.projected_area(0)
    u0 = float(np.float32(np.sqrt(np.float32(0.05) * np.float32(0.05))))
    cds = []
    for _ in range(2):
        L.step(200)
        fx = L.cv_force(W.C5_CV_HALF, W.C5_CV_HALF, W.C5_CV_HALF)[0]
        cds.append(fx / (0.5 * u0 * u0 * float(area)))
    L.close()
    assert res["area"] == area
    assert np.float64(res["cd_prev"]).tobytes() == np.float64(cds[0]).tobytes(), (res["cd_prev"], cds[0])
    assert np.float64(res["cd"]).tobytes() == np.float64(cds[1]).tobytes(), (res["cd"], cds[1])


def test_served_memory_admission():
    cases = small_cases()
    # a budget that holds the largest case but not two of them at once
    big = max(m["cells"] for _, m in cases)
    budget = int(big * 200 + (6 << 20))
    res, st = served(cases, memory_budget=budget)
    assert all(r["state_name"] == "done" for r in res)
    assert st["peak_bytes"] <= budget
    assert st["peak_resident"] <= 3
    # a case that does not fit the budget on its own fails instead of waiting forever
    huge_cfg, _ = W.c5_request(np.random.Generator(np.random.PCG64(3)), thermal_frac=0.0, size=1.0)
    res2, _ = served([(huge_cfg, None)], memory_budget=budget // 8)
    assert res2[0]["state_name"] == "failed" and res2[0]["error"] == -5
