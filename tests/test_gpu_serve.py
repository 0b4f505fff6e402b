"""GPU: served mode (include/aeros_serve.h) against the single-lattice solver.

A served case advances in batched launches beside other cases; its block
count depends on its own size only, and it counts its rounds from its own
admission, so it must compute exactly what it computes alone, whenever it
joins, and its drag must equal the single-lattice solver's drag on the same
fields.
"""
import time

import numpy as np
import pytest

import aeros_lbm as A
import workloads as W

pytestmark = pytest.mark.gpu


def small_cases(seed=11):
    """Five small requests of different sizes (free-slip side walls, Re from the
    stream's draw); the third and fifth are thermal."""
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
    ids = [srv.submit(A.default_config(**cfg, precision=kw["precision"])) for cfg, _ in cases]
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


def single_drag(cfg, precision, samples, every):
    """The single-lattice solver stepped `every` steps at a time: the drag
    coefficient after each call (lbmControlVolumeForce on the served control
    volume, over 1/2 U0^2 A)."""
    L = A.Lattice(A.default_config(**cfg, precision=precision))
    area = L.projected_area(0)
    u0 = float(np.float32(np.sqrt(np.float32(cfg["U0"]) * np.float32(cfg["U0"]))))
    q = 0.5 * u0 * u0 * float(area)
    fx = []
    for _ in range(samples):
        L.step(every)
        fx.append(L.cv_force(W.C5_CV_HALF, W.C5_CV_HALF, W.C5_CV_HALF)[0])
    L.close()
    return fx, q, area


@pytest.mark.parametrize("precision", ["fp16", "fp32"])
def test_served_alone_equals_mix(precision):
    cases = small_cases()
    mix, st = served(cases, precision=precision)
    assert st["peak_resident"] >= 2
    for k in (0, 2):   # a plain and a thermal case, each alone
        alone, _ = served([cases[k]], precision=precision)
        print(f"{precision} case {k}: alone cd {alone[0]['cd']:.9f}, in the mix {mix[k]['cd']:.9f}", flush=True)
        same(alone[0], mix[k])
    for r in mix:
        assert r["state_name"] == "done" and r["steps"] == 400 and r["rounds"] == 2
        assert np.isfinite(r["cd"])   # transient after 400 steps: the sign is not settled


def test_served_drag_equals_single_lattice():
    cfg, meta = small_cases()[1]
    (res,), _ = served([(cfg, meta)])
    fx, q, area = single_drag(cfg, "fp16", 2, 200)
    assert res["area"] == area
    assert np.float64(res["cd_prev"]).tobytes() == np.float64(fx[0] / q).tobytes(), (res["cd_prev"], fx[0] / q)
    assert np.float64(res["cd"]).tobytes() == np.float64(fx[1] / q).tobytes(), (res["cd"], fx[1] / q)


def test_served_round_averaged_drag():
    """cd_average: a round's Cd is the mean of the drag sampled after each tick."""
    cfg, meta = small_cases()[3]
    (res,), _ = served([(cfg, meta)], tick_steps=20, cd_average=1)
    fx, q, _ = single_drag(cfg, "fp16", 20, 20)
    means = []
    for r in range(2):
        s = 0.0
        for v in fx[r * 10:(r + 1) * 10]:
            s += v
        means.append((s / 10) / q)
    assert np.float64(res["cd_prev"]).tobytes() == np.float64(means[0]).tobytes(), (res["cd_prev"], means[0])
    assert np.float64(res["cd"]).tobytes() == np.float64(means[1]).tobytes(), (res["cd"], means[1])


def test_requests_join_mid_round():
    """A request submitted while another is part-way through its rounds joins at
    the next tick (it does not wait for a round boundary); both results equal
    their runs alone."""
    cases = small_cases()
    a, b = cases[3], cases[1]
    kw = dict(precision="fp16", round_steps=200, tick_steps=20, cd_tolerance=1e30, max_steps=4000,
              cv_half=W.C5_CV_HALF)
    srv = A.Server(**kw)
    ia = srv.submit(A.default_config(**a[0], precision="fp16"))
    t0 = time.time()
    while True:
        ra = srv.poll(ia)
        if ra["state_name"] == "running" and srv.stats()["ticks"] - ra["tick_admit"] >= 3:
            break
        assert time.time() - t0 < 60.0, "first request never started"
        time.sleep(0.0005)
    ib = srv.submit(A.default_config(**b[0], precision="fp16"))
    assert srv.wait_all(120.0)
    ra, rb = srv.poll(ia), srv.poll(ib)
    srv.close()
    joined = rb["tick_admit"] - ra["tick_admit"]
    print(f"second request joined {joined} ticks after the first (step {20 * joined % 200} of its round)",
          flush=True)
    assert joined >= 3 and rb["tick_admit"] < ra["tick_done"]
    (alone_a,), _ = served([a], **kw)
    (alone_b,), _ = served([b], **kw)
    same(ra, alone_a)
    same(rb, alone_b)


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
