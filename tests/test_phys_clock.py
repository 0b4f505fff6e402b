"""Host stepping driver, host side: the fixed-tick physics clock and
updateLBM's step rule.

The clock's known answers are those the reference's own clock suite checks
(aeros/tests/phys_clock_headless.cpp), case by case; the step rule follows
aeros/src/lbm.cpp:1762-1776.  No GPU is used.
"""
import math

import pytest

import aeros_lbm as A


def clock(**kw):
    return A.PhysClock(**kw)


def test_defaults():                                   # phys_clock_headless.cpp:47-56
    c = clock()
    assert c.fixed == 1 and c.hz == 240.0 and c.max_substeps == 64 and c.particle_substeps == 8
    assert c.time_scale == 1.0


def test_sixty_fps_at_240_hz():                        # :58-71
    c = clock(hz=240.0)
    t, _ = c.advance(1.0 / 60.0, 0.0, 60.0)
    assert t == 4 and c.part_steps == 4
    assert abs(c.frame_sim_dt() - 1.0 / 60.0) <= 1e-4
    assert abs(c.particle_dt() - 1.0 / 240.0) <= 1e-5
    assert abs(c.display_ratio - 4.0) <= 1e-4


def test_2000_hz():                                    # :73-83
    c = clock(hz=2000.0)
    t, _ = c.advance(1.0 / 60.0, 0.0, 60.0)
    assert t == 33 and c.part_steps == c.particle_substeps
    assert abs(c.display_ratio - 2000.0 / 60.0) <= 1e-3


def test_tick_rate_independent_of_fps():               # :85-110
    def run(fps):
        c = clock(hz=240.0)
        dt = float(__import__("numpy").float32(1.0) / __import__("numpy").float32(fps))
        ticks = sum(c.advance(dt, 0.0, fps)[0] for _ in range(int(fps * 2.0)))
        return ticks, c.sim_time
    r30, r60, r144 = run(30.0), run(60.0), run(144.0)
    assert abs(r30[0] - r60[0]) <= 1 and abs(r60[0] - r144[0]) <= 1
    assert abs(r30[1] - 2.0) <= 0.01 and abs(r144[1] - 2.0) <= 0.01
    assert abs(r60[0] - 480) <= 2


def test_lattice_steps_follow_simulated_time():        # :112-140
    c = clock(hz=240.0)
    total = sum(c.advance(1.0 / 60.0, 0.005, 60.0)[1] for _ in range(600))
    assert total == 2000 and c.total_lattice_steps == total
    assert abs(c.sim_time - 10.0) <= 0.01 and 0.0 <= c.lattice_accum < 1.0
    c = clock(hz=60.0)
    assert c.advance(1.0 / 60.0, 0.005, 60.0)[1] == 3


def test_solver_off_and_nan_lattice_dt():              # :141-150
    c = clock(hz=240.0)
    t, lbm = c.advance(1.0 / 60.0, 0.0, 60.0)
    assert t == 4 and lbm == 0
    assert c.advance(1.0 / 60.0, float("nan"), 60.0)[1] == 0


def test_spiral_of_death_limit():                      # :152-168
    c = clock(hz=240.0, max_substeps=64, max_frame_dt=5.0)
    t, _ = c.advance(1.0, 0.0, 1.0)
    assert t == 64 and c.saturated and c.time_scale < 0.5
    assert abs(c.sim_time - 64.0 / 240.0) <= 1e-4
    assert 0.7 < c.accumulator < 0.75
    assert abs(c.sim_time + c.accumulator - c.real_time) <= 1e-3


def test_long_frame_is_cut():                          # :169-185
    c = clock(hz=240.0)
    c.advance(10.0, 0.0, 0.1)
    assert abs(c.real_time - 0.25) <= 1e-4 and c.saturated
    assert abs(c.dropped_time - 9.75) <= 1e-3 and abs(c.wall_scale - 0.025) <= 1e-3
    assert abs(c.time_scale - 1.0) <= 1e-4 and c.ticks == 60


def test_recovery_after_long_frame():                  # :186-203
    c = clock(hz=240.0)
    c.advance(2.0, 0.0, 0.5)
    for _ in range(600):
        c.advance(1.0 / 60.0, 0.0, 60.0)
    assert not c.saturated and abs(c.time_scale - 1.0) <= 1e-4
    assert 0.25 < c.wall_scale < 1.0 and abs(c.dropped_time - 1.75) <= 1e-3
    assert abs(c.wall_scale - 10.25 / 12.0) <= 1e-3


def test_garbage_inputs():                             # :205-237
    c = clock()
    t1 = c.advance(float("nan"), 0.0, 60.0)[0]
    t2 = c.advance(-5.0, 0.0, 60.0)[0]
    t3 = c.advance(float("inf"), 0.0, 60.0)[0]
    assert t1 == 0 and t2 == 0 and 0 <= t3 <= c.max_substeps
    assert math.isfinite(c.sim_time) and math.isfinite(c.time_scale) and math.isfinite(c.achieved_hz)
    c = clock()
    for hz in (0.0, -100.0, float("nan")):
        c.set(hz=hz)
        c.sanitize()
        assert c.hz >= 1.0
    c.set(hz=1e9)
    c.sanitize()
    assert c.hz <= 5000.0
    for v, want in ((0, 1), (1000000, 4096)):
        c.set(max_substeps=v)
        c.sanitize()
        assert c.max_substeps == want
    for v, want in ((-3, 1), (10000, 64)):
        c.set(particle_substeps=v)
        c.sanitize()
        assert c.particle_substeps == want
    c.set(accumulator=float("nan"), lattice_accum=-7.0)
    c.sanitize()
    assert c.accumulator == 0.0 and c.lattice_accum == 0.0


def test_per_frame_mode():                             # :239-251
    c = clock(fixed=0)
    t, lbm = c.advance(1.0 / 60.0, 0.005, 60.0)
    assert t == 1 and lbm == -1 and c.part_steps == 1
    assert abs(c.time_scale - 1.0) <= 1e-6 and c.particle_dt() == 0.0


def test_jittery_frames_do_not_drift():                # :253-274
    import numpy as np
    c = clock(hz=500.0)
    ticks, real = 0, 0.0
    for i in range(3000):
        dt = float(np.float32(0.0069) + np.float32(0.0261) * np.float32((i * 37) % 100) / np.float32(100.0))
        ticks += c.advance(dt, 0.002, 1.0 / dt)[0]
        real += dt
    assert abs(c.sim_time - real) <= 0.01 and abs(ticks - real * 500.0) <= 2.0
    assert abs(c.total_lattice_steps - int(real / 0.002)) <= 1
    assert 0.999 < c.time_scale < 1.001


def test_high_fps_skips_frames():                      # :276-290
    c = clock(hz=240.0)
    res = [c.advance(0.001, 0.0, 1000.0)[0] for _ in range(1000)]
    assert sum(res) == 240 and res.count(0) == 760
    assert abs(c.sim_time - 1.0) <= 0.01 and c.part_steps >= 0


def test_monotonic_and_reset():                        # :292-310
    c = clock(hz=240.0)
    prev = 0.0
    for _ in range(200):
        c.advance(1.0 / 60.0, 0.005, 60.0)
        assert c.sim_time >= prev
        prev = c.sim_time
    assert c.total_ticks > 0
    c.reset()
    assert (c.sim_time, c.real_time, c.accumulator, c.lattice_accum, c.total_ticks, c.total_lattice_steps) == (0, 0, 0, 0, 0, 0)
    assert c.hz == 240.0 and c.fixed


@pytest.mark.parametrize("dt,force,spf,want", [
    (0.016, 10, 3, 10), (0.016, 100, 3, 64), (0.016, 0, 3, 0),     # forced, clamped to 64
    (0.016, -1, 3, 3), (0.016, -1, 0, 1), (0.016, -1, 80, 50),     # stepsPerFrame in [1, 50]
    (0.03, -1, 3, 6), (0.03, -1, 40, 50),                          # slow frame doubles, capped
])
def test_update_step_rule(dt, force, spf, want):
    assert A.update_steps(dt, force, spf) == want
