// Fixed-tick physics clock of the host stepping driver: how many physics
// ticks, and how many lattice steps, a rendered frame of a given length
// advances.  Follows PhysClock (aeros/src/phys_clock.h:48-250): a frame's
// real time accumulates, whole ticks of 1/hz are taken from it (at most
// max_substeps, frames longer than max_frame_dt are cut), and lattice steps
// follow the simulated time at lattice_dt seconds per step with the
// fractional remainder carried to the next frame.
#include <cmath>

#include "lbm_common.h"

namespace {

float finite_or(float v, float fallback) { return std::isfinite(v) ? v : fallback; }

void clamp_settings(aeros_phys_clock& c) {   // PhysClock::sanitize (phys_clock.h:92-105)
    c.hz = finite_or(c.hz, 1.0f);
    c.hz = std::fmin(std::fmax(c.hz, 1.0f), 5000.0f);
    c.max_substeps = c.max_substeps < 1 ? 1 : (c.max_substeps > 4096 ? 4096 : c.max_substeps);
    c.particle_substeps = c.particle_substeps < 1 ? 1 : (c.particle_substeps > 64 ? 64 : c.particle_substeps);
    if (!std::isfinite(c.max_frame_dt) || c.max_frame_dt < 1e-4f) c.max_frame_dt = 1e-4f;
    if (c.max_frame_dt > 5.0f) c.max_frame_dt = 5.0f;
    if (!std::isfinite(c.accumulator) || c.accumulator < 0.0f) c.accumulator = 0.0f;
    if (!std::isfinite(c.lattice_accum) || c.lattice_accum < 0.0f) c.lattice_accum = 0.0f;
}

float ratio_or_one(double num, double den) { return den > 1e-9 ? (float)(num / den) : 1.0f; }

}  // namespace

extern "C" {

void aeros_clock_defaults(aeros_phys_clock* c) {
    if (!c) return;
    *c = aeros_phys_clock{};
    c->fixed = 1;
    c->hz = 240.0f;
    c->max_substeps = 64;
    c->particle_substeps = 8;
    c->max_frame_dt = 0.25f;
    c->tick_dt = 1.0f / 240.0f;
    c->time_scale = c->wall_scale = c->display_ratio = 1.0f;
}

void aeros_clock_reset(aeros_phys_clock* c) {   // state only; the settings stay
    if (!c) return;
    c->accumulator = c->lattice_accum = 0.0f;
    c->sim_time = c->real_time = c->dropped_time = 0.0;
    c->total_ticks = c->total_lattice_steps = 0;
    c->ticks = c->lattice_steps = c->part_steps = 0;
    c->saturated = 0;
}

void aeros_clock_sanitize(aeros_phys_clock* c) {
    if (c) clamp_settings(*c);
}

int32_t aeros_clock_advance(aeros_phys_clock* c, float frame_dt, float lattice_dt, float fps, int32_t* lbm_steps) {
    if (!c) return 0;
    clamp_settings(*c);
    if (!std::isfinite(frame_dt) || frame_dt < 0.0f) frame_dt = 0.0f;
    const double true_dt = frame_dt;           // before the cut: what is dropped is reported
    const bool cut = frame_dt > c->max_frame_dt;
    if (cut) frame_dt = c->max_frame_dt;
    c->tick_dt = 1.0f / c->hz;

    if (!c->fixed) {   // one tick per frame; the solver picks its own steps per frame
        c->ticks = 1;
        c->part_steps = 1;
        c->lattice_steps = 0;
        if (lbm_steps) *lbm_steps = -1;
        c->sim_time += frame_dt;
        c->real_time += frame_dt;
        if (cut) c->dropped_time += true_dt - (double)frame_dt;
        c->total_ticks += 1;
        c->time_scale = 1.0f;
        c->wall_scale = ratio_or_one(c->sim_time, c->real_time + c->dropped_time);
        c->achieved_hz = fps > 0.5f ? fps : 0.0f;
        c->display_ratio = 1.0f;
        c->saturated = cut;
        return c->ticks;
    }

    c->accumulator += frame_dt;
    c->real_time += frame_dt;
    if (cut) c->dropped_time += true_dt - (double)frame_dt;
    // whole ticks by multiplying with hz and a 1e-4-tick allowance, so that a
    // frame of exactly k ticks never yields k - 1 from rounding
    int n = (int)((double)c->accumulator * (double)c->hz + 1e-4);
    if (n < 0) n = 0;
    c->saturated = cut;
    if (n > c->max_substeps) {
        n = c->max_substeps;
        c->saturated = 1;
    }
    c->accumulator -= (float)n * c->tick_dt;
    if (c->accumulator < 0.0f) c->accumulator = 0.0f;
    if (c->accumulator > c->max_frame_dt) {   // a backlog beyond one long frame is dropped
        c->dropped_time += (double)c->accumulator - (double)c->max_frame_dt;
        c->accumulator = c->max_frame_dt;
    }
    const double sim_dt = (double)n * (double)c->tick_dt;
    c->sim_time += sim_dt;
    c->ticks = n;
    c->total_ticks += n;

    c->lattice_steps = 0;
    if (std::isfinite(lattice_dt) && lattice_dt > 1e-9f) {
        c->lattice_accum += (float)(sim_dt / (double)lattice_dt);
        int k = (int)c->lattice_accum;
        k = k < 0 ? 0 : (k > 4096 ? 4096 : k);
        c->lattice_steps = k;
        c->lattice_accum -= (float)k;
        if (c->lattice_accum < 0.0f) c->lattice_accum = 0.0f;
        c->total_lattice_steps += k;
    } else {
        c->lattice_accum = 0.0f;
    }
    if (lbm_steps) *lbm_steps = c->lattice_steps;

    c->part_steps = n < c->particle_substeps ? n : c->particle_substeps;
    c->time_scale = ratio_or_one(c->sim_time, c->real_time);
    c->wall_scale = ratio_or_one(c->sim_time, c->real_time + c->dropped_time);
    c->achieved_hz = frame_dt > 1e-9f ? (float)((double)n / (double)frame_dt) : 0.0f;
    c->display_ratio = fps > 0.5f ? c->hz / fps : 0.0f;
    return n;
}

float aeros_clock_particle_dt(const aeros_phys_clock* c) {   // one particle sub-step of this frame
    if (!c || !c->fixed || c->part_steps <= 0) return 0.0f;
    float dt = (float)((double)c->ticks * (double)c->tick_dt) / (float)c->part_steps;
    if (!std::isfinite(dt) || dt < 0.0f) return 0.0f;
    return dt > c->max_frame_dt ? c->max_frame_dt : dt;
}

float aeros_clock_frame_sim_dt(const aeros_phys_clock* c) {   // simulated seconds of this frame
    if (!c || !c->fixed) return 0.0f;
    const float dt = (float)((double)c->ticks * (double)c->tick_dt);
    return (!std::isfinite(dt) || dt < 0.0f) ? 0.0f : dt;
}

}  // extern "C"
