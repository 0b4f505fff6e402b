// Reference-side binding: what a maintainer adds to aeros/src to run the LBM
// hot path on an MI355X through libaeros_lbm.so (include/aeros_lbm.h).
//
// It is compiled inside the reference tree (it includes the reference's
// lbm.h, thermal_lbm.h, globals.h and atmosphere.h) and linked with
// -laeros_lbm.  lbm.cpp calls these functions at the top of initLBM,
// resetLBM, shutdownLBM, stepLBMCPU, updateLBM, lbmControlVolumeForce and
// lbmProjectedAreaCells and falls back to its CPU code when one returns
// false, the pattern the reference already uses for its CUDA helpers
// (aeros/src/voxel_grid.cpp:559-571, aeros/src/forces.cpp:274-289).  After
// every step call the shim copies the device fields into the std::vector
// globals, so the samplers, forces, streamlines and the UI read them
// unchanged.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <glm/glm.hpp>

#include "atmosphere.h"
#include "globals.h"
#include "lbm.h"
#include "thermal_lbm.h"
#include "aeros_lbm.h"

namespace {

aeros_lbm* g_hip = nullptr;

// Storage precision, AEROS_LBM_PRECISION = fp32 (default) | fp64 | fp16 | bf16 |
// fp32_fp64c | fp16_eq | bf16_eq.
int precision_from_env() {
    const char* p = std::getenv("AEROS_LBM_PRECISION");
    if (!p) return AEROS_FP32;
    if (!std::strcmp(p, "fp64")) return AEROS_FP64;
    if (!std::strcmp(p, "fp16")) return AEROS_FP16;
    if (!std::strcmp(p, "bf16")) return AEROS_BF16;
    if (!std::strcmp(p, "fp32_fp64c")) return AEROS_FP32_FP64C;
    if (!std::strcmp(p, "fp16_eq")) return AEROS_FP16_EQ;
    if (!std::strcmp(p, "bf16_eq")) return AEROS_BF16_EQ;
    return AEROS_FP32;
}

void report(const char* where) {
    char msg[512] = {0};
    const int code = aeros_lbm_last_error(msg, (int32_t)sizeof msg);
    std::fprintf(stderr, "[aeros-hip] %s failed (%d): %s\n", where, code, msg);
}

// initLBM's inputs (aeros/src/lbm.cpp:405-560) into the ABI's config.
void fill_config(aeros_lbm_config& c) {
    aeros_lbm_config_defaults(&c);
    c.voxels = g_voxelData.empty() ? nullptr : g_voxelData.data();
    c.vox_n[0] = g_voxNx; c.vox_n[1] = g_voxNy; c.vox_n[2] = g_voxNz;
    c.vox_min[0] = g_voxMinX; c.vox_min[1] = g_voxMinY; c.vox_min[2] = g_voxMinZ;
    c.vox_max[0] = g_voxMaxX; c.vox_max[1] = g_voxMaxY; c.vox_max[2] = g_voxMaxZ;
    c.voxel_resolution = voxelResolution;
    c.flow_speed = flowSpeed;
    c.flow_azimuth_deg = flowAzimuth;
    c.flow_elevation_deg = flowElevation;
    c.air_density = airDensity;
    c.air_temperature = airTemperature;
    c.max_dim = maxDim;
    c.frontal_area = computeLBMRefArea();
    c.ground_effect = aeroGroundEffect ? 1 : 0;
    c.ground_effect_height = aeroGroundHeight;
    c.tau = lbmParams.tau;
    c.U0 = lbmParams.U0;
    c.smagorinsky_c = lbmParams.smagorinskyC;
    c.convergence_threshold = lbmParams.convergenceThreshold;
    c.domain_padding = lbmParams.domainPadding;
    c.inlet_turbulence = lbmParams.inletTurbulence;
    c.ground_height = lbmParams.groundHeight;
    c.body_force[0] = lbmParams.bodyForceX;
    c.body_force[1] = lbmParams.bodyForceY;
    c.body_force[2] = lbmParams.bodyForceZ;
    c.use_les = lbmParams.useTurbulence ? 1 : 0;
    c.use_trt = lbmParams.useTRT ? 1 : 0;
    c.convective_outlet = lbmParams.useConvectiveOutlet ? 1 : 0;
    c.periodic_x = lbmParams.periodicX ? 1 : 0;
    c.periodic_z = lbmParams.periodicZ ? 1 : 0;
    c.auto_viscosity = lbmParams.autoViscosity ? 1 : 0;
    c.use_ground = lbmParams.useGround ? 1 : 0;
    c.thermal = thermalParams.enabled ? 1 : 0;
    c.rayleigh = thermalParams.rayleigh;
    c.prandtl = thermalParams.prandtl;
    c.fixed_wall_temps = thermalParams.fixedWallTemps ? 1 : 0;
    c.hot_temp = thermalParams.hotTemp;
    c.cold_temp = thermalParams.coldTemp;
    c.perturbation = thermalParams.perturbation;
    c.precision = precision_from_env();
}

// The scalar globals of aeros/src/lbm.h:77-125 and aeros/src/thermal_lbm.h:46-58.
void apply_scalars(const aeros_lbm_scalars& s) {
    lbmCurrentStep = s.step;
    lbmConverged = s.converged != 0;
    lbmAvgRho = s.avg_rho;
    lbmAvgKineticEnergy = s.avg_ke;
    lbmMaxVelocityLB = s.max_u_lb;
    lbmMaxVelocityWorld = s.max_u_world;
    lbmTKE = s.tke;
    lbmConvergence = s.convergence;
    lbmNegativePops = s.neg_pops;
    lbmBadPops = s.bad_pops;
    lbmRhoOutOfRange = s.rho_out_of_range;
    lbmRhoNonFinite = s.rho_non_finite;
    lbmULimited = s.u_limited;
    lbmPressureNonFinite = s.pressure_non_finite;
    lbmParams.tau = s.tau;
    lbmParams.viscosity = s.viscosity;
    lbmParams.U0 = s.U0;
    lbmReynolds = s.reynolds_phys;
    lbmReynoldsLB = s.reynolds_lb;
    lbmBlockageRatio = s.blockage;
    lbmPressureMin = s.pressure_min;
    lbmPressureMax = s.pressure_max;
    lbmResidualRho = s.residual_rho;
    lbmResidualU = s.residual_u;
    lbmTimeMs = s.time_ms;
    thermalReady = s.thermal_ready != 0;
    thermalNusselt = s.nusselt;
    thermalMeanT = s.thermal_mean_t;
    thermalAlpha = s.thermal_alpha;
    thermalGravityLb = s.thermal_gravity;
    thermalRayleigh = s.thermal_rayleigh;
    thermalPrandtl = s.thermal_prandtl;
}

void pull_scalars() {
    aeros_lbm_scalars s;
    if (aeros_lbm_get_scalars(g_hip, &s) == AEROS_OK) apply_scalars(s);
}

void pull_field(int32_t id, std::vector<float>& dst, size_t n) {
    dst.resize(n);
    if (aeros_lbm_read_field(g_hip, id, dst.data()) != AEROS_OK) dst.assign(n, 0.0f);
}

// The field globals the samplers and the visualizer read (aeros/src/lbm.h:132-154).
void pull_fields() {
    const size_t n = (size_t)lbmNx * (size_t)lbmNy * (size_t)lbmNz;
    pull_field(AEROS_FIELD_RHO, lbmRho, n);
    pull_field(AEROS_FIELD_UX, lbmUx, n);
    pull_field(AEROS_FIELD_UY, lbmUy, n);
    pull_field(AEROS_FIELD_UZ, lbmUz, n);
    pull_field(AEROS_FIELD_UX_WORLD, lbmUxWorld, n);
    pull_field(AEROS_FIELD_UY_WORLD, lbmUyWorld, n);
    pull_field(AEROS_FIELD_UZ_WORLD, lbmUzWorld, n);
    pull_field(AEROS_FIELD_PRESSURE, lbmPressure, n);
    pull_field(AEROS_FIELD_TKE, lbmTKEField, n);
    pull_field(AEROS_FIELD_STRAIN, lbmStrainMag, n);
    if (thermalReady) {
        pull_field(AEROS_FIELD_T, thermalT, n);
        pull_field(AEROS_FIELD_BUOYANCY, thermalBuoyancyY, n);
    }
}

void pull_geometry() {
    int32_t d[3];
    float mn[3], cs[3];
    aeros_lbm_dims(g_hip, d, mn, cs);
    lbmNx = d[0]; lbmNy = d[1]; lbmNz = d[2];
    lbmMinX = mn[0]; lbmMinY = mn[1]; lbmMinZ = mn[2];
    lbmCellSizeX = cs[0]; lbmCellSizeY = cs[1]; lbmCellSizeZ = cs[2];
    lbmMaxX = mn[0] + cs[0] * (float)d[0];
    lbmMaxY = mn[1] + cs[1] * (float)d[1];
    lbmMaxZ = mn[2] + cs[2] * (float)d[2];
    const size_t n = (size_t)d[0] * (size_t)d[1] * (size_t)d[2];
    std::vector<float> tmp(n);
    lbmIsSolid.assign(n, 0);
    lbmIsGround.assign(n, 0);
    if (aeros_lbm_read_field(g_hip, AEROS_FIELD_SOLID, tmp.data()) == AEROS_OK)
        for (size_t i = 0; i < n; ++i) lbmIsSolid[i] = tmp[i] != 0.0f;
    if (aeros_lbm_read_field(g_hip, AEROS_FIELD_GROUND, tmp.data()) == AEROS_OK)
        for (size_t i = 0; i < n; ++i) lbmIsGround[i] = tmp[i] != 0.0f;
}

// Asynchronous frames: whether they run, and the newest frame in the globals.
bool g_frames = false;
int64_t g_frame_seen = 0;

// The fields pull_fields copies, as they arrive with a finished frame.
struct FrameField {
    int32_t id;
    std::vector<float>* dst;
};
std::vector<FrameField> frame_fields() {
    std::vector<FrameField> v = {{AEROS_FIELD_RHO, &lbmRho}, {AEROS_FIELD_UX, &lbmUx}, {AEROS_FIELD_UY, &lbmUy},
                                 {AEROS_FIELD_UZ, &lbmUz}, {AEROS_FIELD_UX_WORLD, &lbmUxWorld},
                                 {AEROS_FIELD_UY_WORLD, &lbmUyWorld}, {AEROS_FIELD_UZ_WORLD, &lbmUzWorld},
                                 {AEROS_FIELD_PRESSURE, &lbmPressure}, {AEROS_FIELD_TKE, &lbmTKEField},
                                 {AEROS_FIELD_STRAIN, &lbmStrainMag}, {AEROS_FIELD_VORTICITY, &lbmVorticityMag},
                                 {AEROS_FIELD_QCRIT, &lbmQCriterion}};
    if (thermalReady) {
        v.push_back({AEROS_FIELD_T, &thermalT});
        v.push_back({AEROS_FIELD_BUOYANCY, &thermalBuoyancyY});
    }
    return v;
}

void stop_frames() {
    if (g_hip && g_frames) aeros_lbm_frames_stop(g_hip);
    g_frames = false;
    g_frame_seen = 0;
}

}  // namespace

// initLBM (aeros/src/lbm.cpp:405-886) on the device.
bool aerosHipInitLBM() {
    stop_frames();
    if (g_hip) { aeros_lbm_destroy(g_hip); g_hip = nullptr; }
    aeros_lbm_config c;
    fill_config(c);
    int err = 0;
    g_hip = aeros_lbm_create(&c, &err);
    if (!g_hip) { report("aeros_lbm_create"); lbmInitialized = false; return false; }
    pull_geometry();
    pull_scalars();
    pull_fields();
    lbmInitialized = true;
    return true;
}

// resetLBM (aeros/src/lbm.cpp:888-895).
bool aerosHipResetLBM() {
    if (!g_hip) return aerosHipInitLBM();
    stop_frames();
    if (aeros_lbm_reset(g_hip) != AEROS_OK) { report("aeros_lbm_reset"); return false; }
    pull_scalars();
    pull_fields();
    return true;
}

// shutdownLBM (aeros/src/lbm.cpp:897-905).
void aerosHipShutdownLBM() {
    stop_frames();
    if (g_hip) aeros_lbm_destroy(g_hip);
    g_hip = nullptr;
    lbmInitialized = false;
}

// stepLBMCPU(steps) (aeros/src/lbm.cpp:907-1709).
bool aerosHipStepLBM(int steps) {
    if (!g_hip) return false;
    if (aeros_lbm_step(g_hip, steps) != AEROS_OK) { report("aeros_lbm_step"); return false; }
    pull_scalars();
    pull_fields();
    return true;
}

// updateLBM(dt, forceSteps) (aeros/src/lbm.cpp:1734-1779).
bool aerosHipUpdateLBM(float dt, int forceSteps) {
    if (!g_hip) return false;
    if (aeros_lbm_update(g_hip, dt, forceSteps, lbmParams.stepsPerFrame) != AEROS_OK) {
        report("aeros_lbm_update");
        return false;
    }
    pull_scalars();
    pull_fields();
    return true;
}

// updateLBM without blocking the frame loop: the frame is queued on the
// library's frame worker (the same steps and 30-step vorticity pass), and the
// globals take the newest finished frame, one or two frames behind, from
// pinned host memory.  Any synchronous call (stepLBMCPU, a reset) still
// works and waits for the frame in progress.
bool aerosHipUpdateLBMAsync(float dt, int forceSteps) {
    if (!g_hip) return false;
    if (!g_frames) {
        uint32_t mask = 0;
        for (const FrameField& f : frame_fields()) mask |= 1u << f.id;
        if (aeros_lbm_frames_start(g_hip, mask) != AEROS_OK) { report("aeros_lbm_frames_start"); return false; }
        g_frames = true;
    }
    if (aeros_lbm_frame_submit(g_hip, dt, forceSteps, lbmParams.stepsPerFrame) != AEROS_OK) {
        report("aeros_lbm_frame_submit");
        return false;
    }
    aeros_frame fr;
    if (aeros_lbm_frame_latest(g_hip, &fr) != AEROS_OK) { report("aeros_lbm_frame_latest"); return false; }
    if (fr.frame > g_frame_seen) {
        const size_t n = (size_t)lbmNx * (size_t)lbmNy * (size_t)lbmNz;
        for (const FrameField& f : frame_fields())
            if (fr.field[f.id]) f.dst->assign(fr.field[f.id], fr.field[f.id] + n);
        apply_scalars(fr.scalars);
        g_frame_seen = fr.frame;
    }
    return true;
}

// lbmLatticeDtWorld (aeros/src/lbm.cpp:1717-1732).
bool aerosHipLatticeDtWorld(float* out) {
    return g_hip && aeros_lbm_lattice_dt(g_hip, out) == AEROS_OK;
}

// lbmControlVolumeForce (aeros/src/lbm.cpp:2183-2232): the double sums, cast like the reference's return.
bool aerosHipControlVolumeForce(float hx, float hy, float hz, glm::vec3* out) {
    double f[3];
    if (!g_hip || aeros_lbm_cv_force(g_hip, hx, hy, hz, f) != AEROS_OK) return false;
    *out = glm::vec3((float)f[0], (float)f[1], (float)f[2]);
    return true;
}

// lbmProjectedAreaCells (aeros/src/lbm.cpp:2236-2256).
bool aerosHipProjectedAreaCells(int axis, float* out) {
    return g_hip && aeros_lbm_projected_area(g_hip, axis, out) == AEROS_OK;
}

// computeLBMVorticityAndQ (aeros/src/lbm.cpp:2005-2093) on the device, then
// the three fields it writes into the reference's globals.
bool aerosHipComputeVorticityAndQ() {
    if (!g_hip || aeros_lbm_vorticity_q(g_hip) != AEROS_OK) { report("aeros_lbm_vorticity_q"); return false; }
    const size_t n = (size_t)lbmNx * (size_t)lbmNy * (size_t)lbmNz;
    pull_field(AEROS_FIELD_VORTICITY, lbmVorticityMag, n);
    pull_field(AEROS_FIELD_QCRIT, lbmQCriterion, n);
    pull_field(AEROS_FIELD_STRAIN, lbmStrainMag, n);
    return true;
}

// Batched forms of the per-point samplers getLBMPressureWorld /
// getLBMDensityWorld / getLBMVorticityWorld / getLBMQWorld / getLBMTKEWorld /
// getLBMStrainWorld and getLBMVelocityWorld (aeros/src/lbm.cpp:1845-1999),
// for callers that sample many points per frame (streamlines, maps):
// positions as glm::vec3, results evaluated on the device.
static_assert(sizeof(glm::vec3) == 3 * sizeof(float), "glm::vec3 must be three packed floats");
bool aerosHipSampleField(int32_t field, const std::vector<glm::vec3>& pos, std::vector<float>& out) {
    out.assign(pos.size(), 0.0f);
    if (!g_hip) return false;
    return aeros_lbm_sample(g_hip, field, reinterpret_cast<const float*>(pos.data()), (int64_t)pos.size(),
                            out.data(), nullptr) == AEROS_OK;
}
bool aerosHipSampleVelocityWorld(const std::vector<glm::vec3>& pos, std::vector<glm::vec3>& out) {
    out.assign(pos.size(), glm::vec3(0.0f));
    if (!g_hip) return false;
    return aeros_lbm_sample_velocity(g_hip, 1, reinterpret_cast<const float*>(pos.data()), (int64_t)pos.size(),
                                     reinterpret_cast<float*>(out.data())) == AEROS_OK;
}
