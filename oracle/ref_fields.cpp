// TEST INFRASTRUCTURE ONLY.  Runs the reference solver (compiled from its
// own sources under the reference tree by oracle/ref.mk) on small cases and
// writes its per-cell fields as raw little-endian float32 files into
// oracle/_ref/fields/, from which tests/golden/make_golden.py records SHA-256
// digests: the per-cell pin of the oracle's FP32 arithmetic (convective
// outlet, LES with Van Driest damping near walls, thermal coupling,
// Dirichlet wall rows) that the summed checksums only pin in aggregate.
//
// The link-time stubs the reference's solver files need (its GLFW clock, the
// CUDA and renderer hooks) come from the reference's own bench translation
// unit, built with its main renamed; this file adds none.
//
//   oracle/_ref/ref_fields <case> <out-dir>
//   cases: bench12_les  bench12_les_thermal  bench16  rb_walls  rb_storm
//          bench12_auto  bench12_auto_clamped (tau from Re, autoViscosity on;
//          the standard atmosphere at 1500 m; the first at a flow speed whose
//          Reynolds number puts tau inside [0.51, 1.5], the second at the
//          bench's 20 m/s, where tau clamps to 0.51).  After init these print
//          the scalars initLBM derived (tau, viscosity, Re_phys, Re_LB and
//          the air state) to 9 significant digits.
//   oracle/_ref/ref_fields atmosphere -   the air temperature and density of
//          the reference's standard atmosphere at 0 and 1500 m.
//
//   oracle/_ref/ref_fields box_les <out-dir> <nx> <ny> <nz> <tau> <steps> <dump-step>
//   a lattice-aligned box (padding 0) with the solid mask read from
//   <out-dir>/mask.u8 (nx*ny*nz bytes, x fastest), TRT + Smagorinsky Cs 0.1,
//   U0 0.08 along +x, convective outlet; prints the control-volume drag
//   coefficient (box 0.45/0.40/0.40) every 50 steps and dumps the fields after
//   <dump-step> steps.
//
//   oracle/_ref/ref_box box_time <out-dir> <nx> <ny> <nz> <tau> <steps>
//   the same case timed: one warm-up step, then one call of <steps> steps;
//   prints the wall time and the OpenMP thread count.  ref_box is this file
//   built with -O3 (the reference bench's flags) for timing.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "atmosphere.h"
#include "globals.h"
#include "lbm.h"
#include "thermal_lbm.h"

namespace {

// The reference bench's sphere (aeros/tests/bench_lbm.cpp:43-79): radius 0.35
// in [-1, 1]^3 on N^3 voxels, 25 % padding, tau 0.6, U0 0.08, convective outlet.
void sphere_case(int n, bool les, bool thermal) {
    g_voxNx = g_voxNy = g_voxNz = n;
    g_voxMinX = g_voxMinY = g_voxMinZ = -1.0f;
    g_voxMaxX = g_voxMaxY = g_voxMaxZ = 1.0f;
    g_voxelData.assign((size_t)n * n * n, 0);
    g_distanceField.assign((size_t)n * n * n, 0.0f);
    const float radius = 0.35f, h = 2.0f / (float)n;
    for (int z = 0; z < n; ++z)
        for (int y = 0; y < n; ++y)
            for (int x = 0; x < n; ++x) {
                const float px = -1.0f + (x + 0.5f) * h, py = -1.0f + (y + 0.5f) * h, pz = -1.0f + (z + 0.5f) * h;
                const float d = std::sqrt(px * px + py * py + pz * pz) - radius;
                g_voxelData[((size_t)z * n + y) * n + x] = d <= 0.0f ? 1 : 0;
                g_distanceField[((size_t)z * n + y) * n + x] = d / h;
            }
    minBB = glm::vec3(-radius);
    maxBB = glm::vec3(radius);
    center = glm::vec3(0.0f);
    maxDim = 2.0f * radius;
    voxelResolution = n;
    flowSpeed = 20.0f; flowAzimuth = 0.0f; flowElevation = 0.0f;
    altitude = 0.0f; timeScale = 1.0f; aeroMachEffects = false;
    lbmParams.enabled = true; lbmParams.useGPU = false;
    lbmParams.U0 = 0.08f; lbmParams.stepsPerFrame = 1; lbmParams.tau = 0.6f;
    lbmParams.useTurbulence = les; lbmParams.useGround = false; lbmParams.inletTurbulence = 0.0f;
    lbmParams.domainPadding = 0.25f; lbmParams.autoViscosity = false;
    lbmParams.useConvectiveOutlet = true; lbmParams.periodicX = false;
    lbmParams.bodyForceX = lbmParams.bodyForceY = lbmParams.bodyForceZ = 0.0f;
    thermalParams.enabled = thermal;
    thermalParams.rayleigh = 2500.0f;
    thermalParams.prandtl = 1.0f;
}

// The reference's Rayleigh-Benard box (aeros/tests/thermal_headless.cpp:54-99,
// 295-304): 32x18x8, solid rows y = 0 and y = 17, periodic x and z, tau 0.65,
// fixed wall temperatures 1 / 0, Ra 2500, Pr 1, perturbation 0.02.
void rb_case() {
    const int nx = 32, ny = 18, nz = 8;
    g_voxNx = nx; g_voxNy = ny; g_voxNz = nz;
    g_voxMinX = g_voxMinY = g_voxMinZ = -1.0f;
    g_voxMaxX = g_voxMaxY = g_voxMaxZ = 1.0f;
    g_voxelData.assign((size_t)nx * ny * nz, 0);
    g_distanceField.assign((size_t)nx * ny * nz, 0.0f);
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x) {
            g_voxelData[((size_t)z * ny + 0) * nx + x] = 1;
            g_voxelData[((size_t)z * ny + ny - 1) * nx + x] = 1;
        }
    minBB = glm::vec3(-1.0f);
    maxBB = glm::vec3(1.0f);
    center = glm::vec3(0.0f);
    maxDim = 2.0f;
    voxelResolution = nx;
    flowSpeed = 1.0f; flowAzimuth = 0.0f; flowElevation = 0.0f;
    altitude = 0.0f; timeScale = 1.0f; aeroMachEffects = false;
    lbmParams.enabled = true; lbmParams.useGPU = false;
    lbmParams.U0 = 0.0f; lbmParams.stepsPerFrame = 1; lbmParams.tau = 0.65f;
    lbmParams.useTurbulence = false; lbmParams.useGround = false; lbmParams.inletTurbulence = 0.0f;
    lbmParams.domainPadding = 0.0f; lbmParams.autoViscosity = false;
    lbmParams.useConvectiveOutlet = false; lbmParams.periodicX = true; lbmParams.periodicZ = true;
    lbmParams.bodyForceX = lbmParams.bodyForceY = lbmParams.bodyForceZ = 0.0f;
    thermalParams.enabled = true;
    thermalParams.rayleigh = 2500.0f;
    thermalParams.prandtl = 1.0f;
    thermalParams.fixedWallTemps = true;
    thermalParams.hotTemp = 1.0f;
    thermalParams.coldTemp = 0.0f;
    thermalParams.perturbation = 0.02f;
}

// A lattice-aligned box over the caller's mask, with the bench ellipsoid's
// flow settings (TRT + Smagorinsky, U0 0.08, convective outlet).
bool box_case(const std::string& dir, int nx, int ny, int nz, float tau) {
    g_voxNx = nx; g_voxNy = ny; g_voxNz = nz;
    g_voxMinX = g_voxMinY = g_voxMinZ = 0.0f;
    g_voxMaxX = (float)nx; g_voxMaxY = (float)ny; g_voxMaxZ = (float)nz;
    const size_t n = (size_t)nx * ny * nz;
    g_voxelData.assign(n, 0);
    g_distanceField.assign(n, 0.0f);
    std::vector<unsigned char> m(n);
    FILE* f = std::fopen((dir + "/mask.u8").c_str(), "rb");
    if (!f) return false;
    const bool ok = std::fread(m.data(), 1, n, f) == n;
    std::fclose(f);
    if (!ok) return false;
    for (size_t i = 0; i < n; ++i) g_voxelData[i] = m[i] ? 1 : 0;
    minBB = glm::vec3(0.0f);
    maxBB = glm::vec3((float)nx, (float)ny, (float)nz);
    center = 0.5f * maxBB;
    maxDim = (float)nx;
    voxelResolution = nx;
    flowSpeed = 20.0f; flowAzimuth = 0.0f; flowElevation = 0.0f;
    altitude = 0.0f; timeScale = 1.0f; aeroMachEffects = false;
    lbmParams.enabled = true; lbmParams.useGPU = false;
    lbmParams.U0 = 0.08f; lbmParams.stepsPerFrame = 1; lbmParams.tau = tau;
    lbmParams.useTurbulence = true; lbmParams.smagorinskyC = 0.1f; lbmParams.useTRT = true;
    lbmParams.useGround = false; lbmParams.inletTurbulence = 0.0f;
    lbmParams.domainPadding = 0.0f; lbmParams.autoViscosity = false;
    lbmParams.useConvectiveOutlet = true; lbmParams.periodicX = false; lbmParams.periodicZ = false;
    lbmParams.bodyForceX = lbmParams.bodyForceY = lbmParams.bodyForceZ = 0.0f;
    thermalParams.enabled = false;
    return true;
}

bool dump(const std::string& dir, const char* name, const std::vector<float>& v) {
    const std::string path = dir + "/" + name + ".f32";
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(v.data(), sizeof(float), v.size(), f) == v.size();
    return std::fclose(f) == 0 && ok;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <case> <out-dir>\n", argv[0]);
        return 2;
    }
    const std::string which = argv[1], dir = argv[2];
    if (which == "atmosphere") {   // the air state updateAtmosphereParams derives at a few altitudes
        for (float alt : {0.0f, 1500.0f}) {
            altitude = alt;
            updateAtmosphereParams();
            std::printf("altitude %.9g air_temperature %.9g air_density %.9g\n", altitude, airTemperature, airDensity);
        }
        return 0;
    }
    int steps = 0, dump_step = -1;
    const bool timing = which == "box_time";
    if (timing) {
        if (argc < 8) { std::fprintf(stderr, "box_time needs nx ny nz tau steps\n"); return 2; }
        if (!box_case(dir, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), (float)std::atof(argv[6]))) {
            std::fprintf(stderr, "cannot read %s/mask.u8\n", dir.c_str());
            return 1;
        }
        steps = std::atoi(argv[7]);
    }
    else if (which == "box_les") {
        if (argc < 9) { std::fprintf(stderr, "box_les needs nx ny nz tau steps dump-step\n"); return 2; }
        if (!box_case(dir, std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), (float)std::atof(argv[6]))) {
            std::fprintf(stderr, "cannot read %s/mask.u8\n", dir.c_str());
            return 1;
        }
        steps = std::atoi(argv[7]);
        dump_step = std::atoi(argv[8]);
    }
    else if (which == "bench12_les") { sphere_case(12, true, false); steps = 90; }
    else if (which == "bench12_les_thermal") { sphere_case(12, true, true); steps = 90; }
    else if (which == "bench16") { sphere_case(16, false, false); steps = 150; }
    else if (which == "rb_walls") { rb_case(); steps = 200; }
    else if (which == "rb_storm") {   // buoyancy so large that the populations overflow: the NaN-storm reset
        rb_case();
        thermalParams.rayleigh = 1.0e30f;
        steps = 12;
    }
    else if (which == "bench12_auto" || which == "bench12_auto_clamped") {
        sphere_case(12, true, false);
        lbmParams.autoViscosity = true;
        altitude = 1500.0f;
        updateAtmosphereParams();
        if (which == "bench12_auto") flowSpeed = 0.0004f;
        steps = 60;
    }
    else { std::fprintf(stderr, "unknown case %s\n", which.c_str()); return 2; }
    shutdownLBM();
    initLBM();
    if (!lbmInitialized) { std::fprintf(stderr, "init failed\n"); return 1; }
    if (lbmParams.autoViscosity)
        std::printf("auto tau %.9g viscosity %.9g re_phys %.9g re_lb %.9g air_temperature %.9g air_density %.9g "
                    "flow_speed %.9g\n", lbmParams.tau, lbmParams.viscosity, lbmReynolds, lbmReynoldsLB,
                    airTemperature, airDensity, flowSpeed);
    auto dump_all = [&]() {
        bool ok = dump(dir, "rho", lbmRho) && dump(dir, "ux", lbmUx) && dump(dir, "uy", lbmUy) && dump(dir, "uz", lbmUz) &&
                  dump(dir, "tke", lbmTKEField) && dump(dir, "strain", lbmStrainMag) && dump(dir, "ux_world", lbmUxWorld) &&
                  dump(dir, "pressure", lbmPressure);
        if (ok && thermalParams.enabled) ok = dump(dir, "T", thermalT) && dump(dir, "buoyancy", thermalBuoyancyY);
        return ok;
    };
    if (timing) {   // box_time: one warm-up step, then one call of `steps` steps
        stepLBMCPU(1);
        const auto t0 = std::chrono::steady_clock::now();
        stepLBMCPU(steps);
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        int threads = 1;
#ifdef _OPENMP
        threads = omp_get_max_threads();
#endif
        std::printf("box_time %dx%dx%d steps %d seconds %.6f threads %d step %d\n", lbmNx, lbmNy, lbmNz, steps, sec,
                    threads, lbmCurrentStep);
        shutdownLBM();
        return 0;
    }
    if (dump_step >= 0) {   // box_les: drag along the way, fields at dump_step
        const float area = lbmProjectedAreaCells(0);
        bool ok = true;
        for (int i = 1; i <= steps; ++i) {
            stepLBMCPU(1);
            if (i % 50 == 0) {
                const glm::vec3 F = lbmControlVolumeForce(0.45f, 0.40f, 0.40f);
                float umax = 0.0f;
                for (size_t c = 0; c < lbmUx.size(); ++c)
                    umax = std::max(umax, std::sqrt(lbmUx[c] * lbmUx[c] + lbmUy[c] * lbmUy[c] + lbmUz[c] * lbmUz[c]));
                std::printf("call %d step %d Cd %.6f A %.0f max|u| %.5f\n", i, lbmCurrentStep,
                            area > 0 ? F.x / (0.5f * 0.08f * 0.08f * area) : 0.0f, area, umax);
                std::fflush(stdout);
            }
            if (i == dump_step) ok = ok && dump_all();
        }
        shutdownLBM();
        return ok ? 0 : 1;
    }
    for (int i = 0; i < steps; ++i) stepLBMCPU(1);   // one call per step, as the frame loop does
    bool ok = dump(dir, "rho", lbmRho) && dump(dir, "ux", lbmUx) && dump(dir, "uy", lbmUy) && dump(dir, "uz", lbmUz) &&
              dump(dir, "tke", lbmTKEField) && dump(dir, "strain", lbmStrainMag) && dump(dir, "ux_world", lbmUxWorld) &&
              dump(dir, "pressure", lbmPressure);
    if (ok && thermalParams.enabled) ok = dump(dir, "T", thermalT) && dump(dir, "buoyancy", thermalBuoyancyY);
    std::printf("%s %dx%dx%d steps %d step %d\n", which.c_str(), lbmNx, lbmNy, lbmNz, steps, lbmCurrentStep);
    shutdownLBM();
    return ok ? 0 : 1;
}
