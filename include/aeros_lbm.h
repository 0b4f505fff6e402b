/*
 * aeros_lbm.h — C ABI of the MI355X (gfx950) lattice Boltzmann solver that
 * replaces the CPU hot path of Aeros Engine 1.38.1.
 *
 * Every entry point is plain C: opaque handles, plain pointers and sizes, an
 * int status (0 = success, negative = error; the message is kept per thread
 * and fetched with aeros_lbm_last_error).  Nothing throws across the ABI.
 * Caller buffers are borrowed only for the duration of a call.  Calls on one
 * handle are serialised on that handle's HIP stream; different handles are
 * independent.  Each entry point cites the reference interface it replaces
 * (paths relative to the reference repository root).
 *
 * Lattice conventions shared with the reference (they are an interface, not
 * an implementation choice):
 *   - cells are x-fastest, cell = (z*Ny + y)*Nx + x   (aeros/src/lbm.cpp:125)
 *   - D3Q19 direction order: 0 rest; 1..6 = +x,-x,+y,-y,+z,-z;
 *     7..10 = (+x+y),(-x-y),(+x-y),(-x+y); 11..14 = (+x+z),(-x-z),(+x-z),(-x+z);
 *     15..18 = (+y+z),(-y-z),(+y-z),(-y+z)   (aeros/src/lbm.cpp:92-109)
 *   - D3Q7 thermal directions are the first seven of the above
 *     (aeros/src/thermal_lbm.cpp:31-41).
 */
#ifndef AEROS_LBM_H
#define AEROS_LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AEROS_LBM_ABI_VERSION 1

/* status codes */
#define AEROS_OK               0
#define AEROS_ERR_ARG         -1   /* bad argument (null pointer, size mismatch, ...) */
#define AEROS_ERR_NOMEM       -2   /* device or host allocation failed              */
#define AEROS_ERR_HIP         -3   /* HIP runtime error                             */
#define AEROS_ERR_STATE       -4   /* handle not in a usable state                  */
#define AEROS_ERR_TOO_LARGE   -5   /* domain exceeds the configured cell cap        */
#define AEROS_ERR_UNSUPPORTED -6   /* option combination not built                  */
#define AEROS_ERR_NO_DEVICE   -7   /* no usable gfx950 device                       */

/* population storage precision */
#define AEROS_FP32 0   /* float populations, float collision (the reference's arithmetic) */
#define AEROS_FP64 1   /* double populations, double collision                           */
#define AEROS_FP16 2   /* IEEE half deviations f_i - w_i, float collision                */
#define AEROS_BF16 3   /* bfloat16 deviations f_i - w_i, float collision                 */
#define AEROS_FP32_FP64C 4   /* float deviations f_i - w_i and velocity stencil, double
                                collision and thermal lattice: FP32's memory traffic, FP64's
                                arithmetic (no counterpart in the reference)               */
#define AEROS_FP16_EQ 5   /* IEEE half of 2^12 (f_i - e_i), e_i the inflow equilibrium
                             f_i^eq(1, u_in): the free stream stores near zero, and the
                             scale keeps small deviations out of the subnormal range     */
#define AEROS_BF16_EQ 6   /* bfloat16 of f_i - e_i, e_i as for AEROS_FP16_EQ              */

/* field identifiers for aeros_lbm_read_field */
#define AEROS_FIELD_RHO        0   /* lbmRho        (aeros/src/lbm.h:132)          */
#define AEROS_FIELD_UX         1   /* lbmUx, lattice units, includes + F/(2 rho)   */
#define AEROS_FIELD_UY         2
#define AEROS_FIELD_UZ         3
#define AEROS_FIELD_UX_WORLD   4   /* lbmUxWorld, m/s (post-pass)                   */
#define AEROS_FIELD_UY_WORLD   5
#define AEROS_FIELD_UZ_WORLD   6
#define AEROS_FIELD_PRESSURE   7   /* lbmPressure, Pa (post-pass)                   */
#define AEROS_FIELD_TKE        8   /* lbmTKEField, m^2/s^2                          */
#define AEROS_FIELD_STRAIN     9   /* lbmStrainMag as the solver stores it (|S| LB) */
#define AEROS_FIELD_SOLID     10   /* lbmIsSolid as 0/1                              */
#define AEROS_FIELD_GROUND    11   /* lbmIsGround as 0/1                             */
#define AEROS_FIELD_WALLDIST  12   /* u8 wall distance (0..255) as float            */
#define AEROS_FIELD_T         13   /* thermalT      (aeros/src/thermal_lbm.h:46)     */
#define AEROS_FIELD_BUOYANCY  14   /* thermalBuoyancyY                              */
#define AEROS_FIELD_VORTICITY 15   /* lbmVorticityMag, 1/s (aeros_lbm_vorticity_q)  */
#define AEROS_FIELD_QCRIT     16   /* lbmQCriterion, 1/s^2 (aeros_lbm_vorticity_q)  */
#define AEROS_FIELD_COUNT     17

/* bit masks for aeros_lbm_read */
#define AEROS_READ_RHO (1u << 0)
#define AEROS_READ_UX  (1u << 1)
#define AEROS_READ_UY  (1u << 2)
#define AEROS_READ_UZ  (1u << 3)
#define AEROS_READ_T   (1u << 4)

typedef struct aeros_lbm aeros_lbm;

/*
 * Everything initLBM reads (aeros/src/lbm.cpp:405-886): the set-mask producer's
 * voxel grid and box (g_voxelData, g_vox*; aeros/src/globals.h:150-161), the
 * flow and atmosphere globals, LBMParams (aeros/src/lbm.h:11-75) and
 * ThermalParams (aeros/src/thermal_lbm.h:33-41), plus the MI355X-side storage
 * precision and cell cap.  aeros_lbm_config_defaults fills the reference's
 * defaults.
 */
typedef struct aeros_lbm_config {
    int32_t struct_size;          /* = sizeof(aeros_lbm_config) */
    /* voxel grid (1 = solid), resampled by cell-centre world coordinates */
    const int32_t* voxels;        /* may be NULL: empty domain of voxel_resolution^3 */
    int32_t vox_n[3];
    float   vox_min[3];
    float   vox_max[3];
    int32_t voxel_resolution;
    /* flow / atmosphere globals */
    float   flow_speed;           /* m/s                 */
    float   flow_azimuth_deg;
    float   flow_elevation_deg;
    float   air_density;          /* kg/m^3              */
    float   air_temperature;      /* K                   */
    float   max_dim;              /* model size, world   */
    float   frontal_area;         /* computeLBMRefArea() result (world units^2) for the
                                     blockage ratio; 0 = unknown, blockage reported as 0 */
    int32_t ground_effect;        /* aeroGroundEffect    */
    float   ground_effect_height; /* aeroGroundHeight    */
    /* LBMParams */
    float   tau;
    float   U0;
    float   smagorinsky_c;
    float   convergence_threshold;
    float   domain_padding;
    float   inlet_turbulence;
    float   ground_height;
    float   body_force[3];
    int32_t use_les;
    int32_t use_trt;
    int32_t convective_outlet;
    int32_t periodic_x;
    int32_t periodic_z;
    int32_t auto_viscosity;
    int32_t use_ground;
    /* ThermalParams */
    int32_t thermal;
    float   rayleigh;
    float   prandtl;
    int32_t fixed_wall_temps;
    float   hot_temp;
    float   cold_temp;
    float   perturbation;
    /* MI355X side */
    int32_t precision;            /* AEROS_FP32 / FP64 / FP16 / BF16 / FP32_FP64C          */
    int64_t cell_cap;             /* 0 = the reference's caps (8 Mi padded, 20 Mi hard);
                                     >0 = explicit cap on Nx*Ny*Nz                          */
    int32_t device;               /* HIP device ordinal                                    */
    int32_t free_slip;            /* faces closed by a free-slip (specular) wall, a mask of
                                     AEROS_FACE_*; none may lie on a periodic axis.  No
                                     counterpart in the reference, whose non-periodic faces
                                     copy the cell's own population in place.  Acts on the
                                     D3Q19 lattice; the D3Q7 thermal lattice keeps its own
                                     shell rules.                                          */
} aeros_lbm_config;

/* Domain faces (aeros_lbm_config.free_slip). */
#define AEROS_FACE_X0 1   /* x = 0      */
#define AEROS_FACE_X1 2   /* x = Nx - 1 */
#define AEROS_FACE_Y0 4   /* y = 0      */
#define AEROS_FACE_Y1 8   /* y = Ny - 1 */
#define AEROS_FACE_Z0 16  /* z = 0      */
#define AEROS_FACE_Z1 32  /* z = Nz - 1 */

/* Scalars the reference keeps in globals (aeros/src/lbm.h:77-125,
 * aeros/src/thermal_lbm.h:46-58).  Values are those after the last step. */
typedef struct aeros_lbm_scalars {
    int32_t step;                 /* lbmCurrentStep        */
    int32_t converged;            /* lbmConverged          */
    int32_t resets;               /* divergence resets so far */
    int32_t nx, ny, nz;
    float   avg_rho;              /* lbmAvgRho             */
    float   avg_ke;               /* lbmAvgKineticEnergy   */
    float   max_u_lb;             /* lbmMaxVelocityLB      */
    float   max_u_world;          /* lbmMaxVelocityWorld   */
    float   tke;                  /* lbmTKE                */
    float   convergence;          /* lbmConvergence        */
    int64_t neg_pops;             /* lbmNegativePops       */
    int64_t bad_pops;             /* lbmBadPops            */
    int64_t rho_out_of_range;     /* lbmRhoOutOfRange      */
    int64_t rho_non_finite;       /* lbmRhoNonFinite       */
    int64_t u_limited;            /* lbmULimited           */
    int64_t pressure_non_finite;  /* lbmPressureNonFinite  */
    float   tau;                  /* lbmParams.tau after init */
    float   viscosity;            /* lbmParams.viscosity   */
    float   U0;                   /* lbmParams.U0 after init */
    float   reynolds_phys;        /* lbmReynolds           */
    float   reynolds_lb;          /* lbmReynoldsLB         */
    float   blockage;             /* lbmBlockageRatio      */
    float   pressure_min;         /* lbmPressureMin        */
    float   pressure_max;         /* lbmPressureMax        */
    float   residual_rho;         /* lbmResidualRho        */
    float   residual_u;           /* lbmResidualU          */
    float   time_ms;              /* lbmTimeMs (host wall time of the last step call) */
    int64_t fluid_cells;
    /* thermal */
    float   nusselt;              /* thermalNusselt        */
    float   thermal_mean_t;       /* thermalMeanT          */
    float   thermal_alpha;        /* thermalAlpha          */
    float   thermal_gravity;      /* thermalGravityLb      */
    float   thermal_tau;          /* tau_T                 */
    float   thermal_rayleigh;     /* thermalRayleigh       */
    float   thermal_prandtl;      /* thermalPrandtl        */
    int32_t thermal_ready;        /* thermalReady          */
    int32_t thermal_ylo, thermal_yhi;
} aeros_lbm_scalars;

/* Backend description, the analogue of CudaBackendInfo (aeros/src/cuda_api.h:29-46). */
typedef struct aeros_backend_info {
    int32_t compiled_with_hip;
    int32_t device_count;
    int32_t usable;               /* 1 = a test kernel ran on the device */
    int32_t gfx_major, gfx_minor, gfx_stepping;
    int32_t runtime_version;
    int32_t driver_version;
    int64_t total_memory_mb;
    int32_t compute_units;
    int32_t last_error;
    char    device_name[128];
    char    arch_name[64];
    char    status[512];
} aeros_backend_info;

/* ---------------------------------------------------------------- setup */

/* Reference defaults of LBMParams/ThermalParams/globals (aeros/src/lbm.h:11-75,
 * aeros/src/thermal_lbm.h:33-41, aeros/src/globals.cpp:59-145). */
void aeros_lbm_config_defaults(aeros_lbm_config* cfg);

/* initLBM (aeros/src/lbm.h:156, aeros/src/lbm.cpp:405-886): builds the domain
 * from the voxel box plus padding, resamples the mask, computes the wall
 * distance, initialises the lattice to equilibrium, picks tau and (if asked)
 * initialises the thermal lattice.  Returns NULL on failure with *err set. */
aeros_lbm* aeros_lbm_create(const aeros_lbm_config* cfg, int* err);

/* shutdownLBM (aeros/src/lbm.h:158). */
int aeros_lbm_destroy(aeros_lbm* h);

/* resetLBM (aeros/src/lbm.h:157): shutdown + init with the handle's config. */
int aeros_lbm_reset(aeros_lbm* h);

/* Domain geometry (lbmNx/Ny/Nz, lbmMin*, lbmCellSize*; aeros/src/lbm.h:127-130). */
int aeros_lbm_dims(const aeros_lbm* h, int32_t dims[3], float box_min[3], float cell_size[3]);

/* Set-mask at lattice resolution (the voxel -> lattice transfer of
 * aeros/src/lbm.cpp:596-622 already done by the caller): 1 = solid.  The
 * wall distance is recomputed and the lattice re-initialised to equilibrium,
 * as initLBM does when the voxel grid changes (aeros/src/lbm.cpp:1736-1760).
 * The host buffer is copied during the call. */
int aeros_lbm_set_mask(aeros_lbm* h, const uint8_t* solid, int32_t nx, int32_t ny, int32_t nz);

/* ---------------------------------------------------------------- step */

/* stepLBMCPU(steps) (aeros/src/lbm.h:160, aeros/src/lbm.cpp:907-1709),
 * including the thermal step (aeros/src/thermal_lbm.cpp:199-355), the
 * divergence reset and the end-of-call post-pass. */
int aeros_lbm_step(aeros_lbm* h, int32_t steps);

/* updateLBM(dt, forceSteps) (aeros/src/lbm.h:164, aeros/src/lbm.cpp:1734-1779):
 * forceSteps >= 0 is clamped to 64; -1 uses stepsPerFrame with the frame-time
 * rule.  steps_per_frame is LBMParams::stepsPerFrame. */
int aeros_lbm_update(aeros_lbm* h, float dt, int32_t force_steps, int32_t steps_per_frame);
/* The lattice steps updateLBM runs for these arguments (host rule only). */
int32_t aeros_lbm_update_steps(float dt, int32_t force_steps, int32_t steps_per_frame);
/* lbmLatticeDtWorld (aeros/src/lbm.h:171, aeros/src/lbm.cpp:1717-1732):
 * physical seconds one lattice step represents, cell size / (flowSpeed / |u_in|). */
int aeros_lbm_lattice_dt(aeros_lbm* h, float* out);

/* ------------------------------------------------- host stepping driver */

/* Fixed-tick physics clock (PhysClock, aeros/src/phys_clock.h:48-250):
 * settings first (as the reference's defaults: fixed, 240 Hz, 64 ticks per
 * frame, 8 particle sub-steps, frames cut at 0.25 s), then state and the
 * last frame's measurements. */
typedef struct aeros_phys_clock {
    int32_t fixed;                /* 0 = one tick per frame, the solver's own steps per frame */
    float   hz;
    int32_t max_substeps;
    int32_t particle_substeps;
    float   max_frame_dt;
    float   accumulator;          /* real seconds not yet ticked */
    float   lattice_accum;        /* fraction of a lattice step carried over */
    double  sim_time, real_time, dropped_time;
    int64_t total_ticks, total_lattice_steps;
    int32_t ticks, lattice_steps, part_steps;
    float   tick_dt, achieved_hz, time_scale, wall_scale, display_ratio;
    int32_t saturated;            /* the simulation fell behind real time this frame */
} aeros_phys_clock;

void aeros_clock_defaults(aeros_phys_clock* c);
void aeros_clock_reset(aeros_phys_clock* c);        /* state only: the settings stay */
void aeros_clock_sanitize(aeros_phys_clock* c);     /* clamps settings and accumulators */
/* One rendered frame of frame_dt seconds: returns the physics ticks and puts
 * the lattice steps (lattice_dt seconds each; <= 0 or NaN: none) in
 * *lbm_steps (-1 when fixed = 0: the solver chooses). */
int32_t aeros_clock_advance(aeros_phys_clock* c, float frame_dt, float lattice_dt, float fps, int32_t* lbm_steps);
float aeros_clock_particle_dt(const aeros_phys_clock* c);
float aeros_clock_frame_sim_dt(const aeros_phys_clock* c);

/* Asynchronous frames (no counterpart in the reference, whose updateLBM
 * blocks the frame loop).  After aeros_lbm_frames_start a worker thread runs
 * each submitted frame as aeros_lbm_update does (the same step rule, then
 * computeLBMVorticityAndQ whenever the step count is a multiple of 30), copies
 * the fields of field_mask (bits 1 << AEROS_FIELD_*; device fields only, not
 * SOLID/GROUND/WALLDIST) into a device staging buffer and from there, on a
 * copy stream overlapped with the next frame's steps, into pinned host
 * memory.  Three host slots rotate: the caller holds one, one holds the
 * newest finished frame, one is being written, so neither side waits for the
 * other.  Every other call on the handle waits for the frame in progress. */
typedef struct aeros_frame {
    int64_t frame;                /* 1, 2, ... in submission order; 0 = none finished yet */
    int32_t steps;                /* lattice steps that frame ran */
    aeros_lbm_scalars scalars;    /* after that frame */
    const float* field[AEROS_FIELD_COUNT];   /* Nx*Ny*Nz floats each, NULL if not in the mask */
} aeros_frame;
int aeros_lbm_frames_start(aeros_lbm* h, uint32_t field_mask);
/* Queues one frame (at most two wait; a third submission blocks until the
 * worker takes one).  Returns the first error a finished frame met, if any. */
int aeros_lbm_frame_submit(aeros_lbm* h, float dt, int32_t force_steps, int32_t steps_per_frame);
/* Never blocks: the newest finished frame, whose buffers stay valid until the
 * next call of aeros_lbm_frame_latest or aeros_lbm_frames_stop. */
int aeros_lbm_frame_latest(aeros_lbm* h, aeros_frame* out);
/* Blocks until every submitted frame has finished and been copied. */
int aeros_lbm_frames_wait(aeros_lbm* h);
int aeros_lbm_frames_stop(aeros_lbm* h);           /* waits, then frees the buffers */

/* ---------------------------------------------------------------- readback */

/* Synchronous device->host copy of one field into a caller buffer of
 * Nx*Ny*Nz floats (globals of aeros/src/lbm.h:132-154). */
int aeros_lbm_read_field(aeros_lbm* h, int32_t field, float* out);
/* Same, in double (exact for AEROS_FP64 handles, widened otherwise). */
int aeros_lbm_read_field_f64(aeros_lbm* h, int32_t field, double* out);
/* The visualizer's bundle: any subset of rho, u, T (null pointers skipped). */
int aeros_lbm_read(aeros_lbm* h, uint32_t fields, float* rho, float* ux, float* uy, float* uz, float* T);
/* Scalars (aeros/src/lbm.h:77-125, aeros/src/thermal_lbm.h:46-58). */
int aeros_lbm_get_scalars(aeros_lbm* h, aeros_lbm_scalars* out);

/* Raw population readback, reference layout f[cell*19+q] of the state the
 * reference holds after the same number of steps (post-stream), as double. */
int aeros_lbm_read_populations(aeros_lbm* h, double* out);

/* ---------------------------------------------------------------- forces */

/* lbmControlVolumeForce (aeros/src/lbm.h:200, aeros/src/lbm.cpp:2183-2232):
 * momentum balance over a box; out = force on the body before the float cast. */
int aeros_lbm_cv_force(aeros_lbm* h, float half_x, float half_y, float half_z, double out[3]);
/* lbmProjectedAreaCells (aeros/src/lbm.h:202, aeros/src/lbm.cpp:2236-2256). */
int aeros_lbm_projected_area(aeros_lbm* h, int32_t axis, float* out);
/* Deterministic FP64 sums over all cells of rho and of ux+uy+uz, the bench's
 * checksum (aeros/tests/bench_lbm.cpp:103-107). */
int aeros_lbm_checksums(aeros_lbm* h, double* rho_sum, double* u_sum);
/* The deterministic FP64 mass (rho sum over all cells) alone, the quantity
 * the FP64 acceptance criterion compares over 120 steps. */
int aeros_lbm_rho_sum(aeros_lbm* h, double* rho_sum);

/* ---------------------------------------------------------------- thermal */

/* thermalUpload (aeros/src/thermal_lbm.h:67): populations to g_eq(T, 0). */
int aeros_lbm_thermal_upload(aeros_lbm* h, const float* T);
/* thermalResetField (aeros/src/thermal_lbm.h:64). */
int aeros_lbm_thermal_reset(aeros_lbm* h);

/* ---------------------------------------------------------------- hooks */

/* lbmEquilibriumMoments (aeros/src/lbm.h:106-107, aeros/src/lbm.cpp:249-266),
 * evaluated by the device's equilibrium code.  Returns 1 on success, 0 on a
 * non-finite input or null output pointer (the reference's bool). */
int aeros_lbm_equilibrium_moments(float rho, float ux, float uy, float uz,
                                  float* sum_feq, float* mom_x, float* mom_y, float* mom_z);

/* ---------------------------------------------------------------- voxels */

/* Solid mask by the 3-ray majority vote at cell centres
 * (meshbvh::BVH::inside, aeros/src/mesh_bvh.h:168-185; loop of
 * aeros/src/voxel_grid.cpp:337-351) on an arbitrary grid: centre of voxel
 * (i,j,k) = box_min + (idx + 0.5) * cell.  tris holds 9 floats per triangle;
 * non-finite and degenerate triangles are dropped as the reference does.
 * out receives nx*ny*nz int32 (1 = solid). */
int aeros_voxelize(const float* tris, int64_t ntris, int32_t nx, int32_t ny, int32_t nz,
                   const float box_min[3], const float cell[3], int32_t* out);

/* The same vote on the device (replaces the per-voxel ray loop of
 * buildVoxelGrid, aeros/src/voxel_grid.cpp:337-351; the BVH of
 * aeros/src/mesh_bvh.h:101-160 is built on the host and uploaded): one
 * thread per voxel, identical float slab and double-precision triangle
 * tests, so out is bit-identical to aeros_voxelize.  Synchronous; borrows
 * tris and out for the call only.  AEROS_ERR_NO_DEVICE without a gfx950
 * device (no host fallback). */
int aeros_voxelize_gpu(const float* tris, int64_t ntris, int32_t nx, int32_t ny, int32_t nz,
                       const float box_min[3], const float cell[3], int32_t device, int32_t* out);

/* ---------------------------------------------------------------- post-processing on the device */

/* computeLBMVorticityAndQ (aeros/src/lbm.cpp:2005-2093): from the world
 * velocity of the last step, |omega| and Q on the interior cells (0 in solid
 * cells), and sqrt(2 S:S) in world units written over AEROS_FIELD_STRAIN, as
 * the reference overwrites lbmStrainMag. */
int aeros_lbm_vorticity_q(aeros_lbm* h);

/* The world-space samplers the visualizer calls per point, batched:
 * sampleFieldTrilinear over fluid cells (aeros/src/lbm.cpp:1871-1925) for
 * field = AEROS_FIELD_PRESSURE (getLBMPressureWorld), RHO (getLBMDensityWorld:
 * clamped to [0.8, 1.2], 1 when no fluid corner), TKE, STRAIN, VORTICITY,
 * QCRIT (0 when no fluid corner).  pos holds npos (x, y, z) world points;
 * out npos values; valid (may be NULL) 1 where the sampler found fluid data. */
int aeros_lbm_sample(aeros_lbm* h, int32_t field, const float* pos, int64_t npos, float* out, uint8_t* valid);
/* aeros_lbm_sample without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, per-frame probe and streamline
 * loops that must not call hipMalloc).  No counterpart in the reference.  Same
 * fields, arguments and results as aeros_lbm_sample; the positions, values and
 * flags are staged on the device in a fixed library array (1.06 MiB per
 * device, held until the process ends) instead of per-call buffers.  At most
 * 65536 points per call (AEROS_ERR_TOO_LARGE beyond).  VORTICITY and QCRIT
 * need their device buffers to exist: AEROS_ERR_STATE until the first
 * aeros_lbm_vorticity_q (or aeros_lbm_update that runs it) on the handle.
 * Calls on one device share the array; the library runs them one at a time,
 * and the array keeps the last call's contents. */
int aeros_lbm_sample_noalloc(aeros_lbm* h, int32_t field, const float* pos, int64_t npos, float* out,
                             uint8_t* valid);
/* getLBMVelocityLB (world = 0, lattice units) or getLBMVelocityWorld
 * (world = 1: free stream outside the lattice box, 0 in solid cells, else
 * scaled by flowSpeed/U0 and capped at 2 flowSpeed), aeros/src/lbm.cpp:
 * 1781-1869.  out receives 3 floats per point. */
int aeros_lbm_sample_velocity(aeros_lbm* h, int32_t world, const float* pos, int64_t npos, float* out);

/* ---------------------------------------------------------------- STL */

/* Format of an STL buffer, restating stlformat::detect
 * (aeros/src/stl_format.h:86-144): ASCII when it starts with "solid" and its
 * text holds facet/vertex; binary when the count at byte 80 promises
 * triangles the buffer holds (a tail is allowed and counted). */
#define AEROS_STL_UNKNOWN 0
#define AEROS_STL_BINARY  1
#define AEROS_STL_ASCII   2
/* reasons */
#define AEROS_STL_EMPTY            1
#define AEROS_STL_SOLID_KEYWORDS   2   /* "solid" and facet/vertex: ASCII           */
#define AEROS_STL_BINARY_EXACT     3
#define AEROS_STL_BINARY_TAIL      4   /* binary with bytes after the triangles     */
#define AEROS_STL_TEXT_NO_SOLID    5   /* text with keywords but no "solid": ASCII  */
#define AEROS_STL_TEXT_NO_KEYWORDS 6   /* text without facet/vertex: unknown        */
#define AEROS_STL_TRUNCATED        7   /* header promises more triangles than held  */
#define AEROS_STL_UNRECOGNISED     8

typedef struct aeros_stl_info {
    int32_t  format;
    int32_t  reason;
    uint32_t declared_triangles;   /* count at byte 80 (binary)            */
    int32_t  pad;
    int64_t  payload_bytes;        /* 50 per declared triangle (binary)    */
    int64_t  trailing_bytes;       /* bytes after the last triangle        */
} aeros_stl_info;

int aeros_stl_detect(const uint8_t* data, int64_t size, aeros_stl_info* out);

/* loadSTL (aeros/src/stl_loader.cpp:30-205) on a buffer: *tris and *normals
 * receive library-allocated arrays of 9 floats per triangle (vertex normals,
 * the facet normal repeated; free both with aeros_stl_free).  Binary: facet
 * normals sanitised and normalised, triangles with a non-finite coordinate
 * dropped.  ASCII: vertices read in threes, non-finite ones dropped, normals
 * recomputed from the edges.  *dropped counts what was dropped.  Returns
 * AEROS_ERR_ARG when the format is unknown or no triangle remains. */
int aeros_stl_parse(const uint8_t* data, int64_t size, float** tris, float** normals, int64_t* ntris,
                    int64_t* dropped);
/* The same from a file (10 B .. 2 GiB, as loadSTL accepts). */
int aeros_stl_load(const char* path, float** tris, float** normals, int64_t* ntris, int64_t* dropped);
void aeros_stl_free(float* p);
/* Binary STL writer (no counterpart in the reference, which only reads):
 * 80-byte header (header text, zero padded), count, 50 bytes per triangle
 * with the unit normal of its edges ((0,0,1) when degenerate). */
int aeros_stl_write_binary(const char* path, const float* tris, int64_t ntris, const char* header);

/* buildVoxelGrid's grid plan (aeros/src/voxel_grid.cpp:225-300): bbox plus
 * 0.14*maxDim margin, nx = res, ny/nz proportional.  Fills dims and box. */
int aeros_voxel_plan(const float bb_min[3], const float bb_max[3], float max_dim, int32_t res,
                     int32_t dims[3], float box_min[3], float box_max[3]);

/* ---------------------------------------------------------------- diagnostics */

/* Kernel timing of the step kernel (no counterpart in the reference, whose
 * only timer is the wall time of a stepLBMCPU call, aeros/src/lbm.cpp:916,1707):
 * enable = 1 brackets every step-kernel launch with HIP events on the
 * handle's stream and accumulates their elapsed time; enable = 2 brackets
 * the back-to-back launches of each step call with one event pair (the
 * per-launch pair itself delays the next dispatch by several microseconds);
 * enable = 0 stops and leaves the totals readable; enable = -1 resets them.
 * total_ms and launches may be NULL. */
int aeros_lbm_kernel_timing(aeros_lbm* h, int32_t enable, double* total_ms, int64_t* launches);

/* cudaConsumeError analogue (aeros/src/cuda_api.h:53-56): copies the last
 * error message of this thread, clears it and returns its code (0 = none). */
int aeros_lbm_last_error(char* buf, int32_t size);
/* cudaProbeBackend analogue (aeros/src/cuda_api.h:29-46): enumerates devices
 * and runs a test kernel. */
int aeros_probe(aeros_backend_info* out);
/* cudaReleaseAll analogue (aeros/src/cuda_api.h:51): frees library-wide
 * device buffers (handles stay owned by their callers). */
void aeros_release_all(void);

/* The existing C-ABI hook the voxelizer calls (aeros/src/cuda_api.h:15-18,
 * aeros/src/voxel_grid.cpp:552).  Kept with its name and signature; the
 * signed distance field it carries is not used by the LBM and is ignored. */
void setVoxelData(const int* voxel, const float* dist, int nx, int ny, int nz,
                  float mnX, float mnY, float mnZ, float csX, float csY, float csZ);

#ifdef __cplusplus
}
#endif

#endif /* AEROS_LBM_H */
